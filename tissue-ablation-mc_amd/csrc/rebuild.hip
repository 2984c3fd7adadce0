// rebuild.hip -- ablation mask, property/optics rebuild, stencil face
// coefficients, MC tally normalisation, diagnostics.
//
// Ablation mask.  setupThermalCoeff (3dFD.f90:330-360) sweeps the grid in
// place, x fastest, and zeroes a voxel when the sum of its six neighbours'
// rhokap is zero, reading the already-updated -x/-y/-z neighbours and the
// not-yet-updated +x/+y/+z ones.  With B(v) = [T(v) >= T_ablate] or
// [rhokap(v) = 0 before the sweep] and U(v) = the +x,+y,+z neighbours were
// empty before the sweep, the sweep's result is
//     Z(v) = B(v) or (U(v) and Z(v-x) and Z(v-y) and Z(v-z)).
// If B(v) is false then rhokap(v) > 0 before the sweep, so v-x (whose +x
// neighbour is v) has U(v-x) false and Z(v-x) = B(v-x); likewise v-y, v-z.
// Hence Z(v) = B(v) or (U(v) and B(v-x) and B(v-y) and B(v-z)): one
// data-parallel pass, bit-exact with the sequential sweep (halo voxels are
// empty, so they count as B = 1 and U-true).
#include "tam_internal.h"

// The property fits and face coefficients are FP64 restatements of the
// Fortran expressions; no fused multiply-add, so they round like the
// reference's separate multiply and add.
#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double air_thermal_cond(double T)      // thermalConst_mod.f90:11-25
{
    return -0.188521 * exp(-0.000367259 * (T - 273.15)) + 0.212453;
}
__device__ __forceinline__ double air_density(double T)           // :28-38
{
    return 101.325e3 / (287.058 * T);
}
constexpr double kW0 = 0.75, kProt = 1.0 - 0.75;                   // :5-6
__device__ __forceinline__ double skin_density(double W) { return 1000.0 / (W + 0.649 * kProt); }
__device__ __forceinline__ double skin_heatcap(double W) { return 1000.0 * (4.2 * W + 1.09 * kProt); }
__device__ __forceinline__ double skin_cond(double W, double rho) { return rho * (6.28e-4 * W + 1.17e-4 * kProt); }

__device__ __forceinline__ void vox_ijk(long v, int nx, int ny, int& i, int& j, int& k)
{
    i = (int)(v % nx) + 1;
    const long r = v / nx;
    j = (int)(r % ny) + 1;
    k = (int)(r / ny) + 1;
}

__global__ __launch_bounds__(256) void water_kernel(RebuildArgs a)   // 3dFD.f90:328
{
    const long nv = (long)a.nx * a.ny * a.nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    double w = kW0 - kW0 * (a.Q[v] / a.qvapor);
    w = fmin(fmin(w, kW0), a.W[v]);
    a.W[v] = fmax(w, 0.0);
}

__global__ __launch_bounds__(256) void mask_kernel(RebuildArgs a)
{
    const long nv = (long)a.nx * a.ny * a.nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    int i, j, k;
    vox_ijk(v, a.nx, a.ny, i, j, k);
    const long c = i + a.px * j + a.pxy * k;
    const double* rk = a.rhokap;
    auto hot = [&](long cc) { return (a.T64 ? a.T64[cc] : a.t_ref + (double)a.dT[cc]) >= a.ablate_K; };
    const bool b = hot(c) || rk[c] == 0.0;
    const bool u = rk[c + 1] == 0.0 && rk[c + a.px] == 0.0 && rk[c + a.pxy] == 0.0;
    const bool bx = (i == 1) || hot(c - 1) || rk[c - 1] == 0.0;
    const bool by = (j == 1) || hot(c - a.px) || rk[c - a.px] == 0.0;
    const bool bz = (k == 1) || hot(c - a.pxy) || rk[c - a.pxy] == 0.0;
    a.mask[v] = (uint8_t)(b || (u && bx && by && bz));
}

__global__ __launch_bounds__(256) void props_kernel(RebuildArgs a)   // 3dFD.f90:334-357
{
    const long nv = (long)a.nx * a.ny * a.nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    int i, j, k;
    vox_ijk(v, a.nx, a.ny, i, j, k);
    const long c = i + a.px * j + a.pxy * k;
    if (a.mask[v]) {
        const double T = a.T64 ? a.T64[c] : a.t_ref + (double)a.dT[c];
        const double rho = air_density(T);
        const double hc = 1.006e3;
        a.density[c] = rho;
        a.heatcap[c] = hc;
        a.rhokap[c] = 0.0;
        a.kappa[c] = air_thermal_cond(T);
        a.coeff[c] = a.delt / (air_density(T) * hc);
        a.mut[v] = 0.f;
        a.mat[v] = (uint8_t)a.air_id;
        if (a.cls) a.cls[c] = (uint8_t)(TAM_CLASS_REMOVED | (a.cls[c] & TAM_CLASS_DUE));
    } else if (a.mode == TAM_THERMAL_WATER) {
        const double W = a.W[v];
        const double rho = skin_density(W);
        const double rk = W * a.mu_water + a.mu_protein;
        const double hc = skin_heatcap(W);
        a.density[c] = rho;
        a.rhokap[c] = rk;
        a.heatcap[c] = hc;
        a.kappa[c] = skin_cond(W, rho);
        a.coeff[c] = a.delt / (rho * hc);
        a.mut[v] = (float)rk;
    }
}

// Face coefficients G = delt * s * kbar/(rhobar*cbar) / h^2 (3dFD.f90:119-166),
// one thread per halo cell; the face is between c and its + neighbour.
__global__ __launch_bounds__(256) void coeff_kernel(CoeffArgs a)
{
    const long n = (long)(a.nx + 2) * (a.ny + 2) * (a.nz + 2);
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % a.px);
    const long r = c / a.px;
    const int j = (int)(r % (a.ny + 2));
    const int k = (int)(r / (a.ny + 2));
    const double* K = a.kappa;
    const double* D = a.density;
    const double* H = a.heatcap;
    auto face = [&](long o, double ih2) {
        const double kb = .5 * (K[c] + K[c + o]);
        const double db = .5 * (D[c] + D[c + o]);
        const double hb = .5 * (H[c] + H[c + o]);
        return (float)(a.delt * (a.dscale * (kb / (db * hb)) * ih2));
    };
    const bool jin = j >= 1 && j <= a.ny, kin = k >= 1 && k <= a.nz, iin = i >= 1 && i <= a.nx;
    // a removed voxel is an interior cell whose rhokap is 0 (halo cells are
    // empty too, but they carry the boundary condition)
    auto removed = [&](long cc, bool interior) { return a.removed_inert && interior && a.rhokap[cc] == 0.0; };
    const bool rc = removed(c, iin && jin && kin);
    a.gx[c] = (i <= a.nx && jin && kin && !rc && !removed(c + 1, i + 1 <= a.nx && jin && kin))
                  ? face(1, a.idx2) : 0.f;
    a.gy[c] = (j <= a.ny && iin && kin && !rc && !removed(c + a.px, j + 1 <= a.ny && iin && kin))
                  ? face(a.px, a.idy2) : 0.f;
    a.gz[c] = (k <= a.nz && iin && jin && !rc && !removed(c + a.pxy, k + 1 <= a.nz && iin && jin))
                  ? face(a.pxy, a.idz2) : 0.f;
    if (a.robin && k == a.nz && iin && jin) {          // convective top (extension)
        const double kk = K[c] / a.dz;
        const double den = kk + 0.5 * a.h_top;
        const double beta = (kk - 0.5 * a.h_top) / den;
        const double gam = a.h_top * a.t_amb / den;
        a.robin_coef[(i - 1) + (long)a.nx * (j - 1)] =
            make_float2((float)beta, (float)(beta * a.t_ref + gam - a.t_ref));
    }
}

// jmeanGLOBAL = scale * mu_a * L (mcpolar.f90:174); src = coeff * jmean.
__global__ __launch_bounds__(256) void normalise_kernel(int nx, int ny, int nz, long px, long pxy,
                                                        const double* tally, const double* rhokap,
                                                        const uint8_t* mat, const double* mua_frac,
                                                        const CycleHdr* dyn, double scale,
                                                        const double* coeff, double* jmean, float* src)
{
    const long nv = (long)nx * ny * nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    if (dyn) {
        if (!dyn->mc_on) return;
        scale = dyn->scale;
    }
    int i, j, k;
    vox_ijk(v, nx, ny, i, j, k);
    const long c = i + px * j + pxy * k;
    const double mua = rhokap[c] * mua_frac[mat[v]];
    const double jm = tally[v] * mua * scale;
    jmean[v] = jm;
    src[c] = (float)(coeff[c] * jm);
}

__global__ __launch_bounds__(256) void energy_kernel(int nx, int ny, int nz, long px, long pxy,
                                                     const float* dT, const double* T64, double t_ref,
                                                     const double* heatcap, double mass, double* out)
{
    __shared__ double red[4];
    const long nv = (long)nx * ny * nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (v < nv) {
        int i, j, k;
        vox_ijk(v, nx, ny, i, j, k);
        const long c = i + px * j + pxy * k;
        e = heatcap[c] * mass * (T64 ? T64[c] - t_ref : (double)dT[c]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

__global__ void fill_f64_kernel(double* p, long n, double v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_f32_kernel(float* p, long n, float v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u64_kernel(unsigned long long* p, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0ull;
}
// acc += src over n counters (one block; n is a handful)
__global__ void acc_u64_kernel(unsigned long long* acc, const unsigned long long* src, long n)
{
    for (long i = threadIdx.x; i < n; i += blockDim.x) acc[i] += src[i];
}

// Streaming copy (16-byte loads and stores): the measured-bandwidth
// reference beside the HBM spec, and the calibration for byte counters.
__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ in, float4* __restrict__ out, long n4)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long)gridDim.x * 256) out[i] = in[i];
}

inline unsigned grid_for(long n) { return (unsigned)((n + 255) / 256); }
inline unsigned grid_fill(long n)
{
    long b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

}  // namespace

hipError_t launch_water_update(const RebuildArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(water_kernel, dim3(grid_for((long)a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_removal_mask(const RebuildArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(mask_kernel, dim3(grid_for((long)a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_apply_props(const RebuildArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(props_kernel, dim3(grid_for((long)a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_coeffs(const CoeffArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(coeff_kernel, dim3(grid_for((long)(a.nx + 2) * (a.ny + 2) * (a.nz + 2))), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_normalise(int nx, int ny, int nz, long px, long pxy, const double* tally,
                            const double* rhokap, const uint8_t* mat, const double* mua_frac,
                            const CycleHdr* dyn, double scale, const double* coeff, double* jmean,
                            float* src, hipStream_t s)
{
    hipLaunchKernelGGL(normalise_kernel, dim3(grid_for((long)nx * ny * nz)), dim3(256), 0, s, nx, ny, nz, px,
                       pxy, tally, rhokap, mat, mua_frac, dyn, scale, coeff, jmean, src);
    return hipGetLastError();
}
// k^3 block sums of an interior FP64 field: one workgroup per block
__global__ __launch_bounds__(256) void block_sum_kernel(const double* f, int nx, int ny, int k, double* out)
{
    __shared__ double red[256];
    const int bx = nx / k, by = ny / k;
    const long b = blockIdx.x;
    const int ib = (int)(b % bx), jb = (int)((b / bx) % by), kb = (int)(b / ((long)bx * by));
    double s = 0.0;
    const long kkk = (long)k * k * k;                  // k may be a whole 4096-voxel edge
    for (long q = threadIdx.x; q < kkk; q += 256) {
        const int i = (int)(q % k), j = (int)((q / k) % k), l = (int)(q / ((long)k * k));
        s += f[(long)(ib * k + i) + (long)nx * ((jb * k + j) + (long)ny * (kb * k + l))];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0];
}

hipError_t launch_block_sums(const double* f, int nx, int ny, int nz, int k, double* out, hipStream_t s)
{
    const long nb = (long)(nx / k) * (ny / k) * (nz / k);
    hipLaunchKernelGGL(block_sum_kernel, dim3((unsigned)nb), dim3(256), 0, s, f, nx, ny, k, out);
    return hipGetLastError();
}
hipError_t launch_energy(int nx, int ny, int nz, long px, long pxy, const float* dT, const double* T64,
                         double t_ref, const double* heatcap, double mass_voxel, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(energy_kernel, dim3(grid_for((long)nx * ny * nz)), dim3(256), 0, s, nx, ny, nz, px, pxy,
                       dT, T64, t_ref, heatcap, mass_voxel, out);
    return hipGetLastError();
}
namespace {
__global__ __launch_bounds__(256) void tmask_kernel(const double* __restrict__ thres,
                                                    const double* __restrict__ omega, double thr0, double thr1,
                                                    double thr2, uint8_t* __restrict__ tmask, long nvox)
{
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nvox) return;
    const unsigned m = (thres[v] != 0.0 ? 1u : 0u) | (thres[v + nvox] != 0.0 ? 2u : 0u) |
                       (thres[v + 2 * nvox] != 0.0 ? 4u : 0u);
    const double om = omega[v];
    const bool due = (!(m & 1u) && om >= thr0) || (!(m & 2u) && om >= thr1) || (!(m & 4u) && om >= thr2);
    tmask[v] = (uint8_t)(m | (due ? TAM_TMASK_DUE : 0u));
}

}  // namespace

namespace {
__global__ __launch_bounds__(256) void class_flags_kernel(uint8_t* cls, int nx, int ny, int nz, long px, long pxy)
{
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= (long)nx * ny * nz) return;
    const int i = (int)(v % nx) + 1;
    const long r = v / nx;
    const int j = (int)(r % ny) + 1, k = (int)(r / ny) + 1;
    const long c = i + px * j + pxy * k;
    // only bit 7 of a byte changes here, so the neighbours' classes read the
    // same whether or not their own flag was rewritten first; the due bit stays
    const unsigned m = cls[c] & TAM_CLASS_MASK, due = cls[c] & TAM_CLASS_DUE;
    const bool hom = (cls[c - 1] & TAM_CLASS_MASK) == m && (cls[c + 1] & TAM_CLASS_MASK) == m &&
                     (cls[c - px] & TAM_CLASS_MASK) == m && (cls[c + px] & TAM_CLASS_MASK) == m &&
                     (cls[c - pxy] & TAM_CLASS_MASK) == m && (cls[c + pxy] & TAM_CLASS_MASK) == m;
    cls[c] = (uint8_t)(m | due | (hom ? TAM_CLASS_HOM : 0u));
}

__global__ __launch_bounds__(256) void class_due_kernel(uint8_t* cls, const uint8_t* tmask, int nx, int ny, int nz,
                                                        long px, long pxy)
{
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= (long)nx * ny * nz) return;
    const int i = (int)(v % nx) + 1;
    const long r = v / nx;
    const int j = (int)(r % ny) + 1, k = (int)(r / ny) + 1;
    const long c = i + px * j + pxy * k;
    cls[c] = (uint8_t)((cls[c] & ~TAM_CLASS_DUE) | ((tmask[v] & TAM_TMASK_DUE) ? TAM_CLASS_DUE : 0u));
}
}  // namespace

hipError_t launch_class_due(uint8_t* cls, const uint8_t* tmask, int nx, int ny, int nz, long px, long pxy,
                            hipStream_t s)
{
    const long nv = (long)nx * ny * nz;
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(class_due_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, cls, tmask, nx, ny, nz,
                       px, pxy);
    return hipGetLastError();
}

hipError_t launch_class_flags(uint8_t* cls, int nx, int ny, int nz, long px, long pxy, hipStream_t s)
{
    const long nv = (long)nx * ny * nz;
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(class_flags_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, cls, nx, ny, nz, px,
                       pxy);
    return hipGetLastError();
}

hipError_t launch_tmask(const double* thres, const double* omega, double thr0, double thr1, double thr2,
                        uint8_t* tmask, long nvox, hipStream_t s)
{
    if (nvox <= 0) return hipSuccess;
    hipLaunchKernelGGL(tmask_kernel, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, s, thres, omega, thr0, thr1,
                       thr2, tmask, nvox);
    return hipGetLastError();
}


hipError_t launch_fill_f64(double* p, long n, double v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_fill(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_f32(float* p, long n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_fill(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_copy(const void* in, void* out, long bytes, hipStream_t s)
{
    const long n4 = bytes / 16;
    long b = (n4 + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)b), dim3(256), 0, s, (const float4*)in, (float4*)out, n4);
    return hipGetLastError();
}

hipError_t launch_acc_u64(unsigned long long* acc, const unsigned long long* src, long n, hipStream_t s)
{
    hipLaunchKernelGGL(acc_u64_kernel, dim3(1), dim3(64), 0, s, acc, src, n);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, long n, hipStream_t s)
{
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_fill(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
