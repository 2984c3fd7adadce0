// heat.hip -- explicit 7-point heat step and Arrhenius damage on CDNA4.
//
// Replaces the time loop of heat_sim_3D (3dFD.f90:113-215) and Arrhenius
// (3dFD.f90:424-466).
//
// The reference's update, per axis, is  a*T(-) - 2b*T + d*T(+)  with
// b = (a+d)/2 and half-node coefficients a, d built from arithmetic means of
// kappa, density and heatcap (3dFD.f90:119-164).  That is
// sum_faces G_f * (T_nbr - T) with G_f = delt * a_f, so the coefficients are
// built once per property change (rebuild.hip) as FP32 face arrays gx/gy/gz,
// and a step reads: dT (centre, 6 neighbours), 3 own face coefficients
// (the -x/-y/-z ones are the neighbours' +faces), and the source.
//
// Kernel shape: a block owns 256 consecutive (x,y) columns of the interior
// (x fastest, so loads are coalesced along x) and marches KZ planes in z,
// keeping T(k-1), T(k), T(k+1) and gz(k-1) in registers.  The +-x and +-y
// neighbour loads hit the same or adjacent cache lines as the block's own
// rows.  No LDS and no barrier: the exit of a lane never blocks others.
#include "tam_internal.h"

namespace {

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <bool BOIL, bool ROBIN, bool DAMAGE, bool ENERGY>
__global__ __launch_bounds__(256) void heat_step_kernel(HeatArgs a, const float* __restrict__ Tin,
                                                        float* __restrict__ Tout, float lon,
                                                        double time_after, int kz)
{
    const int nxy = a.nx * a.ny;
    const int lin = blockIdx.x * 256 + threadIdx.x;
    double e_out = 0.0;
    if (lin < nxy) {
        const int i = lin % a.nx + 1;
        const int j = lin / a.nx + 1;
        const int k0 = blockIdx.y * kz + 1;
        const int k1 = min(k0 + kz, a.nz + 1);
        const long px = a.px, pxy = a.pxy;
        long c = i + px * j + pxy * k0;
        float tm = Tin[c - pxy];
        float tc = Tin[c];
        float gzm = a.gz[c - pxy];
        const float tneg = -(float)a.t_ref;
        for (int k = k0; k < k1; ++k, c += pxy) {
            float tp;
            if (ROBIN && k == a.nz) {
                const float2 rb = a.robin[lin];
                tp = rb.x * tc + rb.y;
            } else {
                tp = Tin[c + pxy];
            }
            const float gzp = a.gz[c];
            const float gxm = a.gx[c - 1], gxp = a.gx[c];
            const float gym = a.gy[c - px], gyp = a.gy[c];
            const float txm = Tin[c - 1], txp = Tin[c + 1];
            const float tym = Tin[c - px], typ = Tin[c + px];
            const float inc = gxm * (txm - tc) + gxp * (txp - tc) + gym * (tym - tc) +
                              gyp * (typ - tc) + gzm * (tm - tc) + gzp * (tp - tc);
            float tn = tc + inc + lon * a.src[c];
            if (tc < tneg) a.flags[0] = 1;              // 3dFD.f90:179-182
            if (BOIL) {                                  // 3dFD.f90:167-172
                const double T64 = a.t_ref + (double)tc;
                if (T64 >= 100.0 + 273.0) {
                    const long v = (i - 1) + (long)a.nx * ((j - 1) + (long)a.ny * (k - 1));
                    const double q = a.Q[v];
                    if (q < a.qvapor) {
                        const double e = (double)lon * a.jmean[v] * a.delt * a.volume_voxel +
                                         a.heatcap[c] * a.mass_voxel * (double)inc;
                        if (e > 0.0) {
                            a.Q[v] = fmin(q + e, a.qvapor);
                            tn = a.dT_boil;
                        }
                    }
                }
            }
            Tout[c] = tn;
            if (DAMAGE) {                                // 3dFD.f90:450-461, every step
                const double T = a.t_ref + (double)tn;
                if (T >= a.arr_Tlo && T < a.arr_Thi) {
                    const long v = (i - 1) + (long)a.nx * ((j - 1) + (long)a.ny * (k - 1));
                    const long nv = (long)nxy * a.nz;
                    const double om = a.omega[v] + a.delt * a.arr_A * exp(-a.arr_dE / (a.arr_R * T));
                    a.omega[v] = om;
                    if (a.thres[v] == 0.0 && om >= a.thr0) a.thres[v] = time_after;
                    else if (a.thres[v + nv] == 0.0 && om >= a.thr1) a.thres[v + nv] = time_after;
                    else if (a.thres[v + 2 * nv] == 0.0 && om >= a.thr2) a.thres[v + 2 * nv] = time_after;
                }
            }
            if (ENERGY) {
                if (i == 1 || i == a.nx || j == 1 || j == a.ny || k == 1 || k == a.nz) {
                    float e = 0.f;
                    if (i == 1) e += gxm * (txm - tc);
                    if (i == a.nx) e += gxp * (txp - tc);
                    if (j == 1) e += gym * (tym - tc);
                    if (j == a.ny) e += gyp * (typ - tc);
                    if (k == 1) e += gzm * (tm - tc);
                    if (k == a.nz) e += gzp * (tp - tc);
                    e_out -= a.heatcap[c] * a.mass_voxel * (double)e;
                }
            }
            tm = tc;
            tc = tp;
            gzm = gzp;
        }
    }
    if (ENERGY) {
        const double w = wave_sum_d(e_out);
        if ((threadIdx.x & 63) == 0 && w != 0.0) unsafeAtomicAdd(&a.energy[0], w);
    }
}

// Arrhenius at the reference cadence: every interior voxel, the full
// threshold chain (3dFD.f90:447-464).
__global__ __launch_bounds__(256) void arrhenius_kernel(HeatArgs a, const float* __restrict__ T,
                                                        const double* __restrict__ T64, double delt,
                                                        double time)
{
    const long nv = (long)a.nx * a.ny * a.nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    const int i = (int)(v % a.nx) + 1;
    const long r = v / a.nx;
    const int j = (int)(r % a.ny) + 1;
    const int k = (int)(r / a.ny) + 1;
    const long c = i + a.px * j + a.pxy * k;
    const double t = T64 ? T64[c] : a.t_ref + (double)T[c];
    double om = a.omega[v];
    if (t >= a.arr_Tlo && t < a.arr_Thi && a.rhokap[c] >= 0.0) {
        om = om + delt * a.arr_A * exp(-a.arr_dE / (a.arr_R * t));
        a.omega[v] = om;
    }
    if (a.thres[v] == 0.0 && om >= a.thr0) a.thres[v] = time;
    else if (a.thres[v + nv] == 0.0 && om >= a.thr1) a.thres[v + nv] = time;
    else if (a.thres[v + 2 * nv] == 0.0 && om >= a.thr2) a.thres[v + 2 * nv] = time;
}

template <bool B, bool R, bool D, bool E>
hipError_t launch_t(const HeatArgs& a, const float* Tin, float* Tout, float lon, double t, hipStream_t s)
{
    const int nxy = a.nx * a.ny;
    const int bx = (nxy + 255) / 256;
    // z chunk: enough blocks to fill 256 CUs several times over
    int kz = (int)(((long)a.nz * bx) / 4096);
    if (kz < 2) kz = 2;
    if (kz > 32) kz = 32;
    if (kz > a.nz) kz = a.nz;
    const int bz = (a.nz + kz - 1) / kz;
    hipLaunchKernelGGL((heat_step_kernel<B, R, D, E>), dim3(bx, bz), dim3(256), 0, s, a, Tin, Tout, lon, t, kz);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_heat_step(const HeatArgs& a, const float* Tin, float* Tout, float lon,
                            double time_after, int boil, int robin, int damage, int energy,
                            hipStream_t s)
{
    const int key = (boil ? 8 : 0) | (robin ? 4 : 0) | (damage ? 2 : 0) | (energy ? 1 : 0);
    switch (key) {
#define TAM_CASE(K)                                                                        \
    case K:                                                                                \
        return launch_t<((K) & 8) != 0, ((K) & 4) != 0, ((K) & 2) != 0, ((K) & 1) != 0>( \
            a, Tin, Tout, lon, time_after, s);
        TAM_CASE(0) TAM_CASE(1) TAM_CASE(2) TAM_CASE(3) TAM_CASE(4) TAM_CASE(5) TAM_CASE(6)
        TAM_CASE(7) TAM_CASE(8) TAM_CASE(9) TAM_CASE(10) TAM_CASE(11) TAM_CASE(12) TAM_CASE(13)
        TAM_CASE(14) TAM_CASE(15)
#undef TAM_CASE
    }
    return hipErrorInvalidValue;
}

hipError_t launch_arrhenius(const HeatArgs& a, const float* T, const double* T64, double delt, double time,
                            hipStream_t s)
{
    const long nv = (long)a.nx * a.ny * a.nz;
    hipLaunchKernelGGL(arrhenius_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, a, T, T64, delt,
                       time);
    return hipGetLastError();
}
