// heat.hip -- FP32 explicit 7-point heat step and Arrhenius damage on CDNA4.
//
// Replaces the time loop of heat_sim_3D (3dFD.f90:113-215) and Arrhenius
// (3dFD.f90:424-466) for the FP32 path (temperatures held as dT = T - t_ref).
//
// The reference's update, per axis, is  a*T(-) - 2b*T + d*T(+)  with
// b = (a+d)/2 and half-node coefficients a, d built from arithmetic means of
// kappa, density and heatcap (3dFD.f90:119-164), i.e.
//     dT' = dT + sum_faces G_f * (dT_nbr - dT) + laserOn * src,   G_f = delt * a_f.
// The face coefficients G only change when properties do (once per cycle),
// so they are precomputed in one of two forms:
//   * class mode: every halo cell carries a uint8 thermal class (the distinct
//     (kappa, rho, c) triples, plus a "removed" class with no conduction) and
//     G for each (axis, class, class) pair sits in a 3x16x16 LUT in LDS.  A
//     voxel-update then moves 13 bytes (dT read + write, src, class).
//   * array mode (any per-voxel properties): FP32 arrays gx/gy/gz hold each
//     voxel's +x/+y/+z face; the -faces are the neighbours' +faces.
// Both forms hold bit-identical G values (same FP64 formula, then rounded).
//
// Kernel shapes (2.5D blocking): a TX x TY block owns an x-y tile and marches
// KZ planes in z.  stencil_kernel stages each plane of dT (and classes or
// gx/gy) in LDS with a one-cell halo so the +-x/+-y neighbours come from LDS;
// stencil_direct_kernel loads them directly (cache hits); the +-z neighbours
// stay in registers.  stencil_t2_kernel runs two steps per launch (temporal
// blocking, class mode): the first step on a footprint one cell wider than
// the owned tile, handed to the second through LDS.
#include "tam_internal.h"

namespace {

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

enum : unsigned { F_BOIL = 1u, F_ROBIN = 2u, F_DAMAGE = 4u, F_ENERGY = 8u };

// The face sum and the update in one fixed evaluation order (explicit FMAs):
// left to the compiler, a*b + c*d contracts to fma(a, b, c*d) or to
// fma(c, d, a*b) depending on the surrounding code, and every kernel form and
// launch shape must round the same way.
__device__ __forceinline__ float face_sum(float tc, float gxm, float txm, float gxp, float txp, float gym,
                                          float tym, float gyp, float typ, float gzm, float tm, float gzp, float tp)
{
    float s = gxm * (txm - tc);
    s = fmaf(gxp, txp - tc, s);
    s = fmaf(gym, tym - tc, s);
    s = fmaf(gyp, typ - tc, s);
    s = fmaf(gzm, tm - tc, s);
    return fmaf(gzp, tp - tc, s);
}
__device__ __forceinline__ float step_update(float tc, float inc, float lon, float src)
{
    return fmaf(lon, src, tc + inc);
}

// Workgroups go to the 8 XCDs round-robin by linear id, and each XCD has its
// own L2.  Tiles that are neighbours share halo cache lines, so the linear
// block id is remapped to give every XCD one contiguous range of tiles
// (x fastest, then y, then z): neighbours then meet in the same L2 instead of
// each fetching the shared lines from HBM.  The map is a bijection on
// [0, nblocks) for any block count.
struct TileId {
    int x, y, z;
    long id;
};
// 32-bit arithmetic (64-bit divides cost a few hundred scalar instructions
// per wave): every launch here has fewer than 2^32 blocks.
__device__ __forceinline__ TileId xcd_tile()
{
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned nb = gx * gy * gridDim.z;
    const unsigned w = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned q = nb >> 3, r = nb & 7u, x = w & 7u, slot = w >> 3;
    const unsigned t = x * q + (x < r ? x : r) + slot;
    const unsigned txy = t / gx;
    TileId tl;
    tl.x = (int)(t - txy * gx);
    tl.y = (int)(txy % gy);
    tl.z = (int)(txy / gy);
    tl.id = t;
    return tl;
}

// One step's damage at interior voxel v, dT = tn after the step
// (3dFD.f90:450-461): inside the Arrhenius window Omega grows and the
// threshold chain runs; outside it the chain runs only where a time is due
// (due_in: TAM_CLASS_DUE of the class byte, or the mask's TAM_TMASK_DUE
// without classes).  Omega and the mask come from the prefetch when pf.
// Returns whether a time is still due after this step.
__device__ __forceinline__ bool damage_one(const HeatArgs& a, long v, long nv, float tn, double t_after, bool pf,
                                           double om_pf, unsigned tm_pf, bool due_in)
{
    bool chain = due_in;                   // outside the window: only where a time is due
    double om = 0.0;
    bool have = false;
    if (tn >= a.dT_lo_test && tn < a.dT_hi_test) {
        const double T = a.t_ref + (double)tn;
        if (T >= a.arr_Tlo && T < a.arr_Thi) {
            // rate A exp(-dE/RT) = exp(ln A - dE/RT): exponent formed in
            // FP64 (both terms ~230), exponentiated in FP32 (|arg| < ~40,
            // <= 2e-6 relative), accumulated in FP64
            const float arg = (float)(a.lnA - a.arr_dE / (a.arr_R * T));
            om = (pf ? om_pf : a.omega[v]) + a.delt * (double)expf(arg);
            a.omega[v] = om;
            have = true;
            // the threshold chain can only fire once Omega reaches the lowest
            // threshold (and nothing is due below it)
            chain = om >= a.thr_min;
        }
    }
    if (!chain) return false;
    if (!have) om = pf ? om_pf : a.omega[v];
    const unsigned m = thres_chain(a.thres, a.tmask, v, nv, om, a.thr0, a.thr1, a.thr2, t_after,
                                   pf ? tm_pf : a.tmask[v]);
    return (m & TAM_TMASK_DUE) != 0u;
}

template <int TX, int TY, bool LUT>
__global__ __launch_bounds__(TX* TY) void stencil_kernel(const HeatArgs a, const float* __restrict__ Tin,
                                                         float* __restrict__ Tout, float lon, double t_after,
                                                         int kz, unsigned flags)
{
    if (a.dyn_lon) {              // graph replay: this step's values from device memory
        lon = *a.dyn_lon;
        t_after = *a.dyn_tafter;
    }
    __shared__ float sT[TY + 2][TX + 2];
    __shared__ uint8_t sC[LUT ? TY + 2 : 1][LUT ? TX + 2 : 1];
    __shared__ float sGx[LUT ? 1 : TY][LUT ? 1 : TX + 1];
    __shared__ float sGy[LUT ? 1 : TY + 1][LUT ? 1 : TX];
    __shared__ float sLut[LUT ? 3 * 256 + 16 : 1];   // face table + heatcap*mass per class
    __shared__ float2 sRob[LUT ? 16 : 1];

    const float* __restrict__ gxa = a.gx;
    const float* __restrict__ gya = a.gy;
    const float* __restrict__ gza = a.gz;
    const float* __restrict__ srca = a.src;
    const uint8_t* __restrict__ cla = a.cls;
    const long px = a.px, pxy = a.pxy;
    const int nx = a.nx, ny = a.ny, nz = a.nz;

    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const TileId tl = xcd_tile();
    const int i = tl.x * TX + tx + 1, j = tl.y * TY + ty + 1;
    const bool in = i <= nx && j <= ny;
    const bool bnd_ij = i == 1 || i == nx || j == 1 || j == ny;     // lane on an x or y face
    // lanes past the grid edge load the halo cell (a valid address), never store
    const int ic = i <= nx + 1 ? i : nx + 1, jc = j <= ny + 1 ? j : ny + 1;
    const int k0 = tl.z * kz + 1;
    const int k1 = min(k0 + kz, nz + 1);
    const bool robin = flags & F_ROBIN, damage = flags & F_DAMAGE, boil = flags & F_BOIL;
    const bool energy = (flags & F_ENERGY) &&
                        (tl.x == 0 || tl.x == (int)gridDim.x - 1 || tl.y == 0 ||
                         tl.y == (int)gridDim.y - 1 || k0 == 1 || k1 == nz + 1);
    // halo duties of this lane
    const bool hxm = tx == 0, hxp = tx == TX - 1 && i <= nx, hym = ty == 0, hyp = ty == TY - 1 && j <= ny;

    if (LUT) {
        for (int t = threadIdx.x; t < 3 * 256 + 16; t += TX * TY) sLut[t] = a.glut[t];
        if (threadIdx.x < 16) sRob[threadIdx.x] = a.robin_lut[threadIdx.x];
    }

    long c = ic + px * jc + pxy * k0;
    // z registers: dT and class (or gz) of the planes below / at / above
    float tm = Tin[c - pxy], tc = Tin[c], tp = Tin[c + pxy];
    uint8_t cm = 0, cc = 0, cp = 0;
    unsigned ccr = 0, cpr = 0;                  // raw class bytes (TAM_CLASS_DUE) of planes k, k+1
    float gzm = 0.f, gzc = 0.f;
    if (LUT) {
        cm = cla[c - pxy] & TAM_CLASS_MASK;
        ccr = cla[c];
        cpr = cla[c + pxy];
        cc = ccr & TAM_CLASS_MASK;
        cp = cpr & TAM_CLASS_MASK;
    } else {
        gzm = gza[c - pxy]; gzc = gza[c];
    }
    // current-plane loads (x/y halo, coefficients, source).  Halo loads are
    // guarded, never merged with a default: the value is only read by the
    // same lane under the same guard, so no wait is forced on the load.
    float hT0, hT1, hT2, hT3;
    uint8_t hC0, hC1, hC2, hC3;
    float gxc = 0.f, gyc = 0.f, hGx, hGy;
    if (hxm) hT0 = Tin[c - 1];
    if (hxp) hT1 = Tin[c + 1];
    if (hym) hT2 = Tin[c - px];
    if (hyp) hT3 = Tin[c + px];
    if (LUT) {
        if (hxm) hC0 = cla[c - 1] & TAM_CLASS_MASK;
        if (hxp) hC1 = cla[c + 1] & TAM_CLASS_MASK;
        if (hym) hC2 = cla[c - px] & TAM_CLASS_MASK;
        if (hyp) hC3 = cla[c + px] & TAM_CLASS_MASK;
    } else {
        gxc = gxa[c]; gyc = gya[c];
        if (hxm) hGx = gxa[c - 1];
        if (hym) hGy = gya[c - px];
    }
    float s = srca[c];
    float e_acc = 0.f;                          // boundary energy of this lane, J (per launch)
    // damage integral: Omega is fetched one plane ahead for voxels whose
    // temperature is already near the Arrhenius window (others load it late)
    const long nxy = (long)nx * ny;
    long v = (long)(ic - 1) + (long)nx * (jc - 1) + nxy * (k0 - 1);
    const float wlo = a.dT_lo_test - a.dT_win_pad, whi = a.dT_hi_test + a.dT_win_pad;
    bool pom = damage && in && tc >= wlo && tc < whi;
    double om_pf = 0.0;
    unsigned tm_pf = 0;
    if (pom) {
        om_pf = a.omega[v];
        tm_pf = a.tmask[v];
    }

    for (int k = k0; k < k1; ++k, c += pxy, v += nxy) {
        __syncthreads();                       // previous plane's LDS reads are done
        sT[ty + 1][tx + 1] = tc;
        if (hxm) sT[ty + 1][0] = hT0;
        if (hxp) sT[ty + 1][TX + 1] = hT1;
        if (hym) sT[0][tx + 1] = hT2;
        if (hyp) sT[TY + 1][tx + 1] = hT3;
        if (LUT) {
            sC[ty + 1][tx + 1] = cc;
            if (hxm) sC[ty + 1][0] = hC0;
            if (hxp) sC[ty + 1][TX + 1] = hC1;
            if (hym) sC[0][tx + 1] = hC2;
            if (hyp) sC[TY + 1][tx + 1] = hC3;
        } else {
            sGx[ty][tx + 1] = gxc;
            if (hxm) sGx[ty][0] = hGx;
            sGy[ty + 1][tx] = gyc;
            if (hym) sGy[0][tx] = hGy;
        }
        const float tc_k = tc, tm_k = tm, tp_k0 = tp, s_k = s;
        const int cc_k = cc, cm_k = cm, cp_k = cp;
        const unsigned ccr_k = ccr;
        const float gzm_k = gzm, gzc_k = gzc;
        const bool pom_k = pom;
        const double om_k = om_pf;
        const unsigned tmask_k = tm_pf;
        // ---- prefetch plane k+1 (addresses stay inside the halo array: k+2 <= nz+1
        //      whenever k+1 < k1 <= nz+1; the last iteration skips it)
        if (k + 1 < k1) {
            const long n = c + pxy;
            tm = tc; tc = tp; tp = Tin[n + pxy];
            if (hxm) hT0 = Tin[n - 1];
            if (hxp) hT1 = Tin[n + 1];
            if (hym) hT2 = Tin[n - px];
            if (hyp) hT3 = Tin[n + px];
            if (LUT) {
                cm = cc; cc = cp; ccr = cpr;
                cpr = cla[n + pxy];
                cp = cpr & TAM_CLASS_MASK;
                if (hxm) hC0 = cla[n - 1] & TAM_CLASS_MASK;
                if (hxp) hC1 = cla[n + 1] & TAM_CLASS_MASK;
                if (hym) hC2 = cla[n - px] & TAM_CLASS_MASK;
                if (hyp) hC3 = cla[n + px] & TAM_CLASS_MASK;
            } else {
                gzm = gzc; gzc = gza[n];
                gxc = gxa[n]; gyc = gya[n];
                if (hxm) hGx = gxa[n - 1];
                if (hym) hGy = gya[n - px];
            }
            s = srca[n];
            pom = damage && in && tc >= wlo && tc < whi;      // tc now holds plane k+1
            if (pom) {
                om_pf = a.omega[v + nxy];
                tm_pf = a.tmask[v + nxy];
            }
        }
        __syncthreads();
        if (!in) continue;
        // ---- update of voxel (i, j, k)
        const float txm = sT[ty + 1][tx], txp = sT[ty + 1][tx + 2];
        const float tym = sT[ty][tx + 1], typ = sT[ty + 2][tx + 1];
        float gxm, gxp, gym, gyp, gzm_f, gzp_f;
        if (LUT) {
            const int cxm = sC[ty + 1][tx], cxp = sC[ty + 1][tx + 2];
            const int cym = sC[ty][tx + 1], cyp = sC[ty + 2][tx + 1];
            gxm = sLut[cxm * 16 + cc_k];
            gxp = sLut[cc_k * 16 + cxp];
            gym = sLut[256 + cym * 16 + cc_k];
            gyp = sLut[256 + cc_k * 16 + cyp];
            gzm_f = sLut[512 + cm_k * 16 + cc_k];
            gzp_f = sLut[512 + cc_k * 16 + cp_k];
        } else {
            gxm = sGx[ty][tx]; gxp = sGx[ty][tx + 1];
            gym = sGy[ty][tx]; gyp = sGy[ty + 1][tx];
            gzm_f = gzm_k; gzp_f = gzc_k;
        }
        float tp_k = tp_k0;
        if (robin && k == nz) {                 // ghost node of the convective top face
            float2 rb;
            if (LUT) rb = sRob[cc_k];
            else rb = a.robin[(i - 1) + (long)nx * (j - 1)];
            tp_k = fmaf(rb.x, tc_k, rb.y);
        }
        const float inc = face_sum(tc_k, gxm, txm, gxp, txp, gym, tym, gyp, typ, gzm_f, tm_k, gzp_f, tp_k);
        float tn = step_update(tc_k, inc, lon, s_k);
        if (tc_k < a.dT_neg) a.flags[0] = 1;                    // 3dFD.f90:179-182
        if (boil && tc_k >= a.dT_boil_test) {                    // 3dFD.f90:167-172
            const double T64 = a.t_ref + (double)tc_k;
            const long v = (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1));
            const double q = a.Q[v];
            if (T64 >= 100.0 + 273.0 && q < a.qvapor) {
                const double e = (double)lon * a.jmean[v] * a.delt * a.volume_voxel +
                                 a.heatcap[c] * a.mass_voxel * (double)inc;
                if (e > 0.0) {
                    a.Q[v] = fmin(q + e, a.qvapor);
                    tn = a.dT_boil;
                }
            }
        }
        Tout[c] = tn;
        if (damage) {                                               // 3dFD.f90:450-461, every step
            const bool din = LUT ? (ccr_k & TAM_CLASS_DUE) != 0u
                                 : ((pom_k ? tmask_k : (unsigned)a.tmask[v]) & TAM_TMASK_DUE) != 0u;
            const bool dout = damage_one(a, v, nxy * nz, tn, t_after, pom_k, om_k, tmask_k, din);
            if (LUT && dout != din) a.cls[c] = (uint8_t)((ccr_k & ~TAM_CLASS_DUE) | (dout ? TAM_CLASS_DUE : 0u));
        }
        if (energy && (bnd_ij || k == 1 || k == nz)) {
            float e = 0.f;
            if (i == 1) e += gxm * (txm - tc_k);
            if (i == nx) e += gxp * (txp - tc_k);
            if (j == 1) e += gym * (tym - tc_k);
            if (j == ny) e += gyp * (typ - tc_k);
            if (k == 1) e += gzm_f * (tm_k - tc_k);
            if (k == nz) e += gzp_f * (tp_k - tc_k);
            e_acc -= (LUT ? sLut[768 + cc_k] : (float)(a.heatcap[c] * a.mass_voxel)) * e;
        }
    }
    if (energy) {
        // one FP64 slot per block, accumulated across steps (no shared-address
        // atomics); tam_energy sums the slots
        __shared__ double sE[TX * TY / 64];
        const double w = wave_sum_d((double)e_acc);
        if ((threadIdx.x & 63) == 0) sE[threadIdx.x >> 6] = w;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int q = 0; q < TX * TY / 64; ++q) t += sE[q];
            a.eblk[tl.id] += t;
        }
    }
}

// Direct form of the same update: no LDS and no barriers.  Every lane loads
// its own +-x/+-y neighbours (they share cache lines with the tile's own
// rows, so they are L1/L2 hits) from addresses clamped into the interior, so
// the plane loop is straight-line code; +-z stay in registers.
// tl: this block's tile; gx, gy: tiles per row and column of the grid.
template <int TX, int TY, bool LUT>
__device__ __forceinline__ void direct_body(const HeatArgs& a, const float* __restrict__ Tin,
                                            float* __restrict__ Tout, float lon, double t_after, int kz,
                                            unsigned flags, const TileId tl, int gx, int gy)
{
    __shared__ float sLut[LUT ? 3 * 256 + 16 : 1];   // face table + heatcap*mass per class
    __shared__ float2 sRob[LUT ? 16 : 1];
    const float* __restrict__ gxa = a.gx;
    const float* __restrict__ gya = a.gy;
    const float* __restrict__ gza = a.gz;
    const float* __restrict__ srca = a.src;
    const uint8_t* __restrict__ cla = a.cls;
    const long px = a.px, pxy = a.pxy;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i = tl.x * TX + tx + 1, j = tl.y * TY + ty + 1;
    const bool in = i <= nx && j <= ny;
    const bool bnd_ij = i == 1 || i == nx || j == 1 || j == ny;     // lane on an x or y face
    const int ic = i <= nx ? i : nx, jc = j <= ny ? j : ny;     // clamped: every neighbour is valid
    const int k0 = tl.z * kz + 1;
    const int k1 = min(k0 + kz, nz + 1);
    const bool robin = flags & F_ROBIN, damage = flags & F_DAMAGE, boil = flags & F_BOIL;
    const bool energy = (flags & F_ENERGY) &&
                        (tl.x == 0 || tl.x == gx - 1 || tl.y == 0 || tl.y == gy - 1 || k0 == 1 || k1 == nz + 1);
    if (LUT) {
        for (int t = threadIdx.x; t < 3 * 256 + 16; t += TX * TY) sLut[t] = a.glut[t];
        if (threadIdx.x < 16) sRob[threadIdx.x] = a.robin_lut[threadIdx.x];
        __syncthreads();
    }
    long c = ic + px * jc + pxy * k0;
    const long nxy = (long)nx * ny;
    long v = (long)(ic - 1) + (long)nx * (jc - 1) + nxy * (k0 - 1);
    float tm = Tin[c - pxy], tc = Tin[c];
    int cm = 0, cc = 0;
    unsigned ccr = 0;                       // class byte of (i, j, k) with its TAM_CLASS_HOM bit
    float gzm = 0.f;
    if (LUT) {
        cm = cla[c - pxy] & TAM_CLASS_MASK;
        ccr = cla[c];
        cc = ccr & TAM_CLASS_MASK;
    } else {
        gzm = gza[c - pxy];
    }
    float e_acc = 0.f;
    for (int k = k0; k < k1; ++k, c += pxy, v += nxy) {
        const float tp0 = Tin[c + pxy];
        const float txm = Tin[c - 1], txp = Tin[c + 1], tym = Tin[c - px], typ = Tin[c + px];
        const float s_k = srca[c];
        // Omega issued with the plane's loads when T is already near the window
        const bool pom = damage && in && tc >= a.dT_lo_test - a.dT_win_pad && tc < a.dT_hi_test + a.dT_win_pad;
        double om_pf = 0.0;
        unsigned tm_pf = 0;
        if (pom) {
            om_pf = a.omega[v];
            tm_pf = a.tmask[v];
        }
        float gxm, gxp, gym, gyp, gzp;
        int cp = 0;
        unsigned cpr = 0;
        if (LUT) {
            cpr = cla[c + pxy];
            cp = cpr & TAM_CLASS_MASK;
            if (__all((ccr & TAM_CLASS_HOM) != 0)) {
                // every lane's voxel has six same-class neighbours: its own
                // class's x / y / z entries, no neighbour classes to load
                gxm = gxp = sLut[cc * 17];
                gym = gyp = sLut[256 + cc * 17];
                gzm = gzp = sLut[512 + cc * 17];
            } else {
                const int cxm = cla[c - 1] & TAM_CLASS_MASK, cxp = cla[c + 1] & TAM_CLASS_MASK;
                const int cym = cla[c - px] & TAM_CLASS_MASK, cyp = cla[c + px] & TAM_CLASS_MASK;
                gxm = sLut[cxm * 16 + cc];
                gxp = sLut[cc * 16 + cxp];
                gym = sLut[256 + cym * 16 + cc];
                gyp = sLut[256 + cc * 16 + cyp];
                gzp = sLut[512 + cc * 16 + cp];
                gzm = sLut[512 + cm * 16 + cc];
            }
        } else {
            gxm = gxa[c - 1]; gxp = gxa[c]; gym = gya[c - px]; gyp = gya[c]; gzp = gza[c];
        }
        float tp = tp0;
        if (robin && k == nz) {
            float2 rb;
            if (LUT) rb = sRob[cc];
            else rb = a.robin[(ic - 1) + (long)nx * (jc - 1)];
            tp = fmaf(rb.x, tc, rb.y);
        }
        const float inc = face_sum(tc, gxm, txm, gxp, txp, gym, tym, gyp, typ, gzm, tm, gzp, tp);
        float tn = step_update(tc, inc, lon, s_k);
        if (in) {
            if (tc < a.dT_neg) a.flags[0] = 1;                   // 3dFD.f90:179-182
            if (boil && tc >= a.dT_boil_test) {                   // 3dFD.f90:167-172
                const double T64 = a.t_ref + (double)tc;
                const double q = a.Q[v];
                if (T64 >= 100.0 + 273.0 && q < a.qvapor) {
                    const double e = (double)lon * a.jmean[v] * a.delt * a.volume_voxel +
                                     a.heatcap[c] * a.mass_voxel * (double)inc;
                    if (e > 0.0) {
                        a.Q[v] = fmin(q + e, a.qvapor);
                        tn = a.dT_boil;
                    }
                }
            }
            Tout[c] = tn;
            if (damage) {                                                // 3dFD.f90:450-461
                const bool din = LUT ? (ccr & TAM_CLASS_DUE) != 0u
                                     : ((pom ? tm_pf : (unsigned)a.tmask[v]) & TAM_TMASK_DUE) != 0u;
                const bool dout = damage_one(a, v, nxy * nz, tn, t_after, pom, om_pf, tm_pf, din);
                if (LUT && dout != din) a.cls[c] = (uint8_t)((ccr & ~TAM_CLASS_DUE) | (dout ? TAM_CLASS_DUE : 0u));
            }
            if (energy && (bnd_ij || k == 1 || k == nz)) {
                float e = 0.f;
                if (i == 1) e += gxm * (txm - tc);
                if (i == nx) e += gxp * (txp - tc);
                if (j == 1) e += gym * (tym - tc);
                if (j == ny) e += gyp * (typ - tc);
                if (k == 1) e += gzm * (tm - tc);
                if (k == nz) e += gzp * (tp - tc);
                e_acc -= (LUT ? sLut[768 + cc] : (float)(a.heatcap[c] * a.mass_voxel)) * e;
            }
        }
        tm = tc;
        tc = tp0;
        if (LUT) {
            cm = cc;
            cc = cp;
            ccr = cpr;
        } else {
            gzm = gzp;
        }
    }
    if (energy) {
        __shared__ double sE[TX * TY / 64];
        const double w = wave_sum_d((double)e_acc);
        if ((threadIdx.x & 63) == 0) sE[threadIdx.x >> 6] = w;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int q = 0; q < TX * TY / 64; ++q) t += sE[q];
            a.eblk[tl.id] += t;
        }
    }
}

template <int TX, int TY, bool LUT>
__global__ __launch_bounds__(TX* TY) void stencil_direct_kernel(const HeatArgs a, const float* __restrict__ Tin,
                                                                float* __restrict__ Tout, float lon,
                                                                double t_after, int kz, unsigned flags)
{
    if (a.dyn_lon) {              // graph replay: this step's values from device memory
        lon = *a.dyn_lon;
        t_after = *a.dyn_tafter;
    }
    direct_body<TX, TY, LUT>(a, Tin, Tout, lon, t_after, kz, flags, xcd_tile(), (int)gridDim.x, (int)gridDim.y);
}

// One step's damage at interior voxel v, dT = tn after the step
// (3dFD.f90:450-461, the same arithmetic as the one-step kernels).  Omega and
// the threshold mask are loaded on first use (from the prefetch when pom)
// and kept in registers across the two steps of a launch; `dirty` says that
// Omega changed and must be stored.  Outside the window the chain runs only
// where a time is due (`due`: from the class byte for the launch's first
// step, from the first step's chain for the second).  Returns whether a time
// is still due after the step.
__device__ __forceinline__ bool damage_step(const HeatArgs& a, long v, long nv, float tn, double t_after, bool due,
                                            bool pom, double om_pf, unsigned tm_pf, double& om, bool& have,
                                            bool& dirty, unsigned& m, bool& have_m)
{
    bool chain = due;                      // outside the window: only where a time is due
    if (tn >= a.dT_lo_test && tn < a.dT_hi_test) {
        const double T = a.t_ref + (double)tn;
        if (T >= a.arr_Tlo && T < a.arr_Thi) {
            const float arg = (float)(a.lnA - a.arr_dE / (a.arr_R * T));
            if (!have) {
                om = pom ? om_pf : a.omega[v];
                have = true;
            }
            om = om + a.delt * (double)expf(arg);
            dirty = true;
            chain = om >= a.thr_min;
        }
    }
    if (!chain) return false;
    if (!have) {
        om = pom ? om_pf : a.omega[v];
        have = true;
    }
    if (!have_m) {
        m = pom ? tm_pf : a.tmask[v];
        have_m = true;
    }
    m = thres_chain(a.thres, a.tmask, v, nv, om, a.thr0, a.thr1, a.thr2, t_after, m);
    return (m & TAM_TMASK_DUE) != 0u;
}

// Two steps per launch (temporal blocking), class mode without boiling.
//
// A TX x TY block covers a footprint of TX x TY columns and owns the inner
// (TX-2) x (TY-2) of them.  Marching up z, it computes the first step on
// every footprint column (x-y neighbours loaded directly, as in direct_body),
// hands each first-step plane to the second step through a double-buffered
// LDS plane (one barrier per plane), and computes the second step one plane
// behind on the owned columns.  The intermediate field never reaches HBM:
// per two steps a voxel's dT is read and written once, its class and source
// read once, Omega read and written at most once.  The footprint ring
// repeats first-step work: TX*TY / ((TX-2)*(TY-2)) in x-y, (kz+2)/kz in z.
//
// Both steps use the one-step kernels' face coefficients (they only change in
// the rebuild) and arithmetic, in the same order, so the field, Omega and
// threshold times are bit-identical to two one-step launches.  A voxel's
// damage, negative-temperature flag and boundary energy are applied by its
// owning lane only, step 1 before step 2.
// tl: this block's tile (owned (TX-2) x (TY-2) columns); gx, gy: tiles per
// row and column of the grid.
template <int TX, int TY>
__device__ __forceinline__ void t2_body(const HeatArgs& a, const float* __restrict__ Tin, float* __restrict__ Tout,
                                        float lon0, float lon1, double ta0, double ta1, int kz, unsigned flags,
                                        const TileId tl, int gx, int gy)
{
    constexpr int OX = TX - 2, OY = TY - 2;
    __shared__ float sLut[3 * 256 + 16];
    __shared__ float2 sRob[16];
    __shared__ float sT1[2][TY][TX];            // first-step plane, buffer = plane parity
    const uint8_t* __restrict__ cla = a.cls;
    const float* __restrict__ srca = a.src;
    const long px = a.px, pxy = a.pxy;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i = tl.x * OX + tx, j = tl.y * OY + ty;            // footprint column; 0 and n+1 are halo
    const bool real = i >= 1 && i <= nx && j >= 1 && j <= ny;    // the first step updates this column
    const bool own = real && tx >= 1 && tx <= TX - 2 && ty >= 1 && ty <= TY - 2;
    const bool bnd_ij = i == 1 || i == nx || j == 1 || j == ny;
    // dT and class of the lane's own column (a halo column for lanes outside the grid)
    const long ch = (long)min(i, nx + 1) + px * min(j, ny + 1);
    // x-y neighbours and source: the nearest interior column, so every address
    // is valid; lanes that are not `real` discard what they compute from them
    const long ci = (long)min(max(i, 1), nx) + px * min(max(j, 1), ny);
    const int k0 = tl.z * kz + 1;
    const int k1 = min(k0 + kz, nz + 1);
    const bool robin = flags & F_ROBIN, damage = flags & F_DAMAGE;
    const bool energy = (flags & F_ENERGY) &&
                        (tl.x == 0 || tl.x == gx - 1 || tl.y == 0 || tl.y == gy - 1 || k0 == 1 || k1 == nz + 1);
    for (int t = threadIdx.x; t < 3 * 256 + 16; t += TX * TY) sLut[t] = a.glut[t];
    if (threadIdx.x < 16) sRob[threadIdx.x] = a.robin_lut[threadIdx.x];
    __syncthreads();

    const long nxy = (long)nx * ny, nv = nxy * nz;
    const long vcol = (long)(i - 1) + (long)nx * (j - 1);        // interior index of the column (own lanes)
    const float wlo = a.dT_lo_test - a.dT_win_pad, whi = a.dT_hi_test + a.dT_win_pad;
    // first step at plane p: dT0 at p-1, p (z0m, z0c), classes at p-1, p
    int p = k0 - 1;
    float z0m = p >= 1 ? Tin[ch + pxy * (p - 1)] : 0.f;
    float z0c = Tin[ch + pxy * p];
    int cm = p >= 1 ? cla[ch + pxy * (p - 1)] & TAM_CLASS_MASK : 0;
    unsigned ccr = cla[ch + pxy * p];
    // second step at plane k = p-1: dT1 at k-1, k (u1m, u1c) and the first
    // step's face coefficients, source and class at k
    float u1m = 0.f, u1c = 0.f;
    float Gxm = 0.f, Gxp = 0.f, Gym = 0.f, Gyp = 0.f, Gzm = 0.f, Gzp = 0.f, Sk = 0.f;
    int Ck = 0;
    unsigned Ckr = 0;                                         // raw class byte at k (TAM_CLASS_DUE)
    float e_acc = 0.f;
    // loads of plane q (x-y neighbours, source) and of the column at q+1,
    // issued one iteration ahead of their use
    float n_z0p = 0.f, n_txm = 0.f, n_txp = 0.f, n_tym = 0.f, n_typ = 0.f, n_s = 0.f;
    unsigned n_cpr = 0;
    auto issue = [&](int q) {
        if (q <= nz) {                                        // plane q+1 <= nz+1 exists
            n_z0p = Tin[ch + pxy * (q + 1)];
            n_cpr = cla[ch + pxy * (q + 1)];
        }
        if (q >= 1 && q <= nz) {
            const long c = ci + pxy * q;
            n_txm = Tin[c - 1];
            n_txp = Tin[c + 1];
            n_tym = Tin[c - px];
            n_typ = Tin[c + px];
            n_s = srca[c];
        }
    };
    issue(p);
    for (; p <= k1; ++p) {
        const bool interior = p >= 1 && p <= nz;              // else a halo plane: dT fixed
        const int k = p - 1;
        const bool step2 = p > k0;                            // second step at k in this iteration
        // ---- plane p's values (loaded last iteration); plane p+1's loads go out now
        const float z0p = n_z0p, txm = n_txm, txp = n_txp, tym = n_tym, typ = n_typ, s_p = n_s;
        const unsigned cpr = n_cpr;
        if (p < k1) issue(p + 1);
        const long v = vcol + nxy * (k - 1);
        const bool pom = step2 && damage && own && u1c >= wlo && u1c < whi;
        double om_pf = 0.0;
        unsigned tm_pf = 0;
        if (pom) {
            om_pf = a.omega[v];
            tm_pf = a.tmask[v];
        }
        // ---- first step at plane p
        float t1p = z0c;                                      // halo planes and columns keep dT
        float gxm = 0.f, gxp = 0.f, gym = 0.f, gyp = 0.f, gzm = 0.f, gzp = 0.f;
        const int cc = ccr & TAM_CLASS_MASK;
        if (interior) {
            const int cp = cpr & TAM_CLASS_MASK;
            if (__all((ccr & TAM_CLASS_HOM) != 0)) {
                gxm = gxp = sLut[cc * 17];
                gym = gyp = sLut[256 + cc * 17];
                gzm = gzp = sLut[512 + cc * 17];
            } else {
                const long c = ci + pxy * p;
                const int cxm = cla[c - 1] & TAM_CLASS_MASK, cxp = cla[c + 1] & TAM_CLASS_MASK;
                const int cym = cla[c - px] & TAM_CLASS_MASK, cyp = cla[c + px] & TAM_CLASS_MASK;
                gxm = sLut[cxm * 16 + cc];
                gxp = sLut[cc * 16 + cxp];
                gym = sLut[256 + cym * 16 + cc];
                gyp = sLut[256 + cc * 16 + cyp];
                gzp = sLut[512 + cc * 16 + cp];
                gzm = sLut[512 + cm * 16 + cc];
            }
            float tp = z0p;
            if (robin && p == nz) {
                const float2 rb = sRob[cc];
                tp = fmaf(rb.x, z0c, rb.y);
            }
            const float inc = face_sum(z0c, gxm, txm, gxp, txp, gym, tym, gyp, typ, gzm, z0m, gzp, tp);
            if (real) t1p = step_update(z0c, inc, lon0, s_p);
            if (own && p >= k0 && p < k1) {
                if (z0c < a.dT_neg) a.flags[0] = 1;             // 3dFD.f90:179-182
                if (energy && (bnd_ij || p == 1 || p == nz)) {
                    float e = 0.f;
                    if (i == 1) e += gxm * (txm - z0c);
                    if (i == nx) e += gxp * (txp - z0c);
                    if (j == 1) e += gym * (tym - z0c);
                    if (j == ny) e += gyp * (typ - z0c);
                    if (p == 1) e += gzm * (z0m - z0c);
                    if (p == nz) e += gzp * (tp - z0c);
                    e_acc -= sLut[768 + cc] * e;
                }
            }
        }
        // ---- second step at plane k = p-1 (dT1 of plane k is in sT1[k & 1])
        if (step2 && own) {
            const float xm = sT1[k & 1][ty][tx - 1], xp = sT1[k & 1][ty][tx + 1];
            const float ym = sT1[k & 1][ty - 1][tx], yp = sT1[k & 1][ty + 1][tx];
            float tp = t1p;
            if (robin && k == nz) {
                const float2 rb = sRob[Ck];
                tp = fmaf(rb.x, u1c, rb.y);
            }
            const float inc = face_sum(u1c, Gxm, xm, Gxp, xp, Gym, ym, Gyp, yp, Gzm, u1m, Gzp, tp);
            const float tn = step_update(u1c, inc, lon1, Sk);
            if (u1c < a.dT_neg) a.flags[0] = 1;
            Tout[ch + pxy * k] = tn;
            if (damage) {                                       // step 1, then step 2
                double om = 0.0;
                unsigned m = 0;
                bool have = false, dirty = false, have_m = false;
                const bool din = (Ckr & TAM_CLASS_DUE) != 0u;
                bool due = damage_step(a, v, nv, u1c, ta0, din, pom, om_pf, tm_pf, om, have, dirty, m, have_m);
                due = damage_step(a, v, nv, tn, ta1, due, pom, om_pf, tm_pf, om, have, dirty, m, have_m);
                if (dirty) a.omega[v] = om;
                if (due != din)
                    a.cls[ch + pxy * k] = (uint8_t)((Ckr & ~TAM_CLASS_DUE) | (due ? TAM_CLASS_DUE : 0u));
            }
            if (energy && (bnd_ij || k == 1 || k == nz)) {
                float e = 0.f;
                if (i == 1) e += Gxm * (xm - u1c);
                if (i == nx) e += Gxp * (xp - u1c);
                if (j == 1) e += Gym * (ym - u1c);
                if (j == ny) e += Gyp * (yp - u1c);
                if (k == 1) e += Gzm * (u1m - u1c);
                if (k == nz) e += Gzp * (tp - u1c);
                e_acc -= sLut[768 + Ck] * e;
            }
        }
        // plane p's buffer last held plane p-2, whose readers passed the
        // previous iteration's barrier
        sT1[p & 1][ty][tx] = t1p;
        __syncthreads();
        u1m = u1c;
        u1c = t1p;
        z0m = z0c;
        z0c = z0p;
        cm = cc;
        Ckr = ccr;                                            // plane p's byte, before the rotation
        ccr = cpr;
        Gxm = gxm; Gxp = gxp; Gym = gym; Gyp = gyp; Gzm = gzm; Gzp = gzp;
        Sk = s_p;
        Ck = cc;
    }
    if (energy) {
        __shared__ double sE[TX * TY / 64];
        const double w = wave_sum_d((double)e_acc);
        if ((threadIdx.x & 63) == 0) sE[threadIdx.x >> 6] = w;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int q = 0; q < TX * TY / 64; ++q) t += sE[q];
            a.eblk[tl.id] += t;
        }
    }
}

template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void stencil_t2_kernel(const HeatArgs a, const float* __restrict__ Tin,
                                                            float* __restrict__ Tout, float lon0, float lon1,
                                                            double ta0, double ta1, int kz, unsigned flags)
{
    if (a.dyn_lon) {              // graph replay: both steps' values from device memory
        lon0 = a.dyn_lon[0];
        lon1 = a.dyn_lon[1];
        ta0 = a.dyn_tafter[0];
        ta1 = a.dyn_tafter[1];
    }
    t2_body<TX, TY>(a, Tin, Tout, lon0, lon1, ta0, ta1, kz, flags, xcd_tile(), (int)gridDim.x, (int)gridDim.y);
}

// Many grids in one launch: block b works on the instance whose tile range
// [tile_start[i], tile_start[i + 1]) holds b, at that instance's local tile
// (x fastest).  Every instance uses the same tile shape and step flags.
// Returns the instance; tl, gx, gy get the local tile and the instance's
// tiles per row and column for owned tiles of ox x oy columns.
__device__ __forceinline__ int batch_tile(const HeatBatchEntry* __restrict__ tab, int n,
                                          const int* __restrict__ tile_start, int ox, int oy, TileId& tl, int& gx,
                                          int& gy)
{
    const int b = blockIdx.x;
    int lo = 0, hi = n - 1;                       // last i with tile_start[i] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const HeatBatchEntry& e = tab[lo];
    gx = (e.a.nx + ox - 1) / ox;
    gy = (e.a.ny + oy - 1) / oy;
    const int local = b - tile_start[lo];
    tl.x = local % gx;
    tl.y = (local / gx) % gy;
    tl.z = local / (gx * gy);
    tl.id = local;
    return lo;
}

template <int TX, int TY, bool LUT>
__global__ __launch_bounds__(TX* TY) void stencil_batch_kernel(const HeatBatchEntry* __restrict__ tab, int n,
                                                               const int* __restrict__ tile_start, float lon,
                                                               double t_after, int kz, unsigned flags, int parity)
{
    TileId tl;
    int gx, gy;
    const HeatBatchEntry& e = tab[batch_tile(tab, n, tile_start, TX, TY, tl, gx, gy)];
    const float* Tin = parity ? e.nxt : e.cur;
    float* Tout = parity ? e.cur : e.nxt;
    direct_body<TX, TY, LUT>(e.a, Tin, Tout, lon, t_after, kz, flags, tl, gx, gy);
}

// Two steps of every instance in one launch (class mode, as stencil_t2_kernel).
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void stencil_t2_batch_kernel(const HeatBatchEntry* __restrict__ tab, int n,
                                                                  const int* __restrict__ tile_start, float lon0,
                                                                  float lon1, double ta0, double ta1, int kz,
                                                                  unsigned flags, int parity)
{
    TileId tl;
    int gx, gy;
    const HeatBatchEntry& e = tab[batch_tile(tab, n, tile_start, TX - 2, TY - 2, tl, gx, gy)];
    const float* Tin = parity ? e.nxt : e.cur;
    float* Tout = parity ? e.cur : e.nxt;
    t2_body<TX, TY>(e.a, Tin, Tout, lon0, lon1, ta0, ta1, kz, flags, tl, gx, gy);
}

// Arrhenius at the reference cadence: every interior voxel, the full
// threshold chain (3dFD.f90:447-464).
__global__ __launch_bounds__(256) void arrhenius_kernel(HeatArgs a, const float* __restrict__ T,
                                                        const double* __restrict__ T64, double delt,
                                                        double time)
{
    if (a.dyn_arr) {              // graph replay: (delt, time) from device memory
        delt = a.dyn_arr[0];
        time = a.dyn_arr[1];
    }
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
    thres_chain(a.thres, a.tmask, v, nv, om, a.thr0, a.thr1, a.thr2, time, a.tmask[v]);
}

template <int TX, int TY, bool LUT, bool DIRECT>
hipError_t launch_t(const HeatArgs& a, const float* Tin, float* Tout, float lon, double t, int kz,
                    unsigned flags, hipStream_t s)
{
    const dim3 grid((a.nx + TX - 1) / TX, (a.ny + TY - 1) / TY, (a.nz + kz - 1) / kz);
    if (DIRECT)
        hipLaunchKernelGGL((stencil_direct_kernel<TX, TY, LUT>), grid, dim3(TX * TY), 0, s, a, Tin, Tout, lon, t,
                           kz, flags);
    else
        hipLaunchKernelGGL((stencil_kernel<TX, TY, LUT>), grid, dim3(TX * TY), 0, s, a, Tin, Tout, lon, t, kz,
                           flags);
    return hipGetLastError();
}

template <int TX, int TY>
hipError_t launch_t2(const HeatArgs& a, const float* Tin, float* Tout, float lon0, float lon1, double ta0,
                     double ta1, int kz, unsigned flags, hipStream_t s)
{
    const dim3 grid((a.nx + TX - 3) / (TX - 2), (a.ny + TY - 3) / (TY - 2), (a.nz + kz - 1) / kz);
    hipLaunchKernelGGL((stencil_t2_kernel<TX, TY>), grid, dim3(TX * TY), 0, s, a, Tin, Tout, lon0, lon1, ta0, ta1,
                       kz, flags);
    return hipGetLastError();
}

template <int TX, int TY>
hipError_t launch_t2_batch_t(const HeatBatchEntry* tab, int n, const int* tile_start, int total_tiles, float lon0,
                             float lon1, double ta0, double ta1, int kz, unsigned flags, int parity, hipStream_t s)
{
    hipLaunchKernelGGL((stencil_t2_batch_kernel<TX, TY>), dim3((unsigned)total_tiles), dim3(TX * TY), 0, s, tab, n,
                       tile_start, lon0, lon1, ta0, ta1, kz, flags, parity);
    return hipGetLastError();
}

template <int TX, int TY, bool LUT>
hipError_t launch_batch_t(const HeatBatchEntry* tab, int n, const int* tile_start, int total_tiles, float lon,
                          double t, int kz, unsigned flags, int parity, hipStream_t s)
{
    hipLaunchKernelGGL((stencil_batch_kernel<TX, TY, LUT>), dim3((unsigned)total_tiles), dim3(TX * TY), 0, s, tab, n,
                       tile_start, lon, t, kz, flags, parity);
    return hipGetLastError();
}

}  // namespace

int heat_tile_dims(int tile, int* tx, int* ty)
{
    // columns owned by one block (the two-step shapes own their footprint's inner columns)
    static const int d[9][2] = {{64, 4}, {32, 8}, {64, 8}, {64, 4}, {32, 8}, {64, 8}, {30, 6}, {30, 14}, {62, 6}};
    if (tile < 0 || tile > 8) return -1;
    *tx = d[tile][0];
    *ty = d[tile][1];
    return 0;
}

hipError_t launch_heat_batch(const HeatBatchEntry* tab, int n, const int* tile_start, int total_tiles, float lon,
                             double time_after, int boil, int robin, int damage, int energy, int lut, int tile,
                             int kz, int parity, hipStream_t s)
{
    const unsigned f = (boil ? F_BOIL : 0u) | (robin ? F_ROBIN : 0u) | (damage ? F_DAMAGE : 0u) |
                       (energy ? F_ENERGY : 0u);
    switch (tile * 2 + (lut ? 1 : 0)) {
    case 6: return launch_batch_t<64, 4, false>(tab, n, tile_start, total_tiles, lon, time_after, kz, f, parity, s);
    case 7: return launch_batch_t<64, 4, true>(tab, n, tile_start, total_tiles, lon, time_after, kz, f, parity, s);
    case 8: return launch_batch_t<32, 8, false>(tab, n, tile_start, total_tiles, lon, time_after, kz, f, parity, s);
    case 9: return launch_batch_t<32, 8, true>(tab, n, tile_start, total_tiles, lon, time_after, kz, f, parity, s);
    case 10: return launch_batch_t<64, 8, false>(tab, n, tile_start, total_tiles, lon, time_after, kz, f, parity, s);
    case 11: return launch_batch_t<64, 8, true>(tab, n, tile_start, total_tiles, lon, time_after, kz, f, parity, s);
    }
    return hipErrorInvalidValue;      // batches use the direct shapes 3-5 (and launch_heat_batch2)
}

hipError_t launch_heat_batch2(const HeatBatchEntry* tab, int n, const int* tile_start, int total_tiles, float lon0,
                              float lon1, double time_after0, double time_after1, int robin, int damage, int energy,
                              int tile, int kz, int parity, hipStream_t s)
{
    const unsigned f = (robin ? F_ROBIN : 0u) | (damage ? F_DAMAGE : 0u) | (energy ? F_ENERGY : 0u);
    switch (tile) {
    case 6: return launch_t2_batch_t<32, 8>(tab, n, tile_start, total_tiles, lon0, lon1, time_after0, time_after1, kz,
                                            f, parity, s);
    case 7: return launch_t2_batch_t<32, 16>(tab, n, tile_start, total_tiles, lon0, lon1, time_after0, time_after1,
                                             kz, f, parity, s);
    case 8: return launch_t2_batch_t<64, 8>(tab, n, tile_start, total_tiles, lon0, lon1, time_after0, time_after1, kz,
                                            f, parity, s);
    }
    return hipErrorInvalidValue;
}

// launch shapes: 0-2 stage x-y planes in LDS, 3-5 load neighbours directly,
// 6-8 run two steps per launch (stencil_t2_kernel; a single step, or a step
// outside class mode or with boiling, uses the direct shape named beside it)
int heat_num_tiles() { return 9; }
const char* heat_tile_name(int t)
{
    static const char* names[] = {"lds-64x4",    "lds-32x8",     "lds-64x8",    "direct-64x4", "direct-32x8",
                                  "direct-64x8", "two-step-32x8", "two-step-32x16", "two-step-64x8"};
    return (t >= 0 && t < 9) ? names[t] : "?";
}
int heat_tile_steps(int tile) { return tile >= 6 && tile <= 8 ? 2 : 1; }

long heat_max_blocks(int nx, int ny, int nz)
{
    // one-step shapes: tiles of at least 32 x 4; two-step shapes: at least 30 x 6 owned columns
    const long one = (long)((nx + 31) / 32) * ((ny + 3) / 4) * nz;
    const long two = (long)((nx + 29) / 30) * ((ny + 5) / 6) * nz;
    return one > two ? one : two;
}

hipError_t launch_heat_step2(const HeatArgs& a, const float* Tin, float* Tout, float lon0, float lon1,
                             double time_after0, double time_after1, int robin, int damage, int energy, int tile,
                             int kz, hipStream_t s)
{
    const unsigned f = (robin ? F_ROBIN : 0u) | (damage ? F_DAMAGE : 0u) | (energy ? F_ENERGY : 0u);
    if (kz < 1) kz = 1;
    if (kz > a.nz) kz = a.nz;
    if (!a.cls || !a.glut) return hipErrorInvalidValue;      // class mode only
    switch (tile) {
    case 6: return launch_t2<32, 8>(a, Tin, Tout, lon0, lon1, time_after0, time_after1, kz, f, s);
    case 7: return launch_t2<32, 16>(a, Tin, Tout, lon0, lon1, time_after0, time_after1, kz, f, s);
    case 8: return launch_t2<64, 8>(a, Tin, Tout, lon0, lon1, time_after0, time_after1, kz, f, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_heat_step(const HeatArgs& a, const float* Tin, float* Tout, float lon, double time_after,
                            int boil, int robin, int damage, int energy, int lut, int tile, int kz,
                            hipStream_t s)
{
    const unsigned f = (boil ? F_BOIL : 0u) | (robin ? F_ROBIN : 0u) | (damage ? F_DAMAGE : 0u) |
                       (energy ? F_ENERGY : 0u);
    if (kz < 1) kz = 1;
    if (kz > a.nz) kz = a.nz;
    if (tile == 6 || tile == 7) tile = 4;       // two-step shapes: one step on direct-32x8 / direct-64x8
    else if (tile == 8) tile = 5;
    switch (tile * 2 + (lut ? 1 : 0)) {
    case 0: return launch_t<64, 4, false, false>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 1: return launch_t<64, 4, true, false>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 2: return launch_t<32, 8, false, false>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 3: return launch_t<32, 8, true, false>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 4: return launch_t<64, 8, false, false>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 5: return launch_t<64, 8, true, false>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 6: return launch_t<64, 4, false, true>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 7: return launch_t<64, 4, true, true>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 8: return launch_t<32, 8, false, true>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 9: return launch_t<32, 8, true, true>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 10: return launch_t<64, 8, false, true>(a, Tin, Tout, lon, time_after, kz, f, s);
    case 11: return launch_t<64, 8, true, true>(a, Tin, Tout, lon, time_after, kz, f, s);
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
