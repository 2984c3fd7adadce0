// heat64.hip -- FP64 explicit heat step, operation for operation the
// reference's heat_sim_3D (3dFD.f90:113-187), for the reference-water mode.
//
// The boiling branch (3dFD.f90:169-183) clamps T at 373 K until the voxel's
// latent-heat budget is spent: the update is discontinuous at 373 K, and near
// the beam a single step raises T by tens of kelvin.  A temperature that
// differs from the FP64 reference by one FP32 rounding can take the other
// branch and end its plateau a whole step earlier or later.  This path
// therefore keeps absolute T in FP64, evaluates
//     u = a*T(-) - 2b*T + d*T(+),  b = (a+d)/2,   per axis,
//     tempIncrease = delt*((u_xx + u_yy) + u_zz)
// with the same face coefficients a, d (FP64, built by coeff64_kernel in
// the reference's operation order) and no fused multiply-add, so each step
// rounds exactly as the Fortran does.
#include "tam_internal.h"

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <bool BOIL, bool ROBIN, bool DAMAGE, bool ENERGY>
__global__ __launch_bounds__(256) void heat64_kernel(Heat64Args a, const double* __restrict__ Tin,
                                                     double* __restrict__ Tout, double lon, double time_after)
{
    if (a.dyn_lon) {              // graph replay: this step's values from device memory
        lon = (double)*a.dyn_lon;
        time_after = *a.dyn_tafter;
    }
    const long nv = (long)a.nx * a.ny * a.nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    double e_out = 0.0;
    if (v < nv) {
        const int i = (int)(v % a.nx) + 1;
        const long r = v / a.nx;
        const int j = (int)(r % a.ny) + 1;
        const int k = (int)(r / a.ny) + 1;
        const long px = a.px, pxy = a.pxy;
        const long c = i + px * j + pxy * k;
        const double Tc = Tin[c];
        double Tzp = Tin[c + pxy];
        if (ROBIN && k == a.nz) {                      // ghost node of the convective face
            const double kk = a.kappa[c] / a.dz;
            const double den = kk + 0.5 * a.h_top;
            Tzp = ((kk - 0.5 * a.h_top) * Tc + a.h_top * a.t_amb) / den;
        }
        double ad, dd, b;
        ad = a.az[c - pxy]; dd = a.az[c]; b = 0.5 * (ad + dd);
        const double u_zz = ad * Tin[c - pxy] - 2.0 * b * Tc + dd * Tzp;
        ad = a.ay[c - px]; dd = a.ay[c]; b = 0.5 * (ad + dd);
        const double u_yy = ad * Tin[c - px] - 2.0 * b * Tc + dd * Tin[c + px];
        ad = a.ax[c - 1]; dd = a.ax[c]; b = 0.5 * (ad + dd);
        const double u_xx = ad * Tin[c - 1] - 2.0 * b * Tc + dd * Tin[c + 1];
        const double tempIncrease = a.delt * (u_xx + u_yy + u_zz);
        const double jt = a.jmean[v];
        double tn = Tc;
        if (Tc < 0.0) a.flags[0] = 1;                    // 3dFD.f90:179-182
        if (BOIL && tn >= 100. + 273. && a.Q[v] < a.qvapor) {
            const double energyIncrease = lon * jt * a.delt * a.volume_voxel +
                                          a.heatcap[c] * a.mass_voxel * tempIncrease;
            if (energyIncrease > 0.0) {
                a.Q[v] = fmin(a.Q[v] + energyIncrease, a.qvapor);
                tn = 100.0 + 273.0;
            } else {
                tn = tn + tempIncrease + lon * a.coeff[c] * jt;
            }
        } else {
            tn = tn + tempIncrease + lon * a.coeff[c] * jt;
        }
        Tout[c] = tn;
        if (DAMAGE) {
            if (tn >= a.arr_Tlo && tn < a.arr_Thi) {
                const double om = a.omega[v] + a.delt * a.arr_A * exp(-a.arr_dE / (a.arr_R * tn));
                a.omega[v] = om;
                thres_chain(a.thres, a.tmask, v, nv, om, a.thr0, a.thr1, a.thr2, time_after, a.tmask[v]);
            } else {
                // outside the window the reference's chain still runs: a
                // threshold left due by an earlier step is written now
                const unsigned m = a.tmask[v];
                if (m & TAM_TMASK_DUE)
                    thres_chain(a.thres, a.tmask, v, nv, a.omega[v], a.thr0, a.thr1, a.thr2, time_after, m);
            }
        }
        if (ENERGY && (i == 1 || i == a.nx || j == 1 || j == a.ny || k == 1 || k == a.nz)) {
            double e = 0.0;
            if (k == 1) e += a.az[c - pxy] * (Tin[c - pxy] - Tc);
            if (k == a.nz) e += a.az[c] * (Tzp - Tc);
            if (j == 1) e += a.ay[c - px] * (Tin[c - px] - Tc);
            if (j == a.ny) e += a.ay[c] * (Tin[c + px] - Tc);
            if (i == 1) e += a.ax[c - 1] * (Tin[c - 1] - Tc);
            if (i == a.nx) e += a.ax[c] * (Tin[c + 1] - Tc);
            e_out -= a.heatcap[c] * a.mass_voxel * a.delt * e;
        }
    }
    if (ENERGY) {
        const double w = wave_sum_d(e_out);
        if ((threadIdx.x & 63) == 0 && w != 0.0) unsafeAtomicAdd(&a.energy[0], w);
    }
}

// a_f = dscale * kbar/(rhobar*cbar) / h^2 per face, in the reference's
// operation order (3dFD.f90:119-129): one thread per halo cell, + faces.
__global__ __launch_bounds__(256) void coeff64_kernel(Coeff64Args a)
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
        return a.dscale * (kb / (db * hb)) * ih2;
    };
    const bool jin = j >= 1 && j <= a.ny, kin = k >= 1 && k <= a.nz, iin = i >= 1 && i <= a.nx;
    auto removed = [&](long cc, bool interior) { return a.removed_inert && interior && a.rhokap[cc] == 0.0; };
    const bool rc = removed(c, iin && jin && kin);
    a.ax[c] = (i <= a.nx && jin && kin && !rc && !removed(c + 1, i + 1 <= a.nx && jin && kin))
                  ? face(1, a.idx2) : 0.0;
    a.ay[c] = (j <= a.ny && iin && kin && !rc && !removed(c + a.px, j + 1 <= a.ny && iin && kin))
                  ? face(a.px, a.idy2) : 0.0;
    a.az[c] = (k <= a.nz && iin && jin && !rc && !removed(c + a.pxy, k + 1 <= a.nz && iin && jin))
                  ? face(a.pxy, a.idz2) : 0.0;
}

template <bool B, bool R, bool D, bool E>
hipError_t launch_t(const Heat64Args& a, const double* Tin, double* Tout, double lon, double t, hipStream_t s)
{
    const long nv = (long)a.nx * a.ny * a.nz;
    hipLaunchKernelGGL((heat64_kernel<B, R, D, E>), dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, a, Tin,
                       Tout, lon, t);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_heat64_step(const Heat64Args& a, const double* Tin, double* Tout, double lon,
                              double time_after, int boil, int robin, int damage, int energy, hipStream_t s)
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

hipError_t launch_coeffs64(const Coeff64Args& a, hipStream_t s)
{
    const long n = (long)(a.nx + 2) * (a.ny + 2) * (a.nz + 2);
    hipLaunchKernelGGL(coeff64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}
