// mc.hip -- photon-packet transport on CDNA4: one packet per lane.
//
// Replaces the photon loop of mcpolar.f90:151-170 (sourcephCO2,
// sourceph.f90:7-49; tauint1 / wall_dist / update_pos / find,
// inttau2.f90:7-239; HG scattering in place of stokes, stokes.f90:6-153).
//
// Design:
//  * Each lane owns one packet at a time and takes packets p0+gid,
//    p0+gid+stride, ... (static striding: no shared queue, so the packet ->
//    random stream mapping does not depend on scheduling).  A lane whose
//    packet ends launches the next one on its next iteration.
//  * Random numbers: Philox4x32-10, key = seed, counter = (packet id, event,
//    stream).  One 4-word block per event: launch (radius, azimuth, specular
//    test, first tau), interaction (albedo test, HG cos, HG azimuth, next
//    tau), Fresnel test at a z face.  The FP64 oracle draws identically.
//  * A wave executes the union of its lanes' paths, so every iteration has
//    the same shape: geometry for live lanes, then ONE Philox block for every
//    lane that needs random numbers (launching, interacting, or testing
//    Fresnel reflection), then the event itself.
//  * Voxel walk: integer cell + position local to the voxel (no nudge and
//    no bisection as in update_pos/find): a crossing sets the crossed axis
//    coordinate exactly to the entered face.  Voxels with mu_t = 0 (ablated)
//    are crossed in free flight.  Distances use reciprocal direction
//    components, refreshed when the direction changes.
//  * Tally: path length is summed in a register while the packet stays in a
//    voxel and written with one FP64 atomic add per voxel visit (at exit or
//    absorption).  The absorbed fraction per voxel is mu_a * L.
//  * Transcendentals use the hardware approximations (v_log, v_sin/v_cos
//    in revolutions, v_sqrt, v_rcp/v_rsq; ~1 ulp): they only shape sampled
//    distances and directions, whose distributions the parity tests check.
#include "tam_internal.h"

namespace {

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;     // one v_mad_u64_u32 each
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
                       (uint32_t)p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 23-bit uniform on (0,1), exact in FP32: ((x >> 9) + 0.5) * 2^-23.
__device__ __forceinline__ float u01(uint32_t x)
{
    return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f);
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float neg_ln(float u) { return -0.69314718f * __builtin_amdgcn_logf(u); }
// sin and cos of 2*pi*t for t in [0, 1): the hardware takes revolutions
__device__ __forceinline__ void sincos_rev(float t, float* s, float* c)
{
    *s = __builtin_amdgcn_sinf(t);
    *c = __builtin_amdgcn_cosf(t);
}

__device__ __forceinline__ float fresnel(float ni, float nt, float ci)
{
    const float si = fsqrt(fmaxf(0.f, 1.f - ci * ci));
    const float st = ni * rcp(nt) * si;
    if (st >= 1.f) return 1.f;
    const float ct = fsqrt(1.f - st * st);
    const float rs = (ni * ci - nt * ct) * rcp(ni * ci + nt * ct);
    const float rp = (ni * ct - nt * ci) * rcp(ni * ct + nt * ci);
    return 0.5f * (rs * rs + rp * rp);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr float kInf = 3.0e38f;
constexpr uint32_t kMaxEvents = 1u << 24;

enum : int { EV_NONE = 0, EV_CROSS = 1, EV_INTERACT = 2 };

__global__ __launch_bounds__(256) void mc_kernel(McArgs a)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t pid = a.p0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t pend = a.p0 + a.npk;
    const float idx_ = rcp(a.dx), idy_ = rcp(a.dy);

    uint32_t c_abs = 0, c_spec = 0, c_miss = 0, c_cross = 0, c_inter = 0, c_kill = 0;
    uint32_t c_esc[6] = {0, 0, 0, 0, 0, 0};

    bool alive = false;
    int ci = 0, cj = 0, ck = 0;
    float fx = 0.f, fy = 0.f, fz = 0.f;        // position inside the voxel, cm
    float ux = 0.f, uy = 0.f, uz = -1.f;       // direction cosines
    float iux = 0.f, iuy = 0.f, iuz = -1.f;    // their reciprocals (unused where u == 0)
    float tau = 0.f, lacc = 0.f;               // optical depth left; path in this voxel
    uint32_t ev = 0;                           // Philox event counter of this packet
    uint32_t nev = 0;                          // events of this packet (exit guarantee)

    for (;;) {
        if (!alive && pid >= pend) break;

        // ---- geometry of live lanes: distance to the exit face vs. to the
        //      next interaction (tauint1's `taurun + taucell < tau` test)
        int evt = EV_NONE, axis = 0;
        float s = 0.f, mu = 0.f;
        long v = 0;
        bool zexit = false, fres = false;
        float nin = 1.f, nout = 0.f;
        if (alive) {
            v = ci + (long)a.nx * (cj + (long)a.ny * ck);
            mu = a.mut[v];
            const float sx = ux > 0.f ? (a.dx - fx) * iux : (ux < 0.f ? -fx * iux : kInf);
            const float sy = uy > 0.f ? (a.dy - fy) * iuy : (uy < 0.f ? -fy * iuy : kInf);
            const float sz = uz > 0.f ? (a.dz - fz) * iuz : (uz < 0.f ? -fz * iuz : kInf);
            s = sx;
            if (sy < s) { s = sy; axis = 1; }
            if (sz < s) { s = sz; axis = 2; }
            evt = (mu * s < tau) ? EV_CROSS : EV_INTERACT;
            if (evt == EV_CROSS && axis == 2) {
                const bool top = uz > 0.f;
                zexit = top ? (ck == a.nz - 1) : (ck == 0);
                if (zexit) {
                    nout = top ? a.n_zhi : a.n_zlo;
                    nin = a.mtab[a.mat[v]].n;
                    fres = nout > 0.f && nout != nin;
                }
            }
        }

        // ---- one Philox block for every lane that needs random numbers
        const bool need = !alive || evt == EV_INTERACT || fres;
        uint4 r = make_uint4(0u, 0u, 0u, 0u);
        if (need) {
            r = philox4x32_10(make_uint4((uint32_t)pid, (uint32_t)(pid >> 32), ev, a.stream), a.key0, a.key1);
            ++ev;
        }

        if (!alive) {
            // ---- launch (sourcephCO2 with a beam profile), event 0
            const float u0 = u01(r.x), u1 = u01(r.y), u2 = u01(r.z), u3 = u01(r.w);
            const float rad = (a.beam == TAM_BEAM_TOPHAT) ? a.radius * fsqrt(u0)
                                                           : a.radius * fsqrt(0.5f * neg_ln(u0));
            float sn, cs;
            sincos_rev(u1, &sn, &cs);
            const float x = a.cx + rad * cs, y = a.cy + rad * sn;
            ci = (int)floorf(x * idx_);
            cj = (int)floorf(y * idy_);
            ck = a.nz - 1;
            const bool inside = x >= 0.f && y >= 0.f && ci >= 0 && ci < a.nx && cj >= 0 && cj < a.ny;
            if (!inside) {
                ++c_miss;
                pid += stride;
                ev = 0;
                continue;
            }
            const long v0 = ci + (long)a.nx * (cj + (long)a.ny * ck);
            if (u2 < a.mtab[a.mat[v0]].rsp_top) {
                ++c_spec;
                pid += stride;
                ev = 0;
                continue;
            }
            fx = fminf(fmaxf(x - ci * a.dx, 0.f), a.dx);
            fy = fminf(fmaxf(y - cj * a.dy, 0.f), a.dy);
            fz = a.dz;
            ux = 0.f; uy = 0.f; uz = -1.f;
            iux = 0.f; iuy = 0.f; iuz = -1.f;
            tau = neg_ln(u3);
            lacc = 0.f;
            nev = 0;
            alive = true;
            continue;
        }

        if (++nev > kMaxEvents) {
            // a packet that neither escapes nor is absorbed within kMaxEvents
            // events (e.g. total internal reflection in a non-absorbing slab)
            // is abandoned, so every lane's loop ends
            ++c_kill;
            alive = false;
            pid += stride;
            ev = 0;
            continue;
        }

        if (evt == EV_CROSS) {
            tau -= mu * s;
            lacc += s;
            ++c_cross;
            unsafeAtomicAdd(&a.tally[v], (double)lacc);
            lacc = 0.f;
            bool out = false;
            int face = 0;
            if (axis == 0) {
                fy = fminf(fmaxf(fy + s * uy, 0.f), a.dy);
                fz = fminf(fmaxf(fz + s * uz, 0.f), a.dz);
                if (ux > 0.f) { ++ci; fx = 0.f; } else { --ci; fx = a.dx; }
                out = ci < 0 || ci >= a.nx;
                face = ci < 0 ? TAM_FACE_XLO : TAM_FACE_XHI;
            } else if (axis == 1) {
                fx = fminf(fmaxf(fx + s * ux, 0.f), a.dx);
                fz = fminf(fmaxf(fz + s * uz, 0.f), a.dz);
                if (uy > 0.f) { ++cj; fy = 0.f; } else { --cj; fy = a.dy; }
                out = cj < 0 || cj >= a.ny;
                face = cj < 0 ? TAM_FACE_YLO : TAM_FACE_YHI;
            } else {
                fx = fminf(fmaxf(fx + s * ux, 0.f), a.dx);
                fy = fminf(fmaxf(fy + s * uy, 0.f), a.dy);
                const bool top = uz > 0.f;
                if (zexit) {
                    // z face of the grid: Fresnel test, else escape
                    if (fres && u01(r.x) < fresnel(nin, nout, fabsf(uz))) {
                        fz = top ? a.dz : 0.f;
                        uz = -uz;
                        iuz = -iuz;
                    } else {
                        out = true;
                        face = top ? TAM_FACE_ZHI : TAM_FACE_ZLO;
                    }
                } else {
                    if (top) { ++ck; fz = 0.f; } else { --ck; fz = a.dz; }
                }
            }
            if (out) {
                ++c_esc[face];
                alive = false;
                pid += stride;
                ev = 0;
            }
        } else {
            // interaction inside the voxel
            const float t = tau * rcp(mu);
            lacc += t;
            fx = fminf(fmaxf(fx + t * ux, 0.f), a.dx);
            fy = fminf(fmaxf(fy + t * uy, 0.f), a.dy);
            fz = fminf(fmaxf(fz + t * uz, 0.f), a.dz);
            ++c_inter;
            const MatEntry me = a.mtab[a.mat[v]];
            if (!(u01(r.x) < me.albedo)) {
                unsafeAtomicAdd(&a.tally[v], (double)lacc);
                lacc = 0.f;
                ++c_abs;
                alive = false;
                pid += stride;
                ev = 0;
            } else {
                // Henyey-Greenstein deflection, MCML direction-cosine update
                float ct;
                if (me.inv2g == 0.f) {
                    ct = 2.f * u01(r.y) - 1.f;
                } else {
                    const float q = me.omg2 * rcp(1.f - me.g + 2.f * me.g * u01(r.y));
                    ct = (me.opg2 - q * q) * me.inv2g;
                }
                ct = fminf(fmaxf(ct, -1.f), 1.f);
                const float st = fsqrt(1.f - ct * ct);
                float sp, cp;
                sincos_rev(u01(r.z), &sp, &cp);
                float nx_, ny_, nz_;
                if (fabsf(uz) > 0.99999f) {
                    nx_ = st * cp;
                    ny_ = st * sp;
                    nz_ = uz >= 0.f ? ct : -ct;
                } else {
                    const float w2 = 1.f - uz * uz;
                    const float rt = __builtin_amdgcn_rsqf(w2);    // 1/sqrt(1-uz^2)
                    nx_ = st * (ux * uz * cp - uy * sp) * rt + ux * ct;
                    ny_ = st * (uy * uz * cp + ux * sp) * rt + uy * ct;
                    nz_ = -st * cp * (w2 * rt) + uz * ct;
                }
                const float inv = __builtin_amdgcn_rsqf(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
                ux = nx_ * inv; uy = ny_ * inv; uz = nz_ * inv;
                iux = rcp(ux); iuy = rcp(uy); iuz = rcp(uz);
                tau = neg_ln(u01(r.w));
            }
        }
    }

    // ---- per-wave reduction of the fate and work counters ----------------
    const uint32_t vals[TAM_CNT_N] = {c_abs, c_spec, c_esc[0], c_esc[1], c_esc[2], c_esc[3],
                                      c_esc[4], c_esc[5], c_miss, c_cross, c_inter, c_kill};
    const bool leader = (threadIdx.x & 63) == 0;
#pragma unroll
    for (int i = 0; i < TAM_CNT_N; ++i) {
        const uint32_t w = wave_sum(vals[i]);
        if (leader && w) atomicAdd(&a.counters[i], (unsigned long long)w);
    }
}

}  // namespace

hipError_t launch_mc(const McArgs& a, hipStream_t s)
{
    if (a.npk == 0) return hipSuccess;
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || ncu <= 0) ncu = 256;
    // blocks_per_cu resident 256-lane blocks per CU (autotuned); never more
    // lanes than packets
    long blocks = (long)ncu * (a.blocks_per_cu > 0 ? a.blocks_per_cu : 8);
    const long need = (long)((a.npk + 255) / 256);
    if (need < blocks) blocks = need;
    hipLaunchKernelGGL(mc_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
