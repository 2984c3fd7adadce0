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
//    packet ends launches its next one in the same loop iteration, so lanes
//    of a wave stay busy until their share runs out.
//  * Random numbers: Philox4x32-10, key = seed, counter = (packet id, event,
//    stream).  One 4-word block per event: launch (radius, azimuth, specular
//    test, first tau), interaction (albedo test, HG cos, HG azimuth, next
//    tau), Fresnel test at a z face.  The FP64 oracle draws identically.
//  * Voxel walk: integer cell + position local to the voxel (no nudge and
//    no bisection as in update_pos/find): a crossing sets the crossed axis
//    coordinate exactly to the entered face.  Voxels with mu_t = 0 (ablated)
//    are crossed in free flight.
//  * Tally: path length is summed in a register while the packet stays in a
//    voxel and written with one FP64 atomic add per voxel visit (at exit or
//    absorption).  The absorbed fraction per voxel is mu_a * L.
#include "tam_internal.h"

namespace {

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
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

__device__ __forceinline__ float fresnel(float ni, float nt, float ci)
{
    const float si = sqrtf(fmaxf(0.f, 1.f - ci * ci));
    const float st = ni / nt * si;
    if (st >= 1.f) return 1.f;
    const float ct = sqrtf(1.f - st * st);
    const float rs = (ni * ci - nt * ct) / (ni * ci + nt * ct);
    const float rp = (ni * ct - nt * ci) / (ni * ct + nt * ci);
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

__global__ __launch_bounds__(256) void mc_kernel(McArgs a)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t pid = a.p0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t pend = a.p0 + a.npk;

    uint32_t c_abs = 0, c_spec = 0, c_miss = 0, c_cross = 0, c_inter = 0, c_kill = 0;
    uint32_t c_esc[6] = {0, 0, 0, 0, 0, 0};

    bool alive = false;
    int ci = 0, cj = 0, ck = 0;
    float fx = 0.f, fy = 0.f, fz = 0.f;        // position inside the voxel, cm
    float ux = 0.f, uy = 0.f, uz = -1.f;       // direction cosines
    float tau = 0.f, lacc = 0.f;               // optical depth left; path in this voxel
    uint32_t ev = 0;                           // Philox event counter of this packet
    uint32_t nev = 0;                          // events of this packet (exit guarantee)

    for (;;) {
        if (!alive) {
            if (pid >= pend) break;
            // ---- launch (sourcephCO2 with a beam profile) --------------
            const uint4 r = philox4x32_10(make_uint4((uint32_t)pid, (uint32_t)(pid >> 32), 0u, a.stream),
                                          a.key0, a.key1);
            ev = 1;
            const float u0 = u01(r.x), u1 = u01(r.y), u2 = u01(r.z), u3 = u01(r.w);
            const float rad = (a.beam == TAM_BEAM_TOPHAT) ? a.radius * sqrtf(u0)
                                                           : a.radius * sqrtf(-0.5f * __logf(u0));
            float sn, cs;
            sincospif(2.f * u1, &sn, &cs);
            const float x = a.cx + rad * cs, y = a.cy + rad * sn;
            ci = (int)floorf(x / a.dx);
            cj = (int)floorf(y / a.dy);
            ck = a.nz - 1;
            const bool inside = x >= 0.f && y >= 0.f && ci >= 0 && ci < a.nx && cj >= 0 && cj < a.ny;
            if (!inside) {
                ++c_miss;
                pid += stride;
                continue;
            }
            fx = fminf(fmaxf(x - ci * a.dx, 0.f), a.dx);
            fy = fminf(fmaxf(y - cj * a.dy, 0.f), a.dy);
            fz = a.dz;
            const long v0 = ci + (long)a.nx * (cj + (long)a.ny * ck);
            const float nin = a.mtab[a.mat[v0]].n;
            if (a.n_zhi > 0.f && a.n_zhi != nin) {
                const float rr = (nin - a.n_zhi) / (nin + a.n_zhi);
                if (u2 < rr * rr) {
                    ++c_spec;
                    pid += stride;
                    continue;
                }
            }
            ux = 0.f; uy = 0.f; uz = -1.f;
            tau = -logf(u3);
            lacc = 0.f;
            nev = 0;
            alive = true;
        }
        if (++nev > kMaxEvents) {
            // a packet that neither escapes nor is absorbed within kMaxEvents
            // events (e.g. total internal reflection in a non-absorbing slab)
            // is abandoned, so every lane's loop ends
            ++c_kill;
            alive = false;
            pid += stride;
            continue;
        }

        // ---- one event: cross a face or interact inside the voxel ------
        const long v = ci + (long)a.nx * (cj + (long)a.ny * ck);
        const float mu = a.mut[v];
        const float sx = ux > 0.f ? (a.dx - fx) / ux : (ux < 0.f ? -fx / ux : kInf);
        const float sy = uy > 0.f ? (a.dy - fy) / uy : (uy < 0.f ? -fy / uy : kInf);
        const float sz = uz > 0.f ? (a.dz - fz) / uz : (uz < 0.f ? -fz / uz : kInf);
        float s = sx;
        int axis = 0;
        if (sy < s) { s = sy; axis = 1; }
        if (sz < s) { s = sz; axis = 2; }

        if (mu * s < tau) {
            // crossing (tauint1's `taurun + taucell < tau` branch)
            tau -= mu * s;
            lacc += s;
            ++c_cross;
            unsafeAtomicAdd(&a.tally[v], (double)lacc);
            lacc = 0.f;
            if (axis == 0) {
                fy = fminf(fmaxf(fy + s * uy, 0.f), a.dy);
                fz = fminf(fmaxf(fz + s * uz, 0.f), a.dz);
                if (ux > 0.f) { ++ci; fx = 0.f; } else { --ci; fx = a.dx; }
                if (ci < 0 || ci >= a.nx) {
                    ++c_esc[ci < 0 ? TAM_FACE_XLO : TAM_FACE_XHI];
                    alive = false;
                }
            } else if (axis == 1) {
                fx = fminf(fmaxf(fx + s * ux, 0.f), a.dx);
                fz = fminf(fmaxf(fz + s * uz, 0.f), a.dz);
                if (uy > 0.f) { ++cj; fy = 0.f; } else { --cj; fy = a.dy; }
                if (cj < 0 || cj >= a.ny) {
                    ++c_esc[cj < 0 ? TAM_FACE_YLO : TAM_FACE_YHI];
                    alive = false;
                }
            } else {
                fx = fminf(fmaxf(fx + s * ux, 0.f), a.dx);
                fy = fminf(fmaxf(fy + s * uy, 0.f), a.dy);
                const int kold = ck;
                if (uz > 0.f) { ++ck; fz = 0.f; } else { --ck; fz = a.dz; }
                if (ck < 0 || ck >= a.nz) {
                    const bool top = ck >= a.nz;
                    const float nout = top ? a.n_zhi : a.n_zlo;
                    const float nin = a.mtab[a.mat[v]].n;
                    bool reflect = false;
                    if (nout > 0.f && nout != nin) {
                        const uint4 r = philox4x32_10(
                            make_uint4((uint32_t)pid, (uint32_t)(pid >> 32), ev, a.stream), a.key0, a.key1);
                        ++ev;
                        reflect = u01(r.x) < fresnel(nin, nout, fabsf(uz));
                    }
                    if (reflect) {
                        ck = kold;
                        fz = top ? a.dz : 0.f;
                        uz = -uz;
                    } else {
                        ++c_esc[top ? TAM_FACE_ZHI : TAM_FACE_ZLO];
                        alive = false;
                    }
                }
            }
            if (!alive) pid += stride;
        } else {
            // interaction inside the voxel
            const float t = tau / mu;
            lacc += t;
            fx = fminf(fmaxf(fx + t * ux, 0.f), a.dx);
            fy = fminf(fmaxf(fy + t * uy, 0.f), a.dy);
            fz = fminf(fmaxf(fz + t * uz, 0.f), a.dz);
            ++c_inter;
            const uint4 r = philox4x32_10(make_uint4((uint32_t)pid, (uint32_t)(pid >> 32), ev, a.stream),
                                          a.key0, a.key1);
            ++ev;
            const MatEntry me = a.mtab[a.mat[v]];
            if (!(u01(r.x) < me.albedo)) {
                unsafeAtomicAdd(&a.tally[v], (double)lacc);
                lacc = 0.f;
                ++c_abs;
                alive = false;
                pid += stride;
            } else {
                // Henyey-Greenstein deflection, MCML direction-cosine update
                const float g = me.g;
                float ct;
                if (g == 0.f) {
                    ct = 2.f * u01(r.y) - 1.f;
                } else {
                    const float q = (1.f - g * g) / (1.f - g + 2.f * g * u01(r.y));
                    ct = (1.f + g * g - q * q) / (2.f * g);
                }
                ct = fminf(fmaxf(ct, -1.f), 1.f);
                const float st = sqrtf(1.f - ct * ct);
                float sp, cp;
                sincospif(2.f * u01(r.z), &sp, &cp);
                float nx_, ny_, nz_;
                if (fabsf(uz) > 0.99999f) {
                    nx_ = st * cp;
                    ny_ = st * sp;
                    nz_ = uz >= 0.f ? ct : -ct;
                } else {
                    const float tmp = sqrtf(1.f - uz * uz);
                    nx_ = st * (ux * uz * cp - uy * sp) / tmp + ux * ct;
                    ny_ = st * (uy * uz * cp + ux * sp) / tmp + uy * ct;
                    nz_ = -st * cp * tmp + uz * ct;
                }
                const float inv = rsqrtf(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
                ux = nx_ * inv; uy = ny_ * inv; uz = nz_ * inv;
                tau = -logf(u01(r.w));
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
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (ncu <= 0) ncu = 256;
    // 8 blocks of 256 lanes per CU keeps every SIMD with several waves in
    // flight; never more lanes than packets.
    long blocks = (long)ncu * 8;
    const long need = (long)((a.npk + 255) / 256);
    if (need < blocks) blocks = need;
    hipLaunchKernelGGL(mc_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
