// mc.hip -- photon-packet transport on CDNA4: one packet per lane.
//
// Replaces the photon loop of mcpolar.f90:151-170 (sourcephCO2,
// sourceph.f90:7-49; tauint1 / wall_dist / update_pos / find,
// inttau2.f90:7-239; HG scattering in place of stokes, stokes.f90:6-153).
//
// Design:
//  * Each lane owns one packet at a time.  Packets are handed out from a
//    queue head in device memory, a chunk per wave at a time, so waves that
//    draw short histories take more packets and the grid drains together.
//    The random stream of a packet depends only on its index, never on the
//    lane or the order in which packets are taken.
//  * Random numbers: Philox4x32-10, key = seed, counter = (packet id, event,
//    stream).  One 4-word block per event: launch (radius, azimuth, specular
//    test, first tau), interaction (albedo test, HG cos, HG azimuth, next
//    tau), Fresnel test at a z face.  The FP64 oracle draws identically.
//  * Two phases.  A walk step moves every lane that needs no random numbers
//    across one voxel face (or finds that its next interaction lies inside
//    the current voxel, or that it reached a z face of the grid where a
//    Fresnel test is due).  Such a lane then waits with its event pending.
//    The random phase (one Philox block plus the launch, interaction or
//    Fresnel event for every pending lane) runs once at least `rng_batch`
//    lanes of the wave wait, or every lane with a packet does: the costly
//    part then runs with most lanes active instead of for one or two.
//  * Voxel walk (Amanatides-Woo DDA): per packet segment an origin o (cm,
//    grid coordinates), the distance t travelled from it, and the distances
//    Tx, Ty, Tz from o to the next face crossed on each axis.  A crossing
//    takes the smallest T, adds one voxel's span on that axis to it, and
//    steps the cell and the linear voxel index.  Scattering or a reflection
//    starts a new segment.  Voxels with mu_t = 0 (ablated) are crossed in
//    free flight.
//  * Tally: path length is summed in a register while the packet stays in a
//    voxel and written with one FP64 atomic add per voxel visit (at exit or
//    absorption).  The absorbed fraction per voxel is mu_a * L.  Voxels with
//    mu_t = 0 absorb nothing and are crossed without a tally.  The global
//    atomic rate (~2.4e10/s on MI355X whatever the address spread) bounds
//    a kernel that sends every visit to the device tally.  Each
//    block therefore sums the visits to a cube of voxels on the beam axis in
//    LDS (mc_box_kernel), placed in z where the last run's path was largest
//    (box_plane_kernel, box_window_kernel), and adds the cube to the tally
//    once at exit.  By default the box holds fixed-point sums in 16-bit
//    integers, two voxels per 32-bit word (units of 2^-13 of dx + dy + dz):
//    four times the voxels of an FP64 box in the same LDS, integer LDS adds
//    cost less than FP64 ones, and an integer sum does not depend on the
//    order of the adds.  Each add that wraps a 16-bit slot hands 2^16 units
//    to the device tally (box_carry), so no sum is lost.
//  * Transcendentals use the hardware approximations (v_log, v_sin/v_cos
//    in revolutions, v_sqrt, v_rcp/v_rsq; ~1 ulp): they only shape sampled
//    distances and directions, whose distributions the parity tests check.
#include "tam_internal.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace {

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

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// the wave's lowest active lane holds v; every lane gets it
__device__ __forceinline__ uint64_t bcast_first(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// distance from o (inside [lo, lo + d]) to the face crossed first along u
__device__ __forceinline__ float face_dist(float o, float lo, float d, float u, float iu)
{
    return u > 0.f ? (lo + d - o) * iu : (u < 0.f ? (lo - o) * iu : 3.0e38f);
}

constexpr uint32_t kMaxEvents = 1u << 24;   // walk steps + random events of one packet
constexpr uint32_t kChunk = 256;            // packets a wave takes from the queue at once
constexpr int kBoxUnitBits = 13;            // fixed-point box unit: 2^-13 of dx + dy + dz

enum : int { PEND_NONE = 0, PEND_LAUNCH = 1, PEND_INTERACT = 2, PEND_FRESNEL = 3 };

// The transport of one instance by the calling wave's lanes, until its
// packet queue is drained; the fate counters go to a.counters.
// The LDS tally box of a launch: voxels [o, o + n).
struct BoxGeom {
    int o[3], n[3];
};

// A fixed-point box add whose carry test is still to run: the slot (-1:
// none), its word's value before the add, the add, and the tally voxel.
struct BoxAdd {
    int i;
    uint32_t old, q, v;
};

// Path length into the tally: voxels inside the block's LDS box are summed
// there (BT = uint16_t: 16-bit fixed point, two slots per 32-bit word, slot
// bi in half bi & 1; BT = double: FP64), the rest go to the FP64 device tally
// directly.  Returns 1 for a deposit into the box.  A fixed-point add leaves
// its carry test in pa (box_carry), so that the LDS return is awaited after
// the lane's other work; a deposit of 2^16 units or more (a path over 8
// (dx + dy + dz) in one visit) goes to the tally directly.
template <bool BOX, class BT>
__device__ __forceinline__ uint32_t tally_add(const McArgs& a, BT* box, const BoxGeom& g, uint32_t v, int ci,
                                              int cj, int ck, float l, BoxAdd& pa)
{
    if (BOX) {
        const unsigned bx = (unsigned)(ci - g.o[0]), by = (unsigned)(cj - g.o[1]), bz = (unsigned)(ck - g.o[2]);
        if (bx < (unsigned)g.n[0] && by < (unsigned)g.n[1] && bz < (unsigned)g.n[2]) {
            const unsigned bi = bx + (unsigned)g.n[0] * (by + (unsigned)g.n[1] * bz);
            if constexpr (std::is_same<BT, uint16_t>::value) {
                const float qf = rintf(l * a.box_scale);
                if (qf < 65536.f) {
                    pa.i = (int)bi;
                    pa.q = (uint32_t)qf;
                    pa.v = v;
                    pa.old = __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(box) + (bi >> 1),
                                                    pa.q << ((bi & 1u) * 16u), __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_WORKGROUP);
                    return 1u;
                }
            } else {
                __hip_atomic_fetch_add(&box[bi], (BT)l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                return 1u;
            }
        }
    }
    unsafeAtomicAdd(&a.tally[v], (double)l);
    return 0u;
}

// The carry test of a 16-bit box add.  A word holds the slots of two voxels
// that are x-neighbours: the even slot in the low half, the next voxel in the
// high half.  An add to the low half that wraps it carries 1 into the high
// half: the low voxel's 2^16 units go to the tally, and 1 unit comes off the
// high voxel there.  An add (to either half) that takes the word past 2^32
// wraps the high half: the high voxel's 2^16 units go to the tally.  The
// returned old value shows each wrap to exactly the add that made it, so the
// slots plus these carries always equal the sums of the adds.
template <class BT>
__device__ __forceinline__ void box_carry(const McArgs& a, BT* box, BoxAdd& pa)
{
    if constexpr (std::is_same<BT, uint16_t>::value) {
        if (pa.i >= 0) {
            const uint32_t sh = ((uint32_t)pa.i & 1u) * 16u;
            const uint32_t vlo = sh ? pa.v - 1u : pa.v, vhi = vlo + 1u;
            if (sh == 0u && (pa.old & 0xffffu) + pa.q >= 0x10000u) {
                unsafeAtomicAdd(&a.tally[vlo], 65536.0 * a.box_quantum);
                unsafeAtomicAdd(&a.tally[vhi], -a.box_quantum);
            }
            if ((uint64_t)pa.old + ((uint64_t)pa.q << sh) >= 0x100000000ull)
                unsafeAtomicAdd(&a.tally[vhi], 65536.0 * a.box_quantum);
        }
        pa.i = -1;
    }
}

// A box slot's sum in path length
__device__ __forceinline__ double box_path(uint16_t s, const McArgs& a) { return (double)s * a.box_quantum; }
__device__ __forceinline__ double box_path(double s, const McArgs&) { return s; }

template <bool BOX, class BT>
__device__ __forceinline__ void transport(const McArgs& a, BT* box, const BoxGeom& g)
{
    unsigned long long* head = a.counters + TAM_CNT_N;     // queue head, packets taken
    const uint32_t nxy = (uint32_t)a.nx * (uint32_t)a.ny;
    const float idx_ = rcp(a.dx), idy_ = rcp(a.dy);
    const float ztop = a.dz * (float)a.nz;
    const uint32_t batch = (uint32_t)a.rng_batch;
    const uint64_t npk = a.dyn_npk ? (uint64_t)*a.dyn_npk : a.npk;
    const uint32_t stream = a.dyn_stream ? *a.dyn_stream : a.stream;

    uint32_t c_abs = 0, c_spec = 0, c_miss = 0, c_cross = 0, c_inter = 0, c_kill = 0, c_tally = 0, c_box = 0;
    uint32_t c_esc[6] = {0, 0, 0, 0, 0, 0};
    BoxAdd pa{-1, 0u, 0u, 0u};

    // wave-uniform queue state: packets [pool, pool_end) are this wave's
    uint64_t pool = 0, pool_end = 0;

    bool have = false;                         // the lane holds a packet
    int pend = PEND_NONE;
    uint64_t pid = 0;
    uint32_t ev = 0, nev = 0;                  // Philox event counter; steps of this packet
    int ci = 0, cj = 0, ck = 0;
    uint32_t v = 0;                            // linear voxel index of (ci, cj, ck)
    float mu = 0.f;                            // mu_t of voxel v
    int mat = 0;                               // material of v, loaded when an event is pending
    float ox = 0.f, oy = 0.f, oz = 0.f;        // segment origin, cm
    float t = 0.f;                             // distance travelled from the origin
    float Tx = 0.f, Ty = 0.f, Tz = 0.f;        // distance from the origin to the next face per axis
    float ux = 0.f, uy = 0.f, uz = -1.f;
    float iux = 0.f, iuy = 0.f, iuz = -1.f;    // reciprocals (unused where u == 0)
    float tau = 0.f, lacc = 0.f;               // optical depth left; path in this voxel
    float nin = 1.f, nout = 1.f;               // indices at a pending Fresnel test

    for (;;) {
        // ---- lanes without a packet take the next ones from the queue
        const uint64_t wm = ballot(!have);
        if (wm) {
            const uint32_t cnt = (uint32_t)__popcll(wm);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(wm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)wm, 0u));
            const uint64_t avail = pool_end - pool;
            uint64_t mine = pool + rank;
            if (avail < cnt) {
                uint64_t nb = 0;
                if (rank == 0 && !have) nb = atomicAdd(head, (unsigned long long)kChunk);
                // the lowest lane wanting a packet is not always the lowest
                // active lane: pass the chunk base through the wave
                nb = __shfl(nb, (int)(__ffsll((long long)wm) - 1), 64);
                nb = bcast_first(nb);
                if (rank >= avail) mine = nb + (rank - avail);
                pool = nb + (cnt - avail);
                pool_end = nb + kChunk;
                if (pool_end > npk) pool_end = npk > nb ? npk : nb;
                if (pool > pool_end) pool = pool_end;
            } else {
                pool += cnt;
            }
            if (!have) {
                if (mine >= npk) break;            // queue drained: this lane is done
                have = true;
                pend = PEND_LAUNCH;
                pid = a.p0 + mine;
                ev = 0;
                nev = 0;
            }
        }

        const uint64_t am = ballot(true);
        const uint64_t pm = ballot(pend != PEND_NONE);
        if ((uint32_t)__popcll(pm) >= batch || pm == am) {
            // ================= random phase: lanes with a pending event ====
            if (pend != PEND_NONE) {
                const uint4 r = philox4x32_10(make_uint4((uint32_t)pid, (uint32_t)(pid >> 32), ev, stream),
                                              a.key0, a.key1);
                ++ev;
                const int kind = pend;
                pend = PEND_NONE;
                if (kind == PEND_LAUNCH) {
                    // sourcephCO2 with a beam profile (sourceph.f90:7-49)
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
                    // Every launch lane sets the whole walk state: a packet
                    // that misses the grid or is reflected ends here, and its
                    // state is never read (voxel 0 stands in for its loads).
                    // One straight path keeps the state in its registers.
                    v = inside ? (uint32_t)ci + (uint32_t)a.nx * (uint32_t)cj + nxy * (uint32_t)ck : 0u;
                    mu = a.mut[v];
                    const bool spec = inside && u2 < a.mtab[a.mat[v]].rsp_top;
                    c_miss += inside ? 0u : 1u;
                    c_spec += spec ? 1u : 0u;
                    have = inside && !spec;
                    const float xl = (float)ci * a.dx, yl = (float)cj * a.dy;
                    ox = fminf(fmaxf(x, xl), xl + a.dx);
                    oy = fminf(fmaxf(y, yl), yl + a.dy);
                    oz = ztop;
                    ux = 0.f; uy = 0.f; uz = -1.f;
                    iux = 0.f; iuy = 0.f; iuz = -1.f;
                    t = 0.f;
                    Tx = 3.0e38f; Ty = 3.0e38f; Tz = a.dz;
                    tau = neg_ln(u3);
                    lacc = 0.f;
                }
                if (kind == PEND_INTERACT) {
                    // interaction inside voxel v, tau / mu beyond the current point
                    const float s = tau * rcp(mu);
                    lacc += s;
                    const float tt = t + s;
                    const float xl = (float)ci * a.dx, yl = (float)cj * a.dy, zl = (float)ck * a.dz;
                    ox = fminf(fmaxf(ox + tt * ux, xl), xl + a.dx);
                    oy = fminf(fmaxf(oy + tt * uy, yl), yl + a.dy);
                    oz = fminf(fmaxf(oz + tt * uz, zl), zl + a.dz);
                    ++c_inter;
                    const MatEntry me = a.mtab[mat];
                    if (!(u01(r.x) < me.albedo)) {
                        c_box += tally_add<BOX, BT>(a, box, g, v, ci, cj, ck, lacc, pa);
                        box_carry<BT>(a, box, pa);
                        ++c_tally;
                        lacc = 0.f;
                        ++c_abs;
                        have = false;
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
                        t = 0.f;
                        Tx = face_dist(ox, xl, a.dx, ux, iux);
                        Ty = face_dist(oy, yl, a.dy, uy, iuy);
                        Tz = face_dist(oz, zl, a.dz, uz, iuz);
                        tau = neg_ln(u01(r.w));
                    }
                }
                if (kind == PEND_FRESNEL) {
                    // Fresnel test at a z face of the grid (the crossing is tallied)
                    const bool top = uz > 0.f;
                    if (u01(r.x) < fresnel(nin, nout, fabsf(uz))) {
                        // reflect: new segment from the point on the face
                        const float xl = (float)ci * a.dx, yl = (float)cj * a.dy, zl = (float)ck * a.dz;
                        ox = fminf(fmaxf(ox + t * ux, xl), xl + a.dx);
                        oy = fminf(fmaxf(oy + t * uy, yl), yl + a.dy);
                        oz = top ? zl + a.dz : zl;
                        Tx -= t;
                        Ty -= t;
                        t = 0.f;
                        uz = -uz;
                        iuz = -iuz;
                        Tz = a.dz * fabsf(iuz);
                    } else {
                        ++c_esc[top ? TAM_FACE_ZHI : TAM_FACE_ZLO];
                        have = false;
                    }
                }
            }
        }
        // ================= walk phase: face crossings, one per lane and
        // pass, until `batch` lanes wait for a random event or a packet (or
        // every lane does).  Staying in this loop keeps the walk state in
        // place between crossings.
        {
            uint64_t sm = ballot(pend != PEND_NONE || !have);
            while ((uint32_t)__popcll(sm) < batch && sm != am) {
                if (pend == PEND_NONE && have) {
                    if (++nev > kMaxEvents) {
                        // a packet that neither escapes nor is absorbed within
                        // kMaxEvents steps (e.g. total internal reflection in a
                        // non-absorbing slab) is abandoned, so every lane's loop ends
                        ++c_kill;
                        have = false;
                    } else {
                        int axis = 0;
                        float tn = Tx;
                        if (Ty < tn) { tn = Ty; axis = 1; }
                        if (Tz < tn) { tn = Tz; axis = 2; }
                        const float s = fmaxf(tn - t, 0.f);
                        if (mu * s < tau) {
                            tau -= mu * s;
                            lacc += s;
                            ++c_cross;
                            // mu_t = 0 (ablated): nothing is absorbed, and the atomic
                            // unit's throughput is what bounds this kernel
                            if (mu > 0.f) {
                                c_box += tally_add<BOX, BT>(a, box, g, v, ci, cj, ck, lacc, pa);
                                ++c_tally;
                            }
                            lacc = 0.f;
                            t = tn;
                            bool out = false;
                            int face = 0;
                            if (axis == 0) {
                                Tx += a.dx * fabsf(iux);
                                if (ux > 0.f) { ++ci; ++v; out = ci >= a.nx; face = TAM_FACE_XHI; }
                                else          { --ci; --v; out = ci < 0;     face = TAM_FACE_XLO; }
                            } else if (axis == 1) {
                                Ty += a.dy * fabsf(iuy);
                                if (uy > 0.f) { ++cj; v += a.nx; out = cj >= a.ny; face = TAM_FACE_YHI; }
                                else          { --cj; v -= a.nx; out = cj < 0;     face = TAM_FACE_YLO; }
                            } else {
                                const bool top = uz > 0.f;
                                if (top ? (ck == a.nz - 1) : (ck == 0)) {
                                    // z face of the grid: Fresnel test where the index changes
                                    nout = top ? a.n_zhi : a.n_zlo;
                                    nin = a.mtab[a.mat[v]].n;
                                    if (nout > 0.f && nout != nin) pend = PEND_FRESNEL;
                                    else { out = true; face = top ? TAM_FACE_ZHI : TAM_FACE_ZLO; }
                                } else {
                                    Tz += a.dz * fabsf(iuz);
                                    if (top) { ++ck; v += nxy; } else { --ck; v -= nxy; }
                                }
                            }
                            if (out) {
                                ++c_esc[face];
                                have = false;
                            } else if (pend == PEND_NONE) {
                                mu = a.mut[v];
                            }
                            box_carry<BT>(a, box, pa);
                        } else {
                            pend = PEND_INTERACT;
                            mat = a.mat[v];
                        }
                    }
                }
                sm = ballot(pend != PEND_NONE || !have);
            }
        }
    }

    // ---- per-wave reduction of the fate and work counters ----------------
    // (lanes leave the loop one by one as the queue drains; the reduction
    // runs once every lane of the wave has left)
    const uint32_t vals[TAM_CNT_N] = {c_abs, c_spec, c_esc[0], c_esc[1], c_esc[2], c_esc[3],
                                      c_esc[4], c_esc[5], c_miss, c_cross, c_inter, c_kill,
                                      c_tally, c_box};
    const bool leader = (threadIdx.x & 63) == 0;
#pragma unroll
    for (int i = 0; i < TAM_CNT_N; ++i) {
        const uint32_t w = wave_sum(vals[i]);
        if (leader && w) atomicAdd(&a.counters[i], (unsigned long long)w);
    }
}

__global__ __launch_bounds__(256) void mc_kernel(McArgs a)
{
    transport<false, double>(a, nullptr, BoxGeom{});
}

// The same with the block-private LDS tally box: cleared first, added to the
// device tally (one atomic per touched voxel) once every wave of the block
// has drained the queue.  NT = 256, or 1024: one block of 16 waves per CU
// holds one box of the CU's whole LDS instead of four quarter-size copies.
// BT = uint16_t (16-bit fixed point, the default box) or double.
template <int NT, class BT>
__global__ __launch_bounds__(NT) void mc_box_kernel(McArgs a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char box_lds[];
    BT* box = reinterpret_cast<BT*>(box_lds);
    // geometry: chosen by box_window_kernel (inside the grid, at most the
    // LDS this launch was given), or the launch's default
    BoxGeom g;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        g.o[q] = a.box_geom ? __builtin_amdgcn_readfirstlane(a.box_geom[q]) : a.box_o[q];
        g.n[q] = a.box_geom ? __builtin_amdgcn_readfirstlane(a.box_geom[3 + q]) : a.box_n[q];
    }
    const int nbox = g.n[0] * g.n[1] * g.n[2];
    for (int i = threadIdx.x; i < nbox; i += NT) box[i] = (BT)0;
    __syncthreads();
    transport<true, BT>(a, box, g);
    __syncthreads();
    const uint32_t nxy = (uint32_t)a.nx * (uint32_t)a.ny;
    for (int i = threadIdx.x; i < nbox; i += NT) {
        const double s = box_path(box[i], a);
        if (s != 0.0) {
            const int bx = i % g.n[0], by = (i / g.n[0]) % g.n[1], bz = i / (g.n[0] * g.n[1]);
            const uint32_t v = (uint32_t)(g.o[0] + bx) + (uint32_t)a.nx * (uint32_t)(g.o[1] + by) +
                               nxy * (uint32_t)(g.o[2] + bz);
            unsafeAtomicAdd(&a.tally[v], s);
        }
    }
}

// Placement of the LDS tally box in z.  box_plane_kernel: one block per z
// plane sums the last run's tally (path length, in proportion to the voxel
// visits, i.e. the deposits a box would take) over the box's footprint,
// skipping voxels removed since (mu_t = 0: ablated after that run);
// box_window_kernel: one block sums every window of d consecutive planes
// (plane sums staged in LDS, one window per lane) and keeps the largest,
// ties to the higher window; with no path at all the box sits at the top.
struct BoxPlace {
    int sd, d, ox, oy;                             // footprint sd x sd at (ox, oy), d planes
};

__global__ __launch_bounds__(256) void box_plane_kernel(const double* __restrict__ src, const float* __restrict__ mut,
                                                        int nx, int ny, BoxPlace bp, double* __restrict__ psum)
{
    __shared__ double red[4];
    const int k = blockIdx.x;
    double t = 0.0;
    for (int q = threadIdx.x; q < bp.sd * bp.sd; q += 256) {
        const int i = bp.ox + q % bp.sd, j = bp.oy + q / bp.sd;
        const long v = (long)i + (long)nx * (j + (long)ny * k);
        if (mut[v] > 0.f) t += src[v];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) psum[k] = red[0] + red[1] + red[2] + red[3];
}

constexpr int BOX_MAX_NZ = 4096;                  // tam_create's largest grid edge

__global__ __launch_bounds__(256) void box_window_kernel(const double* __restrict__ psum, int nz, BoxPlace bp,
                                                         int* __restrict__ out)
{
    __shared__ double sp[BOX_MAX_NZ];
    __shared__ double wb[4];
    __shared__ int wo[4];
    for (int k = threadIdx.x; k < nz; k += 256) sp[k] = psum[k];
    __syncthreads();
    double best = -1.0;                            // sums are >= 0: any window beats it
    int bo = -1;
    for (int o = threadIdx.x; o <= nz - bp.d; o += 256) {   // window [o, o + d)
        double w = 0.0;
        for (int k = 0; k < bp.d; ++k) w += sp[o + k];
        if (w > best || (w == best && o > bo)) {
            best = w;
            bo = o;
        }
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const double ow = __shfl_xor(best, s, 64);
        const int oo = __shfl_xor(bo, s, 64);
        if (ow > best || (ow == best && oo > bo)) {
            best = ow;
            bo = oo;
        }
    }
    if ((threadIdx.x & 63) == 0) {
        wb[threadIdx.x >> 6] = best;
        wo[threadIdx.x >> 6] = bo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; ++q)
            if (wb[q] > best || (wb[q] == best && wo[q] > bo)) {
                best = wb[q];
                bo = wo[q];
            }
        out[0] = bp.ox;
        out[1] = bp.oy;
        out[2] = best > 0.0 ? bo : nz - bp.d;
        out[3] = bp.sd;
        out[4] = bp.sd;
        out[5] = bp.d;
    }
}

// Many instances in one launch: a wave takes packets from one instance's
// queue until it drains, then moves on to the next instance (starting at its
// own index, so the waves spread over the instances); each instance keeps its
// own queue, tally and counters, and the instance's arguments stay uniform
// over the wave.  With wbox > 0 each wave owns an LDS slice of wbox 16-bit
// fixed-point voxels: for every instance it works on (with the box enabled,
// of either kind) it sums its
// visits to a cube under the top face on that instance's beam axis there,
// and adds the cube to the instance's tally when it moves on.  The waves of
// a block work on different instances, so the slices are per wave and only
// wave-level ordering is needed.
__global__ __launch_bounds__(256) void mc_batch_kernel(const McArgs* __restrict__ tab, int n, int wbox)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_slices[];
    // wbox is a multiple of 8, so every slice starts on a 16-byte boundary
    uint16_t* box = reinterpret_cast<uint16_t*>(lds_slices) + (size_t)(threadIdx.x >> 6) * (size_t)wbox;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    for (int k = 0; k < n; ++k) {
        const int i = (wave + k) % n;
        const McArgs a = tab[i];
        if (a.npk == 0) continue;
        BoxGeom g{};
        if (wbox > 0 && a.box_enable) {
            // a cube of at most wbox voxels inside the grid (sd <= nx, ny; d <= nz)
            int sd = (int)cbrtf((float)wbox);
            while ((long)(sd + 1) * (sd + 1) * (sd + 1) <= wbox) ++sd;
            while (sd > 1 && (long)sd * sd * sd > wbox) --sd;
            sd = min(sd, min(a.nx, a.ny));
            if (sd > 1) sd &= ~1;                          // the two voxels of a word share a row
            const int d = min(a.nz, wbox / (sd * sd));
            g.n[0] = g.n[1] = sd;
            g.n[2] = d;
            g.o[0] = max(0, min((int)lroundf(a.cx / a.dx) - sd / 2, a.nx - sd));
            g.o[1] = max(0, min((int)lroundf(a.cy / a.dy) - sd / 2, a.ny - sd));
            g.o[2] = a.nz - d;
        }
        const int nb = g.n[0] * g.n[1] * g.n[2];
        if (nb < 8) {
            transport<false, double>(a, nullptr, g);
            continue;
        }
        for (int q = lane; q < nb; q += 64) box[q] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   // the zeros land before any add
        __builtin_amdgcn_wave_barrier();
        transport<true, uint16_t>(a, box, g);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   // every lane's adds land before the flush
        __builtin_amdgcn_wave_barrier();
        const uint32_t nxy = (uint32_t)a.nx * (uint32_t)a.ny;
        for (int q = lane; q < nb; q += 64) {
            const double s = box_path(box[q], a);
            if (s != 0.0) {
                const int bx = q % g.n[0], by = (q / g.n[0]) % g.n[1], bz = q / (g.n[0] * g.n[1]);
                const uint32_t v = (uint32_t)(g.o[0] + bx) + (uint32_t)a.nx * (uint32_t)(g.o[1] + by) +
                                   nxy * (uint32_t)(g.o[2] + bz);
                unsafeAtomicAdd(&a.tally[v], s);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   // the flush reads finish before the next zeros
        __builtin_amdgcn_wave_barrier();
    }
}

// Atomic-rate probe: each lane adds 1.0 to `per_lane` xorshift-chosen slots.
__global__ __launch_bounds__(256) void atomic_rate_kernel(double* t, uint32_t slots, uint32_t per_lane)
{
    uint32_t x = (blockIdx.x * 256u + threadIdx.x) * 2654435761u + 0x9E3779B9u;
    for (uint32_t i = 0; i < per_lane; ++i) {
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        unsafeAtomicAdd(&t[x % slots], 1.0);
    }
}

}  // namespace

hipError_t launch_atomic_rate(double* t, long slots, long count, hipStream_t s)
{
    // 2048 blocks of 256 lanes; per_lane adds each
    const long lanes = 2048L * 256L;
    long per = count / lanes;
    if (per < 1) per = 1;
    if (slots < 1 || slots > 0xffffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(atomic_rate_kernel, dim3(2048), dim3(256), 0, s, t, (uint32_t)slots, (uint32_t)per);
    return hipGetLastError();
}

hipError_t launch_mc_batch(const McArgs* dtab, int n, long total_npk, int blocks_per_cu, hipStream_t s)
{
    if (n <= 0 || total_npk <= 0) return hipSuccess;
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || ncu <= 0) ncu = 256;
    const int bpc = blocks_per_cu > 0 ? blocks_per_cu : 8;
    long blocks = (long)ncu * bpc;
    const long need = (total_npk + 255) / 256;
    if (need < blocks) blocks = need;
    // per-wave LDS tally boxes: the block's share of the CU's LDS (1 KiB kept
    // for the kernel's own use), split over its 4 waves
    int lds_cu = 0, lds_blk = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        lds_cu <= 0)
        lds_cu = 64 * 1024;
    if (hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        lds_blk <= 0)
        lds_blk = 64 * 1024;
    const long per_block = std::min((long)lds_blk - 1024L, (long)lds_cu / bpc - 1024L);
    // 16-bit slots, a multiple of 8 per wave
    const int wbox = per_block > 0 ? (int)(per_block / 4 / 2) & ~7 : 0;
    const size_t lds = sizeof(uint16_t) * 4 * (size_t)wbox;
    hipLaunchKernelGGL(mc_batch_kernel, dim3((unsigned)blocks), dim3(256), lds, s, dtab, n, wbox);
    return hipGetLastError();
}

void mc_box_units(McArgs& a)
{
    a.box_scale = (float)std::ldexp(1.0 / ((double)a.dx + a.dy + a.dz), kBoxUnitBits);
    a.box_quantum = 1.0 / (double)a.box_scale;
}

hipError_t mc_prepare(int dev)
{
    int lds_blk = 0;
    hipError_t e = hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess || lds_blk <= 64 * 1024) return e;
    // allow the box kernels the device's whole per-block LDS (outside any graph capture)
    const void* box_kernels[] = {(const void*)mc_box_kernel<256, uint16_t>, (const void*)mc_box_kernel<1024, uint16_t>,
                                 (const void*)mc_box_kernel<256, double>, (const void*)mc_box_kernel<1024, double>};
    for (const void* k : box_kernels) {
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds_blk);
        if (e != hipSuccess) return e;
    }
    return hipFuncSetAttribute((const void*)mc_batch_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds_blk);
}

hipError_t launch_mc(const McArgs& a, hipStream_t s)
{
    const long nvox = (long)a.nx * a.ny * a.nz;
    if (a.npk == 0) return a.clear_tally ? launch_fill_f64(a.tally, nvox, 0.0, s) : hipSuccess;
    int dev = 0, ncu = 0, lds_cu = 0, lds_blk = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || ncu <= 0) ncu = 256;
    // LDS per CU and per block (160 KiB each on MI355X)
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        lds_cu <= 0)
        lds_cu = 64 * 1024;
    if (hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        lds_blk <= 0)
        lds_blk = 64 * 1024;
    // blocks_per_cu 256-lane blocks per CU (autotuned); every lane takes
    // packets from the queue until it is empty, so more blocks than packets
    // only idle
    const int bpc = a.blocks_per_cu > 0 ? a.blocks_per_cu : 8;
    // with the box on, a multiple of 4 runs as bpc / 4 blocks of 1024 lanes:
    // the same waves per CU, one box per 16 waves
    const bool wide = a.box_enable && bpc % 4 == 0;
    const int lanes = wide ? 1024 : 256, bpc_l = wide ? bpc / 4 : bpc;
    long blocks = (long)ncu * bpc_l;
    const long need = (long)((a.npk + lanes - 1) / lanes);
    if (need < blocks) blocks = need;
    // LDS tally box: the resident blocks of a CU share its LDS (160 KiB on
    // MI355X, also the most one block may take), 1 KiB per block kept for
    // the kernel's own use: with 16-bit sums 42^2 x 46 voxels at 1 block per CU
    // (1 or 4 blocks_per_cu), 34^2 x 34 at 2 (2 or 8), 28^2 x 34 at 3; with
    // FP64 27^3, 21^3, 18^2 x 20.  The cube sits on the beam axis (its x edge
    // even with 16-bit sums, so the two voxels of a word share a row).  Its z range is where
    // the last run's path was largest, or the top face (where most of the
    // path of a surface beam lies) before any run.  On the c1 skin cycles a
    // cube caught more deposits past the first cycle than flatter (d ~ s/2,
    // s/8) or taller (d ~ 2s) boxes of the same volume, and than a per-launch
    // choice among depths of 2-32 planes made from the last run's path.
    // The box holds 16-bit fixed-point sums (box_enable 1: four times the
    // voxels of FP64) or FP64 ones (box_enable 2).  Fixed-point unit: 2^-13 of
    // dx + dy + dz, so a deposit is rounded by at most ~1e-4 of a voxel-sized
    // path, and a slot wraps (one carry) every ~40 typical deposits.
    const bool fx = a.box_enable == 1;
    const long esz = fx ? 2 : 8;
    McArgs b = a;
    mc_box_units(b);
    b.box_n[0] = b.box_n[1] = b.box_n[2] = 0;
    b.box_geom = nullptr;
    const long vox = std::min((long)lds_blk - 1024L, (long)lds_cu / bpc_l - 1024L) / esz;
    BoxPlace bp{};
    bp.sd = std::min((int)std::cbrt((double)vox), std::min(a.nx, a.ny));
    if (fx && bp.sd > 1) bp.sd &= ~1;
    bp.d = bp.sd > 0 ? (int)std::min<long>(a.nz, vox / ((long)bp.sd * bp.sd)) : 0;
    // A launch with fewer packets than a quarter of the lanes (the reference's
    // own run: 125 000 packets per cycle of a few deposits each) goes without
    // the box: clearing and flushing the box would cost more than it saves.
    const bool few = (double)a.npk < 0.25 * (double)ncu * bpc_l * lanes;
    if (a.box_enable && !few && bp.sd >= 4 && bp.d >= 1) {
        bp.ox = std::max(0, std::min((int)std::lround(a.cx / a.dx) - bp.sd / 2, a.nx - bp.sd));
        bp.oy = std::max(0, std::min((int)std::lround(a.cy / a.dy) - bp.sd / 2, a.ny - bp.sd));
        b.box_o[0] = bp.ox;
        b.box_o[1] = bp.oy;
        b.box_o[2] = a.nz - bp.d;
        b.box_n[0] = b.box_n[1] = bp.sd;
        b.box_n[2] = bp.d;
        const size_t lds = (size_t)esz * (size_t)bp.sd * (size_t)bp.sd * (size_t)bp.d;
        if (a.box_psum && a.box_geom) {
            hipLaunchKernelGGL(box_plane_kernel, dim3((unsigned)a.nz), dim3(256), 0, s, a.tally, a.mut, a.nx, a.ny,
                               bp, a.box_psum);
            hipLaunchKernelGGL(box_window_kernel, dim3(1), dim3(256), 0, s, a.box_psum, a.nz, bp, a.box_geom);
            b.box_geom = a.box_geom;
        }
        if (a.clear_tally) {
            const hipError_t ef = launch_fill_f64(a.tally, nvox, 0.0, s);
            if (ef != hipSuccess) return ef;
        }
        if (fx) {
            if (wide) hipLaunchKernelGGL((mc_box_kernel<1024, uint16_t>), dim3((unsigned)blocks), dim3(1024), lds, s, b);
            else hipLaunchKernelGGL((mc_box_kernel<256, uint16_t>), dim3((unsigned)blocks), dim3(256), lds, s, b);
        } else {
            if (wide) hipLaunchKernelGGL((mc_box_kernel<1024, double>), dim3((unsigned)blocks), dim3(1024), lds, s, b);
            else hipLaunchKernelGGL((mc_box_kernel<256, double>), dim3((unsigned)blocks), dim3(256), lds, s, b);
        }
    } else {
        if (a.clear_tally) {
            const hipError_t ef = launch_fill_f64(a.tally, nvox, 0.0, s);
            if (ef != hipSuccess) return ef;
        }
        hipLaunchKernelGGL(mc_kernel, dim3((unsigned)blocks), dim3(256), 0, s, b);
    }
    return hipGetLastError();
}
