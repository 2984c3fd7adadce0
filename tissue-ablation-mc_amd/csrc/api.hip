// api.hip -- the C-ABI of libtam.so (declared in include/tam.h).
//
// A tam_ctx owns every device array of one tissue grid, one HIP stream and
// timing events.  Host buffers cross the boundary only in tam_set_*,
// tam_upload and tam_download; everything between (transport, normalisation,
// heat steps, damage, rebuild) stays resident in HBM.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "tam_internal.h"

// Host-side coefficient tables restate the device's FP64 formulas; no
// contraction so both round identically.
#pragma clang fp contract(off)

namespace {

thread_local std::string g_create_err;

int fail(tam_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

}  // namespace

int tam_fail_noctx(int code, const std::string& msg) { return fail(nullptr, code, msg); }

// ---- fixed device buffer of tam_tally_blocks_noalloc -----------------------
// Every GPU the library runs on holds its own copy of `scratch` until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: the device address comes from hipGetSymbolAddress, taken
// once per device with that device current (at tam_create, or at the entry
// point's first call on a device) and kept in scratch_dev[].
static __device__ __align__(256) double scratch[TAM_BLOCKSUM_SCRATCH_MAX];

namespace {

constexpr int kMaxDevices = 64;
// tam_tune_mc keeps the CU-wide LDS tally box when it times within this
// fraction of the fastest candidate
constexpr double kWideBoxMargin = 0.10;
// ... and a random-phase batch of 32 when it times within this fraction of the fastest batch
constexpr double kBatchMargin = 0.03;
std::mutex scratch_mu;
double* scratch_dev[kMaxDevices] = {};

// the calling thread's current device must be `dev`
hipError_t scratch_ptr(int dev, double** out)
{
    *out = nullptr;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_dev[dev] = (double*)p;
    }
    *out = scratch_dev[dev];
    return hipSuccess;
}

}  // namespace

namespace {

int hip_fail(tam_ctx* c, const char* what, hipError_t e)
{
    return fail(c, TAM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, call)                                        \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return hip_fail((ctx), #call, e_); \
    } while (0)

template <class T>
hipError_t dalloc(T** p, long n)
{
    return hipMalloc((void**)p, sizeof(T) * (size_t)(n > 0 ? n : 1));
}

struct Dev {   // make the context's device current for the scope of a call
    int prev = -1;
    explicit Dev(int d)
    {
        hipGetDevice(&prev);
        if (prev != d) hipSetDevice(d);
    }
    ~Dev()
    {
        int cur = -1;
        hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) hipSetDevice(prev);
    }
};

McArgs mc_args(tam_ctx* c, int64_t npackets, uint64_t seed, uint32_t stream)
{
    McArgs a{};
    a.nx = c->p.nx; a.ny = c->p.ny; a.nz = c->p.nz;
    a.dx = (float)(2.0 * c->p.xmax / c->p.nx);
    a.dy = (float)(2.0 * c->p.ymax / c->p.ny);
    a.dz = (float)(2.0 * c->p.zmax / c->p.nz);
    a.cx = (float)c->p.xmax;
    a.cy = (float)c->p.ymax;
    a.radius = (float)c->p.beam_radius;
    a.beam = c->p.beam;
    a.n_zlo = (float)c->p.n_ext_zlo;
    a.n_zhi = (float)c->p.n_ext_zhi;
    a.mut = c->mut; a.mat = c->mat; a.mtab = c->mtab;
    a.tally = c->tally; a.counters = c->counters;
    a.p0 = 0; a.npk = (uint64_t)npackets;
    a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32); a.stream = stream;
    a.blocks_per_cu = c->mc_bpc;
    a.rng_batch = c->mc_batch;
    a.box_enable = c->mc_box;
    a.box_psum = c->box_psum;              // the box follows the last run's path
    a.box_geom = c->box_geom;
    mc_box_units(a);
    return a;
}

HeatArgs heat_args(tam_ctx* c)
{
    HeatArgs a{};
    a.nx = c->p.nx; a.ny = c->p.ny; a.nz = c->p.nz;
    a.px = c->px; a.pxy = c->pxy;
    a.gx = c->gx; a.gy = c->gy; a.gz = c->gz;
    a.src = c->src; a.robin = c->robin;
    a.Q = c->Q; a.jmean = c->jmean; a.heatcap = c->heatcap;
    a.delt = c->p.delt; a.volume_voxel = c->p.volume_voxel; a.mass_voxel = c->p.mass_voxel;
    a.qvapor = c->p.qvapor;
    a.dT_boil = (float)(373.0 - c->p.t_ref);
    a.t_ref = c->p.t_ref;
    a.omega = c->omega; a.thres = c->thres; a.tmask = c->tmask; a.rhokap = c->rhokap;
    a.arr_A = c->p.arr_A; a.arr_dE = c->p.arr_dE; a.arr_R = c->p.arr_R;
    a.arr_Tlo = c->p.arr_Tlo; a.arr_Thi = c->p.arr_Thi;
    a.thr0 = c->p.arr_thr[0]; a.thr1 = c->p.arr_thr[1]; a.thr2 = c->p.arr_thr[2];
    a.energy = c->energy; a.flags = c->flags;
    a.cls = c->cls; a.glut = c->glut; a.robin_lut = c->robin_lut;
    a.eblk = c->eblk;
    a.lnA = std::log(c->p.arr_A);
    a.thr_min = std::fmin(c->p.arr_thr[0], std::fmin(c->p.arr_thr[1], c->p.arr_thr[2]));
    a.dT_win_pad = 2.0f;
    // FP32 pre-tests widened by a margin well above FP32 rounding; the exact
    // comparisons are made in FP64 behind them
    auto margin = [](double x) { return 1e-3 * std::fabs(x) + 1e-3; };
    a.dT_neg = (float)(-c->p.t_ref);
    const double db = 373.0 - c->p.t_ref, dl = c->p.arr_Tlo - c->p.t_ref, dh = c->p.arr_Thi - c->p.t_ref;
    a.dT_boil_test = (float)(db - margin(db));
    a.dT_lo_test = (float)(dl - margin(dl));
    a.dT_hi_test = (float)(dh + margin(dh));
    return a;
}

// The step loops pair steps into two-step launches when the shape is a
// two-step one and the step allows it (FP32 class mode, no boiling: the
// boiling clamp changes Q, which a neighbour block's redundant first step
// would read); otherwise every step runs alone on the shape's one-step twin.
bool two_step(const tam_ctx* c)
{
    return !c->fp64 && c->lut && !c->p.boil && heat_tile_steps(c->tile) == 2;
}

// Thermal classes for the FP32 stencil's class mode: every distinct
// (kappa, density, heatcap) triple over the halo grid gets an id (at most
// TAM_NCLASS-1); removed interior voxels get TAM_CLASS_REMOVED.  The face
// table holds the same FP64 expression as coeff_kernel, rounded to FP32.
// Class mode needs the FP32 path, the table thermal model and inert removed
// voxels (air properties depend on T, so they are not a fixed set).
int classify(tam_ctx* c, const double* K, const double* D, const double* Hc, const double* rk)
{
    c->lut = false;
    if (c->fp64 || c->p.thermal_mode != TAM_THERMAL_TABLE || !c->p.removed_inert) return TAM_OK;
    struct Tri { double k, d, h; };
    std::vector<Tri> u;
    const long H = c->nhalo, px = c->px, pxy = c->pxy;
    std::vector<uint8_t> cls((size_t)H);
    for (long h = 0; h < H; ++h) {
        const int i = (int)(h % px), j = (int)((h / px) % (c->p.ny + 2)), k = (int)(h / pxy);
        const bool interior = i >= 1 && i <= c->p.nx && j >= 1 && j <= c->p.ny && k >= 1 && k <= c->p.nz;
        if (interior && rk[h] == 0.0) {
            cls[(size_t)h] = TAM_CLASS_REMOVED;
            continue;
        }
        size_t m = 0;
        while (m < u.size() && !(u[m].k == K[h] && u[m].d == D[h] && u[m].h == Hc[h])) ++m;
        if (m == u.size()) {
            if (u.size() == TAM_NCLASS - 1) return TAM_OK;      // too many: array mode
            u.push_back(Tri{K[h], D[h], Hc[h]});
        }
        cls[(size_t)h] = (uint8_t)m;
    }
    std::vector<float> lut(3 * 256 + TAM_NCLASS, 0.f);       // + heatcap*mass per class
    std::vector<float2> rob(TAM_NCLASS, make_float2(0.f, 0.f));
    const double ih2[3] = {1.0 / (c->p.dx_m * c->p.dx_m), 1.0 / (c->p.dy_m * c->p.dy_m),
                           1.0 / (c->p.dz_m * c->p.dz_m)};
    for (int ax = 0; ax < 3; ++ax)
        for (size_t m = 0; m < u.size(); ++m)
            for (size_t n = 0; n < u.size(); ++n) {
                const double kb = .5 * (u[m].k + u[n].k);
                const double db = .5 * (u[m].d + u[n].d);
                const double hb = .5 * (u[m].h + u[n].h);
                lut[ax * 256 + m * 16 + n] = (float)(c->p.delt * (c->p.dscale * (kb / (db * hb)) * ih2[ax]));
            }
    for (size_t m = 0; m < u.size(); ++m) lut[3 * 256 + m] = (float)(u[m].h * c->p.mass_voxel);
    for (size_t m = 0; m < u.size(); ++m) {
        const double kk = u[m].k / c->p.dz_m;
        const double den = kk + 0.5 * c->p.h_top;
        const double beta = (kk - 0.5 * c->p.h_top) / den;
        const double gam = c->p.h_top * c->p.t_amb / den;
        rob[m] = make_float2((float)beta, (float)(beta * c->p.t_ref + gam - c->p.t_ref));
    }
    HIPCHK(c, hipMemcpyAsync(c->cls, cls.data(), (size_t)H, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_class_flags(c->cls, c->p.nx, c->p.ny, c->p.nz, c->px, c->pxy, c->stream));
    HIPCHK(c, launch_class_due(c->cls, c->tmask, c->p.nx, c->p.ny, c->p.nz, c->px, c->pxy, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->glut, lut.data(), sizeof(float) * lut.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->robin_lut, rob.data(), sizeof(float2) * rob.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lut = true;
    return TAM_OK;
}

// classify() from the device copies of the properties (after optics change)
int reclassify_from_device(tam_ctx* c)
{
    const size_t H = (size_t)c->nhalo;
    std::vector<double> K(H), D(H), Hc(H), rk(H);
    HIPCHK(c, hipMemcpyAsync(K.data(), c->kappa, 8 * H, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.data(), c->density, 8 * H, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Hc.data(), c->heatcap, 8 * H, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(rk.data(), c->rhokap, 8 * H, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return classify(c, K.data(), D.data(), Hc.data(), rk.data());
}

RebuildArgs rebuild_args(tam_ctx* c)
{
    RebuildArgs a{};
    a.nx = c->p.nx; a.ny = c->p.ny; a.nz = c->p.nz;
    a.px = c->px; a.pxy = c->pxy;
    a.mode = c->p.thermal_mode;
    a.ablate_K = c->p.ablate_K; a.delt = c->p.delt;
    a.mu_water = c->p.mu_water; a.mu_protein = c->p.mu_protein;
    a.qvapor = c->p.qvapor; a.t_ref = c->p.t_ref;
    a.dT = c->cur;
    a.rhokap = c->rhokap; a.kappa = c->kappa; a.density = c->density; a.heatcap = c->heatcap;
    a.coeff = c->coeff; a.W = c->W; a.Q = c->Q; a.mut = c->mut; a.mat = c->mat;
    a.mask = c->mask; a.air_id = c->air_id;
    a.T64 = c->fp64 ? c->cur64 : nullptr;
    a.cls = c->lut ? c->cls : nullptr;
    return a;
}

Heat64Args heat64_args(tam_ctx* c)
{
    Heat64Args a{};
    a.nx = c->p.nx; a.ny = c->p.ny; a.nz = c->p.nz;
    a.px = c->px; a.pxy = c->pxy;
    a.ax = c->ax; a.ay = c->ay; a.az = c->az;
    a.kappa = c->kappa; a.heatcap = c->heatcap; a.coeff = c->coeff; a.jmean = c->jmean; a.Q = c->Q;
    a.delt = c->p.delt; a.volume_voxel = c->p.volume_voxel; a.mass_voxel = c->p.mass_voxel;
    a.qvapor = c->p.qvapor; a.h_top = c->p.h_top; a.t_amb = c->p.t_amb; a.dz = c->p.dz_m;
    a.omega = c->omega; a.thres = c->thres; a.tmask = c->tmask;
    a.arr_A = c->p.arr_A; a.arr_dE = c->p.arr_dE; a.arr_R = c->p.arr_R;
    a.arr_Tlo = c->p.arr_Tlo; a.arr_Thi = c->p.arr_Thi;
    a.thr0 = c->p.arr_thr[0]; a.thr1 = c->p.arr_thr[1]; a.thr2 = c->p.arr_thr[2];
    a.energy = c->energy; a.flags = c->flags;
    return a;
}

hipError_t build_coeffs(tam_ctx* c)
{
    if (!c->fp64) {
        CoeffArgs a{};
        a.nx = c->p.nx; a.ny = c->p.ny; a.nz = c->p.nz;
        a.px = c->px; a.pxy = c->pxy;
        a.delt = c->p.delt; a.dscale = c->p.dscale;
        a.idx2 = 1.0 / (c->p.dx_m * c->p.dx_m);
        a.idy2 = 1.0 / (c->p.dy_m * c->p.dy_m);
        a.idz2 = 1.0 / (c->p.dz_m * c->p.dz_m);
        a.kappa = c->kappa; a.density = c->density; a.heatcap = c->heatcap;
        a.gx = c->gx; a.gy = c->gy; a.gz = c->gz;
        a.robin = c->p.robin_top; a.h_top = c->p.h_top; a.t_amb = c->p.t_amb; a.t_ref = c->p.t_ref;
        a.dz = c->p.dz_m; a.robin_coef = c->robin;
        a.rhokap = c->rhokap; a.removed_inert = c->p.removed_inert;
        hipError_t e = launch_coeffs(a, c->stream);
        // removals change classes: refresh the homogeneous-neighbourhood bits
        if (e == hipSuccess && c->lut) e = launch_class_flags(c->cls, c->p.nx, c->p.ny, c->p.nz, c->px, c->pxy,
                                                              c->stream);
        return e;
    }
    Coeff64Args a{};
    a.nx = c->p.nx; a.ny = c->p.ny; a.nz = c->p.nz;
    a.px = c->px; a.pxy = c->pxy;
    a.dscale = c->p.dscale;
    a.idx2 = 1.0 / (c->p.dx_m * c->p.dx_m);
    a.idy2 = 1.0 / (c->p.dy_m * c->p.dy_m);
    a.idz2 = 1.0 / (c->p.dz_m * c->p.dz_m);
    a.kappa = c->kappa; a.density = c->density; a.heatcap = c->heatcap;
    a.ax = c->ax; a.ay = c->ay; a.az = c->az;
    a.rhokap = c->rhokap; a.removed_inert = c->p.removed_inert;
    return launch_coeffs64(a, c->stream);
}

float elapsed(tam_ctx* c)
{
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    return ms;
}

int check_flags(tam_ctx* c)
{
    int f = 0;
    HIPCHK(c, hipMemcpyAsync(&f, c->flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (f) return fail(c, TAM_ERR_NEGATIVE_TEMP, "negative temperature in the heat step");
    return TAM_OK;
}

// src = coeff * jmean (after a property change the per-step source must use
// the new delt/(rho c), as heat_sim_3D does with coeff(i,j,k)*jtmp)
__global__ void src_kernel(int nx, int ny, int nz, long px, long pxy, const double* coeff,
                           const double* jmean, float* src)
{
    const long nv = (long)nx * ny * nz;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    const int i = (int)(v % nx) + 1;
    const long r = v / nx;
    const int j = (int)(r % ny) + 1;
    const int k = (int)(r / ny) + 1;
    const long cc = i + px * j + pxy * k;
    src[cc] = (float)(coeff[cc] * jmean[v]);
}

int refresh_src(tam_ctx* c)
{
    const long nv = c->nvox;
    hipLaunchKernelGGL(src_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, c->stream, c->p.nx, c->p.ny,
                       c->p.nz, c->px, c->pxy, c->coeff, c->jmean, c->src);
    HIPCHK(c, hipGetLastError());
    return TAM_OK;
}

// ---- captured coupling cycle -------------------------------------------

// Drop the captured graphs (state or launch shape changed; the next
// tam_cycle captures again).
void cycle_reset(tam_ctx* c)
{
    if ((c->cyc_exec[0] || c->cyc_exec[1]) && c->stream) hipStreamSynchronize(c->stream);
    for (int p = 0; p < 2; ++p)
        if (c->cyc_exec[p]) {
            hipGraphExecDestroy(c->cyc_exec[p]);
            c->cyc_exec[p] = nullptr;
        }
    c->cyc_loops = -1;
}

// parameter block: header, float lon[loops] (padded to 8 bytes), double t_after[loops]
size_t cycle_lon_off() { return sizeof(CycleHdr); }
size_t cycle_ta_off(int loops) { return sizeof(CycleHdr) + (((size_t)loops * 4 + 7) & ~(size_t)7); }
size_t cycle_bytes(int loops) { return cycle_ta_off(loops) + (size_t)loops * 8; }

// One coupling cycle on c->stream, reading the per-cycle values from the
// device parameter block.  Swaps the host's ping-pong pointers as the steps
// are enqueued (the caller restores them after a capture) and counts the
// swaps in *swaps: one per launch, so a two-step launch swaps once.
int enqueue_cycle(tam_ctx* c, int* swaps)
{
    *swaps = 0;
    const CycleHdr* hdr = (const CycleHdr*)c->cyc_dev;
    const float* lon = (const float*)(c->cyc_dev + cycle_lon_off());
    const double* ta = (const double*)(c->cyc_dev + cycle_ta_off(c->cyc_loops));
    hipStream_t s = c->stream;
    // transport (mcpolar.f90:151-170) and normalisation (:173-174)
    HIPCHK(c, launch_fill_u64(c->counters, TAM_CNT_SLOTS, s));
    McArgs m = mc_args(c, c->cyc_npk, c->cyc_seed, 0);
    m.clear_tally = 1;                     // after the LDS-box placement has read the last cycle's tally
    m.dyn_npk = &hdr->npk;
    m.dyn_stream = &hdr->stream;
    HIPCHK(c, launch_mc(m, s));
    HIPCHK(c, launch_acc_u64(c->cnt_acc, c->counters, TAM_CNT_N, s));
    HIPCHK(c, launch_normalise(c->p.nx, c->p.ny, c->p.nz, c->px, c->pxy, c->tally, c->rhokap, c->mat,
                               c->mua_frac, hdr, 0.0, c->coeff, c->jmean, c->src, s));
    // heat_sim_3D (3dFD.f90:21-230)
    const int boil = c->p.boil != 0, robin = c->p.robin_top != 0, dmg = c->p.damage_every_step != 0;
    HeatArgs a = heat_args(c);
    Heat64Args a64 = heat64_args(c);
    const bool two = two_step(c);
    for (int k = 0; k < c->cyc_loops; ++k) {
        if (c->fp64) {
            a64.dyn_lon = lon + k;
            a64.dyn_tafter = ta + k;
            HIPCHK(c, launch_heat64_step(a64, c->cur64, c->nxt64, 0.0, 0.0, boil, robin, dmg, 1, s));
            std::swap(c->cur64, c->nxt64);
            ++*swaps;
        } else {
            a.dyn_lon = lon + k;
            a.dyn_tafter = ta + k;
            if (two && k + 1 < c->cyc_loops) {      // steps k and k+1 in one launch
                HIPCHK(c, launch_heat_step2(a, c->cur, c->nxt, 0.f, 0.f, 0.0, 0.0, robin, dmg, 1, c->tile, c->kz,
                                            s));
                ++k;
            } else {
                HIPCHK(c, launch_heat_step(a, c->cur, c->nxt, 0.f, 0.0, boil, robin, dmg, 1, c->lut ? 1 : 0,
                                           c->tile, c->kz, s));
            }
            std::swap(c->cur, c->nxt);
            ++*swaps;
        }
    }
    // Arrhenius at the reference cadence (mcpolar.f90:181)
    if (c->cyc_arr) {
        HeatArgs aa = heat_args(c);
        aa.dyn_arr = hdr->arr;
        HIPCHK(c, launch_arrhenius(aa, c->cur, c->fp64 ? c->cur64 : nullptr, 0.0, 0.0, s));
    }
    // setupThermalCoeff (mcpolar.f90:183, 3dFD.f90:312-361)
    const RebuildArgs r = rebuild_args(c);
    if (c->p.thermal_mode == TAM_THERMAL_WATER) HIPCHK(c, launch_water_update(r, s));
    HIPCHK(c, launch_removal_mask(r, s));
    HIPCHK(c, launch_apply_props(r, s));
    HIPCHK(c, build_coeffs(c));
    return refresh_src(c);
}

// Capture enqueue_cycle from the current ping-pong parity into cyc_exec[par].
int capture_cycle(tam_ctx* c, int par)
{
    float* cur = c->cur; float* nxt = c->nxt;
    double* cur64 = c->cur64; double* nxt64 = c->nxt64;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    int swaps = 0;
    const int rc = enqueue_cycle(c, &swaps);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->cur = cur; c->nxt = nxt; c->cur64 = cur64; c->nxt64 = nxt64;
    if (rc != TAM_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, "hipStreamEndCapture", e);
    const hipError_t ei = hipGraphInstantiate(&c->cyc_exec[par], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->cyc_exec[par] = nullptr;
        return hip_fail(c, "hipGraphInstantiate", ei);
    }
    c->cyc_swap[par] = swaps & 1;
    return TAM_OK;
}

}  // namespace

extern "C" {

int tam_abi_version(void) { return TAM_ABI_VERSION; }

int64_t tam_struct_size(int32_t which)
{
    switch (which) {
    case 0: return (int64_t)sizeof(tam_params);
    case 1: return (int64_t)sizeof(tam_material);
    case 2: return (int64_t)sizeof(tam_mc_stats);
    case 3: return (int64_t)sizeof(tam_pulse);
    case 4: return (int64_t)sizeof(tam_timing);
    }
    return -1;
}

void tam_pulse_advance(tam_pulse* p, double delt)
{
    if (!p) return;
    if (p->use_pulse) {                                   // 3dFD.f90:199-210
        if (p->pulseCount >= p->realPulseLength && p->laser_flag) {
            p->laser_flag = 0;
            p->laserOn = 0.;
            p->pulseCount = 0.;
            p->pulsesDone += 1;
            p->repetitionCount = 0.;
        } else if (p->repetitionCount >= p->repetitionRate_1 && !p->laser_flag &&
                   p->pulsesDone < p->pulsesToDo) {
            p->laser_flag = 1;
            p->laserOn = 1.;
            p->pulseCount = 0.;
            p->repetitionCount = 0.;
        }
    }
    p->pulseCount = p->pulseCount + delt;                 // 3dFD.f90:212-214
    p->repetitionCount = p->repetitionCount + delt;
    p->time = p->time + delt;
}

const char* tam_last_error(const tam_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int tam_create(const tam_params* p, tam_ctx** out)
{
    if (!p || !out) return fail(nullptr, TAM_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1 || p->nx > 4096 || p->ny > 4096 || p->nz > 4096)
        return fail(nullptr, TAM_ERR_ARG, "grid dimensions must be in [1, 4096]");
    if ((double)p->nx * p->ny * p->nz >= 4294967296.0)
        return fail(nullptr, TAM_ERR_ARG, "grid must hold fewer than 2^32 voxels (32-bit voxel index)");
    if (!(p->xmax > 0 && p->ymax > 0 && p->zmax > 0))
        return fail(nullptr, TAM_ERR_ARG, "grid extents must be positive");
    if (!(p->dx_m > 0 && p->dy_m > 0 && p->dz_m > 0 && p->delt > 0))
        return fail(nullptr, TAM_ERR_ARG, "FD spacing and delt must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, TAM_ERR_NODEVICE, "no HIP device");
    if (p->device < 0 || p->device >= ndev) return fail(nullptr, TAM_ERR_ARG, "device ordinal out of range");
    tam_ctx* c = new (std::nothrow) tam_ctx();
    if (!c) return fail(nullptr, TAM_ERR_NOMEM, "host allocation failed");
    c->p = *p;
    c->dev = p->device;
    Dev d(c->dev);
    const long nx = p->nx, ny = p->ny, nz = p->nz;
    c->nvox = nx * ny * nz;
    c->px = nx + 2;
    c->pxy = (nx + 2) * (ny + 2);
    c->nhalo = c->pxy * (nz + 2);
    c->air_id = 0;
    c->fp64 = p->heat_fp64 != 0;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    const long H = c->nhalo, V = c->nvox;
    if (e == hipSuccess) e = dalloc(&c->dT0, H);
    if (e == hipSuccess) e = dalloc(&c->dT1, H);
    if (e == hipSuccess) e = dalloc(&c->gx, H);
    if (e == hipSuccess) e = dalloc(&c->gy, H);
    if (e == hipSuccess) e = dalloc(&c->gz, H);
    if (e == hipSuccess) e = dalloc(&c->src, H);
    if (e == hipSuccess) e = dalloc(&c->robin, nx * ny);
    if (e == hipSuccess) e = dalloc(&c->kappa, H);
    if (e == hipSuccess) e = dalloc(&c->density, H);
    if (e == hipSuccess) e = dalloc(&c->heatcap, H);
    if (e == hipSuccess) e = dalloc(&c->coeff, H);
    if (e == hipSuccess) e = dalloc(&c->rhokap, H);
    if (e == hipSuccess) e = dalloc(&c->mut, V);
    if (e == hipSuccess) e = dalloc(&c->mat, V);
    if (e == hipSuccess) e = dalloc(&c->mask, V);
    if (e == hipSuccess) e = dalloc(&c->cls, H);
    if (e == hipSuccess) e = dalloc(&c->glut, 3 * 256 + TAM_NCLASS);
    if (e == hipSuccess) e = dalloc(&c->robin_lut, TAM_NCLASS);
    c->tile = 1;
    c->kz = 16;
    c->mc_bpc = 8;
    c->mc_batch = 32;
    c->mc_box = 1;
    c->cyc_loops = -1;
    c->neblk = heat_max_blocks(nx, ny, nz);      // one energy slot per block of any launch shape
    if (e == hipSuccess) e = dalloc(&c->eblk, c->neblk);
    if (e == hipSuccess) e = dalloc(&c->mtab, TAM_MAX_MATERIALS + 1);
    if (e == hipSuccess) e = dalloc(&c->mua_frac, TAM_MAX_MATERIALS + 1);
    if (e == hipSuccess) e = dalloc(&c->tally, V);
    if (e == hipSuccess) e = dalloc(&c->jmean, V);
    if (e == hipSuccess) e = dalloc(&c->Q, V);
    if (e == hipSuccess) e = dalloc(&c->W, V);
    if (e == hipSuccess) e = dalloc(&c->omega, V);
    if (e == hipSuccess) e = dalloc(&c->thres, 3 * V);
    if (e == hipSuccess) e = dalloc(&c->tmask, V);
    if (e == hipSuccess) e = dalloc(&c->box_psum, nz);
    if (e == hipSuccess) e = dalloc(&c->box_geom, 6);
    if (e == hipSuccess) e = hipMemsetAsync(c->box_geom, 0, 6 * sizeof(int), c->stream);
    if (e == hipSuccess) e = dalloc(&c->energy, 4);
    if (e == hipSuccess) e = dalloc(&c->flags, 4);
    if (e == hipSuccess) e = dalloc(&c->counters, TAM_CNT_SLOTS);
    if (c->fp64) {
        if (e == hipSuccess) e = dalloc(&c->T64a, H);
        if (e == hipSuccess) e = dalloc(&c->T64b, H);
        if (e == hipSuccess) e = dalloc(&c->ax, H);
        if (e == hipSuccess) e = dalloc(&c->ay, H);
        if (e == hipSuccess) e = dalloc(&c->az, H);
    }
    if (e != hipSuccess) {
        std::string m = std::string("device allocation: ") + hipGetErrorString(e);
        tam_destroy(c);
        return fail(nullptr, e == hipErrorOutOfMemory ? TAM_ERR_NOMEM : TAM_ERR_HIP, m);
    }
    hipStream_t s = c->stream;
    launch_fill_f32(c->dT0, H, 0.f, s);
    launch_fill_f32(c->dT1, H, 0.f, s);
    launch_fill_f32(c->gx, H, 0.f, s);
    launch_fill_f32(c->gy, H, 0.f, s);
    launch_fill_f32(c->gz, H, 0.f, s);
    launch_fill_f32(c->src, H, 0.f, s);
    launch_fill_f32((float*)c->robin, 2 * nx * ny, 0.f, s);
    launch_fill_f64(c->rhokap, H, 0.0, s);
    launch_fill_f64(c->tally, V, 0.0, s);
    launch_fill_f64(c->jmean, V, 0.0, s);
    launch_fill_f64(c->Q, V, 0.0, s);
    launch_fill_f64(c->W, V, 0.75, s);
    launch_fill_f64(c->omega, V, 0.0, s);
    launch_fill_f64(c->thres, 3 * V, 0.0, s);
    launch_tmask(c->thres, c->omega, p->arr_thr[0], p->arr_thr[1], p->arr_thr[2], c->tmask, V, s);
    launch_fill_f64(c->energy, 4, 0.0, s);
    launch_fill_f64(c->eblk, c->neblk, 0.0, s);
    launch_fill_f32((float*)c->flags, 4, 0.f, s);
    launch_fill_f32(c->mut, V, 0.f, s);
    e = hipMemsetAsync(c->mat, 0, (size_t)V, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = mc_prepare(c->dev);
    if (e == hipSuccess) {
        double* sp = nullptr;
        e = scratch_ptr(c->dev, &sp);
    }
    if (e != hipSuccess) {
        std::string m = std::string("device init: ") + hipGetErrorString(e);
        tam_destroy(c);
        return fail(nullptr, TAM_ERR_HIP, m);
    }
    c->cur = c->dT0;
    c->nxt = c->dT1;
    c->cur64 = c->T64a;
    c->nxt64 = c->T64b;
    *out = c;
    return TAM_OK;
}

void tam_destroy(tam_ctx* c)
{
    if (!c) return;
    Dev d(c->dev);
    if (c->stream) hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->dT0, c->dT1, c->gx, c->gy, c->gz, c->src, c->robin, c->kappa, c->density,
                    c->heatcap, c->coeff, c->rhokap, c->mut, c->mat, c->mask, c->mtab, c->mua_frac,
                    c->tally, c->jmean, c->Q, c->W, c->omega, c->thres, c->tmask, c->tn_cls,
                    c->energy, c->flags,
                    c->counters, c->T64a, c->T64b, c->ax, c->ay, c->az, c->cls, c->glut,
                    c->robin_lut, c->eblk, c->box_psum, c->box_geom, c->tn_omega, c->tn_thres, c->tn_Q,
                    c->tn_eblk, c->tn_tmask};
    for (void* q : ptrs)
        if (q) hipFree(q);
    cycle_reset(c);
    if (c->cyc_dev) hipFree(c->cyc_dev);
    if (c->cyc_host) hipHostFree(c->cyc_host);
    if (c->cnt_acc) hipFree(c->cnt_acc);
    for (int i = 0; i < 4; ++i)
        if (c->cyc_ev[i]) hipEventDestroy(c->cyc_ev[i]);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int tam_set_optics(tam_ctx* c, const uint8_t* mat, const tam_material* table, int32_t nmat,
                   const double* rhokap)
{
    if (!c) return TAM_ERR_ARG;
    cycle_reset(c);                 // captured cycles hold the old state / shapes
    if (!mat || !table || nmat < 1 || nmat > TAM_MAX_MATERIALS - 1)
        return fail(c, TAM_ERR_ARG, "need 1..254 materials and a material grid");
    Dev d(c->dev);
    const long V = c->nvox, H = c->nhalo;
    for (long v = 0; v < V; ++v)
        if (mat[v] >= nmat) return fail(c, TAM_ERR_ARG, "material id out of range");
    MatEntry tab[TAM_MAX_MATERIALS + 1];
    double frac[TAM_MAX_MATERIALS + 1];
    std::memset(tab, 0, sizeof(tab));
    std::memset(frac, 0, sizeof(frac));
    for (int m = 0; m < nmat; ++m) {
        const tam_material& t = table[m];
        if (!(t.mua >= 0 && t.mus >= 0 && t.g > -1 && t.g < 1 && t.n > 0))
            return fail(c, TAM_ERR_ARG, "material properties out of range");
        const double mt = t.mua + t.mus;
        if (c->p.thermal_mode == TAM_THERMAL_TABLE && !rhokap && mt <= 0.01)
            return fail(c, TAM_ERR_ARG, "tissue materials need mu_t > 0.01 cm^-1 (the removal threshold)");
        tab[m].albedo = mt > 0 ? (float)(t.mus / mt) : 0.f;
        tab[m].g = (float)t.g;
        tab[m].n = (float)t.n;
        tab[m].mua = (float)t.mua;
        tab[m].omg2 = (float)(1.0 - t.g * t.g);
        tab[m].opg2 = (float)(1.0 + t.g * t.g);
        tab[m].inv2g = t.g != 0.0 ? (float)(1.0 / (2.0 * t.g)) : 0.f;
        frac[m] = mt > 0 ? t.mua / mt : 0.0;
        c->host_tab[m] = t;
    }
    c->air_id = nmat;                     // removed voxels: no absorption, no scattering, n = 1
    tab[nmat].albedo = 0.f; tab[nmat].g = 0.f; tab[nmat].n = 1.f; tab[nmat].mua = 0.f;
    tab[nmat].omg2 = 1.f; tab[nmat].opg2 = 1.f; tab[nmat].inv2g = 0.f;
    frac[nmat] = 0.0;
    for (int m = 0; m <= nmat; ++m) {     // normal-incidence specular reflectance at the top face
        const float nt = (float)c->p.n_ext_zhi, ni = tab[m].n;
        const float rr = (nt > 0.f && nt != ni) ? (ni - nt) / (ni + nt) : 0.f;
        tab[m].rsp_top = rr * rr;
    }
    c->host_tab[nmat] = tam_material{0, 0, 0, 1};
    c->nmat = nmat;
    std::vector<double> rk;
    if (!rhokap) {
        rk.assign((size_t)H, 0.0);
        for (int k = 1; k <= c->p.nz; ++k)
            for (int j = 1; j <= c->p.ny; ++j)
                for (int i = 1; i <= c->p.nx; ++i) {
                    const long v = (i - 1) + (long)c->p.nx * ((j - 1) + (long)c->p.ny * (k - 1));
                    rk[(size_t)(i + c->px * j + c->pxy * k)] = table[mat[v]].mua + table[mat[v]].mus;
                }
        rhokap = rk.data();
    }
    std::vector<float> mut((size_t)V);
    for (int k = 1; k <= c->p.nz; ++k)
        for (int j = 1; j <= c->p.ny; ++j)
            for (int i = 1; i <= c->p.nx; ++i) {
                const long v = (i - 1) + (long)c->p.nx * ((j - 1) + (long)c->p.ny * (k - 1));
                mut[(size_t)v] = (float)rhokap[i + c->px * j + c->pxy * k];
            }
    HIPCHK(c, hipMemcpyAsync(c->mat, mat, (size_t)V, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->mtab, tab, sizeof(tab), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->mua_frac, frac, sizeof(frac), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->rhokap, rhokap, sizeof(double) * (size_t)H, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->mut, mut.data(), sizeof(float) * (size_t)V, hipMemcpyHostToDevice, c->stream));
    if (c->thermal_set) {                               // removed voxels follow rhokap
        HIPCHK(c, build_coeffs(c));
        const int rc = reclassify_from_device(c);
        if (rc != TAM_OK) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->optics_set = true;
    return TAM_OK;
}

int tam_set_thermal(tam_ctx* c, const double* temp, const double* kappa, const double* density,
                    const double* heatcap, const double* Q, const double* W)
{
    if (!c) return TAM_ERR_ARG;
    cycle_reset(c);                 // captured cycles hold the old state / shapes
    if (!temp || !kappa || !density || !heatcap) return fail(c, TAM_ERR_ARG, "null thermal array");
    Dev d(c->dev);
    const long H = c->nhalo, V = c->nvox;
    std::vector<float> dT((size_t)H);
    std::vector<double> coeff((size_t)H, 0.0);
    for (long i = 0; i < H; ++i) {
        if (!(density[i] > 0 && heatcap[i] > 0 && kappa[i] >= 0))
            return fail(c, TAM_ERR_ARG, "thermal properties must be positive");
        dT[(size_t)i] = (float)(temp[i] - c->p.t_ref);
    }
    for (int k = 1; k <= c->p.nz; ++k)
        for (int j = 1; j <= c->p.ny; ++j)
            for (int i = 1; i <= c->p.nx; ++i) {
                const long h = i + c->px * j + c->pxy * k;
                coeff[(size_t)h] = c->p.delt / (density[h] * heatcap[h]);
            }
    const size_t hb = sizeof(double) * (size_t)H;
    HIPCHK(c, hipMemcpyAsync(c->dT0, dT.data(), sizeof(float) * (size_t)H, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->dT1, dT.data(), sizeof(float) * (size_t)H, hipMemcpyHostToDevice, c->stream));
    if (c->fp64) {
        HIPCHK(c, hipMemcpyAsync(c->T64a, temp, hb, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->T64b, temp, hb, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->kappa, kappa, hb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->density, density, hb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->heatcap, heatcap, hb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->coeff, coeff.data(), hb, hipMemcpyHostToDevice, c->stream));
    if (Q) HIPCHK(c, hipMemcpyAsync(c->Q, Q, sizeof(double) * (size_t)V, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(c, launch_fill_f64(c->Q, V, 0.0, c->stream));
    if (W) HIPCHK(c, hipMemcpyAsync(c->W, W, sizeof(double) * (size_t)V, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(c, launch_fill_f64(c->W, V, 0.75, c->stream));
    HIPCHK(c, launch_fill_f64(c->energy, 4, 0.0, c->stream));
    HIPCHK(c, launch_fill_f64(c->eblk, c->neblk, 0.0, c->stream));
    HIPCHK(c, launch_fill_f32((float*)c->flags, 4, 0.f, c->stream));
    c->cur = c->dT0;
    c->nxt = c->dT1;
    c->cur64 = c->T64a;
    c->nxt64 = c->T64b;
    HIPCHK(c, build_coeffs(c));
    HIPCHK(c, refresh_src(c) == TAM_OK ? hipSuccess : hipErrorUnknown);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lut = false;
    if (c->optics_set) {
        std::vector<double> rk((size_t)H);
        HIPCHK(c, hipMemcpy(rk.data(), c->rhokap, sizeof(double) * (size_t)H, hipMemcpyDeviceToHost));
        const int rc = classify(c, kappa, density, heatcap, rk.data());
        if (rc != TAM_OK) return rc;
    }
    c->thermal_set = true;
    return TAM_OK;
}

int tam_mc_run(tam_ctx* c, int64_t npackets, uint64_t seed, uint32_t stream, int32_t accumulate,
               tam_mc_stats* stats)
{
    if (!c) return TAM_ERR_ARG;
    if (npackets < 0) return fail(c, TAM_ERR_ARG, "negative packet count");
    if (!c->optics_set) return fail(c, TAM_ERR_STATE, "tam_set_optics must come first");
    Dev d(c->dev);
    HIPCHK(c, launch_fill_u64(c->counters, TAM_CNT_SLOTS, c->stream));
    McArgs a = mc_args(c, npackets, seed, stream);
    a.clear_tally = accumulate ? 0 : 1;    // after the LDS-box placement has read the last run's tally
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_mc(a, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[TAM_CNT_N];
    HIPCHK(c, hipMemcpyAsync(cnt, c->counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timing.mc_ms = elapsed(c);
    c->src_valid = false;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->packets = (double)npackets;
        stats->absorbed = (double)cnt[TAM_CNT_ABSORBED];
        stats->specular = (double)cnt[TAM_CNT_SPECULAR];
        for (int f = 0; f < 6; ++f) stats->escaped[f] = (double)cnt[TAM_CNT_ESC0 + f];
        stats->missed = (double)cnt[TAM_CNT_MISSED];
        stats->crossings = (double)cnt[TAM_CNT_CROSS];
        stats->interactions = (double)cnt[TAM_CNT_INTER];
        stats->killed = (double)cnt[TAM_CNT_KILLED];
        stats->tally_adds = (double)cnt[TAM_CNT_TALLY];
        stats->box_adds = (double)cnt[TAM_CNT_BOX];
        stats->kernel_ms = c->timing.mc_ms;
    }
    return TAM_OK;
}

int tam_normalise(tam_ctx* c, double scale)
{
    if (!c) return TAM_ERR_ARG;
    if (!c->optics_set || !c->thermal_set) return fail(c, TAM_ERR_STATE, "optics and thermal state must be set");
    Dev d(c->dev);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_normalise(c->p.nx, c->p.ny, c->p.nz, c->px, c->pxy, c->tally, c->rhokap, c->mat,
                               c->mua_frac, nullptr, scale, c->coeff, c->jmean, c->src, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timing.normalise_ms = elapsed(c);
    return TAM_OK;
}

int tam_heat_run(tam_ctx* c, int32_t loops, tam_pulse* pulse)
{
    if (!c) return TAM_ERR_ARG;
    if (loops < 0) return fail(c, TAM_ERR_ARG, "negative loop count");
    if (!c->thermal_set) return fail(c, TAM_ERR_STATE, "tam_set_thermal must come first");
    Dev d(c->dev);
    const HeatArgs a = heat_args(c);
    const Heat64Args a64 = heat64_args(c);
    const int boil = c->p.boil != 0, robin = c->p.robin_top != 0, dmg = c->p.damage_every_step != 0;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    const bool two = two_step(c);
    for (int p = 0; p < loops; ++p) {
        const double lon = (pulse && pulse->use_pulse) ? pulse->laserOn : 1.0;
        double t_after = (pulse ? pulse->time : c->time) + c->p.delt;
        if (two && p + 1 < loops) {                  // steps p and p+1 in one launch
            if (pulse) tam_pulse_advance(pulse, c->p.delt);
            else c->time += c->p.delt;
            const double lon1 = (pulse && pulse->use_pulse) ? pulse->laserOn : 1.0;
            const double t_after1 = (pulse ? pulse->time : c->time) + c->p.delt;
            HIPCHK(c, launch_heat_step2(a, c->cur, c->nxt, (float)lon, (float)lon1, t_after, t_after1, robin, dmg,
                                        1, c->tile, c->kz, c->stream));
            std::swap(c->cur, c->nxt);
            if (pulse) tam_pulse_advance(pulse, c->p.delt);
            else c->time += c->p.delt;
            ++p;
            continue;
        }
        if (c->fp64) {
            HIPCHK(c, launch_heat64_step(a64, c->cur64, c->nxt64, lon, t_after, boil, robin, dmg, 1, c->stream));
            std::swap(c->cur64, c->nxt64);
        } else {
            HIPCHK(c, launch_heat_step(a, c->cur, c->nxt, (float)lon, t_after, boil, robin, dmg, 1, c->lut ? 1 : 0,
                                       c->tile, c->kz, c->stream));
            std::swap(c->cur, c->nxt);
        }
        if (pulse) tam_pulse_advance(pulse, c->p.delt);
        else c->time += c->p.delt;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    int rc = check_flags(c);
    c->timing.heat_ms = elapsed(c);
    c->timing.heat_steps = loops;
    c->timing.heat_step_ms = loops > 0 ? c->timing.heat_ms / loops : 0.0;
    return rc;
}

int tam_arrhenius(tam_ctx* c, double delt, double time)
{
    if (!c) return TAM_ERR_ARG;
    if (!c->thermal_set) return fail(c, TAM_ERR_STATE, "tam_set_thermal must come first");
    Dev d(c->dev);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_arrhenius(heat_args(c), c->cur, c->fp64 ? c->cur64 : nullptr, delt, time, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timing.arrhenius_ms = elapsed(c);
    return TAM_OK;
}

int tam_rebuild(tam_ctx* c)
{
    if (!c) return TAM_ERR_ARG;
    if (!c->thermal_set || !c->optics_set) return fail(c, TAM_ERR_STATE, "optics and thermal state must be set");
    Dev d(c->dev);
    const RebuildArgs a = rebuild_args(c);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (c->p.thermal_mode == TAM_THERMAL_WATER) HIPCHK(c, launch_water_update(a, c->stream));
    HIPCHK(c, launch_removal_mask(a, c->stream));
    HIPCHK(c, launch_apply_props(a, c->stream));
    HIPCHK(c, build_coeffs(c));
    if (refresh_src(c) != TAM_OK) return TAM_ERR_HIP;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timing.rebuild_ms = elapsed(c);
    return TAM_OK;
}

int tam_cycle(tam_ctx* c, int64_t npackets, uint64_t seed, uint32_t stream, double scale, int32_t loops,
              double arr_delt, tam_pulse* pulse, tam_mc_stats* stats)
{
    if (!c) return TAM_ERR_ARG;
    if (npackets < 0 || loops < 0) return fail(c, TAM_ERR_ARG, "negative packet or step count");
    if (!c->optics_set || !c->thermal_set) return fail(c, TAM_ERR_STATE, "optics and thermal state must be set");
    Dev d(c->dev);
    // capture key: a different step count, packet capacity, seed or damage
    // cadence needs a new graph (npackets == 0 reuses the current one)
    const int arr = arr_delt > 0.0 ? 1 : 0;
    const long cap = npackets > 0 ? (long)npackets : (c->cyc_loops >= 0 && c->cyc_npk > 0 ? c->cyc_npk : 1);
    if (loops != c->cyc_loops || cap != c->cyc_npk || seed != c->cyc_seed || arr != c->cyc_arr) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        cycle_reset(c);
        const size_t need = cycle_bytes(loops);
        if (need > c->cyc_bytes) {
            if (c->cyc_dev) HIPCHK(c, hipFree(c->cyc_dev));
            if (c->cyc_host) HIPCHK(c, hipHostFree(c->cyc_host));
            c->cyc_dev = nullptr;
            c->cyc_host = nullptr;
            c->cyc_bytes = 0;
            HIPCHK(c, hipMalloc((void**)&c->cyc_dev, need));
            HIPCHK(c, hipHostMalloc((void**)&c->cyc_host, 4 * need, hipHostMallocDefault));
            c->cyc_bytes = need;
            for (int i = 0; i < 4; ++i) c->cyc_ev_used[i] = false;
        }
        if (!c->cnt_acc) {
            HIPCHK(c, dalloc(&c->cnt_acc, TAM_CNT_N));
            HIPCHK(c, launch_fill_u64(c->cnt_acc, TAM_CNT_N, c->stream));
        }
        for (int i = 0; i < 4; ++i)
            if (!c->cyc_ev[i]) HIPCHK(c, hipEventCreateWithFlags(&c->cyc_ev[i], hipEventDisableTiming));
        c->cyc_loops = loops;
        c->cyc_npk = cap;
        c->cyc_seed = seed;
        c->cyc_arr = arr;
    }
    if (npackets > c->cyc_npk) return fail(c, TAM_ERR_ARG, "packet count above the captured capacity");

    // ---- this cycle's parameter block in a free staging slot
    const int slot = c->cyc_slot;
    if (c->cyc_ev_used[slot]) HIPCHK(c, hipEventSynchronize(c->cyc_ev[slot]));
    char* hb = c->cyc_host + (size_t)slot * c->cyc_bytes;
    CycleHdr* h = (CycleHdr*)hb;
    std::memset(h, 0, sizeof(CycleHdr));
    h->npk = npackets;
    h->stream = stream;
    h->mc_on = npackets > 0 ? 1 : 0;
    h->scale = scale;
    float* lon = (float*)(hb + cycle_lon_off());
    double* ta = (double*)(hb + cycle_ta_off(loops));
    for (int k = 0; k < loops; ++k) {             // the host side of heat_sim_3D's step loop
        lon[k] = (float)((pulse && pulse->use_pulse) ? pulse->laserOn : 1.0);
        ta[k] = (pulse ? pulse->time : c->time) + c->p.delt;
        if (pulse) tam_pulse_advance(pulse, c->p.delt);
        else c->time += c->p.delt;
    }
    h->arr[0] = arr_delt;
    h->arr[1] = pulse ? pulse->time : c->time;

    // ---- graph for the current ping-pong parity (captured on first use)
    const int par = c->fp64 ? (c->cur64 == c->T64a ? 0 : 1) : (c->cur == c->dT0 ? 0 : 1);
    if (!c->cyc_exec[par]) {
        const int rc = capture_cycle(c, par);
        if (rc != TAM_OK) return rc;
    }
    const bool timed = stats != nullptr;
    if (timed) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cyc_dev, hb, cycle_bytes(loops), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->cyc_ev[slot], c->stream));
    c->cyc_ev_used[slot] = true;
    c->cyc_slot = (slot + 1) % 4;
    HIPCHK(c, hipGraphLaunch(c->cyc_exec[par], c->stream));
    if (timed) HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if (c->cyc_swap[par]) {                        // the replay's net ping-pong swap
        std::swap(c->cur, c->nxt);
        std::swap(c->cur64, c->nxt64);
    }
    if (!timed) return TAM_OK;

    // ---- synchronous form: this call's (and earlier untimed calls') counters
    unsigned long long cnt[TAM_CNT_N];
    HIPCHK(c, hipMemcpyAsync(cnt, c->cnt_acc, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, launch_fill_u64(c->cnt_acc, TAM_CNT_N, c->stream));
    const int rc = check_flags(c);                 // synchronizes
    c->timing.cycle_ms = elapsed(c);
    std::memset(stats, 0, sizeof(*stats));
    stats->absorbed = (double)cnt[TAM_CNT_ABSORBED];
    stats->specular = (double)cnt[TAM_CNT_SPECULAR];
    for (int f = 0; f < 6; ++f) stats->escaped[f] = (double)cnt[TAM_CNT_ESC0 + f];
    stats->missed = (double)cnt[TAM_CNT_MISSED];
    stats->crossings = (double)cnt[TAM_CNT_CROSS];
    stats->interactions = (double)cnt[TAM_CNT_INTER];
    stats->killed = (double)cnt[TAM_CNT_KILLED];
    stats->tally_adds = (double)cnt[TAM_CNT_TALLY];
    stats->box_adds = (double)cnt[TAM_CNT_BOX];
    stats->packets = stats->absorbed + stats->specular + stats->missed + stats->killed;
    for (int f = 0; f < 6; ++f) stats->packets += stats->escaped[f];
    return rc;
}

int tam_tally_blocks(tam_ctx* c, int32_t k, double* sums)
{
    if (!c || !sums) return TAM_ERR_ARG;
    if (k < 1 || c->p.nx % k || c->p.ny % k || c->p.nz % k) return fail(c, TAM_ERR_ARG, "k must divide every grid edge");
    Dev d(c->dev);
    const long nb = (long)(c->p.nx / k) * (c->p.ny / k) * (c->p.nz / k);
    double* tmp = nullptr;
    HIPCHK(c, hipMalloc((void**)&tmp, sizeof(double) * (size_t)nb));
    hipError_t e = launch_block_sums(c->tally, c->p.nx, c->p.ny, c->p.nz, k, tmp, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sums, tmp, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(c, "tally blocks", e);
    return TAM_OK;
}

int tam_tally_blocks_noalloc(tam_ctx* c, int32_t k, double* sums)
{
    if (!c || !sums) return TAM_ERR_ARG;
    if (k < 1 || c->p.nx % k || c->p.ny % k || c->p.nz % k) return fail(c, TAM_ERR_ARG, "k must divide every grid edge");
    const long nb = (long)(c->p.nx / k) * (c->p.ny / k) * (c->p.nz / k);
    if (nb > TAM_BLOCKSUM_SCRATCH_MAX)
        return fail(c, TAM_ERR_ARG, "more blocks than the fixed buffer holds (TAM_BLOCKSUM_SCRATCH_MAX)");
    Dev d(c->dev);
    double* sp = nullptr;
    HIPCHK(c, scratch_ptr(c->dev, &sp));
    // every block sum is written (none accumulated), so the buffer needs no clearing
    HIPCHK(c, launch_block_sums(c->tally, c->p.nx, c->p.ny, c->p.nz, k, sp, c->stream));
    HIPCHK(c, hipMemcpyAsync(sums, sp, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TAM_OK;
}

int tam_download(tam_ctx* c, int32_t field, double* out)
{
    if (!c) return TAM_ERR_ARG;
    if (!out) return fail(c, TAM_ERR_ARG, "null output");
    Dev d(c->dev);
    const long H = c->nhalo, V = c->nvox;
    const double* src64 = nullptr;
    long n = 0;
    switch (field) {
    case TAM_FIELD_TEMP: {
        if (c->fp64) {
            HIPCHK(c, hipMemcpyAsync(out, c->cur64, sizeof(double) * (size_t)H, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            return TAM_OK;
        }
        std::vector<float> t((size_t)H);
        HIPCHK(c, hipMemcpyAsync(t.data(), c->cur, sizeof(float) * (size_t)H, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (long i = 0; i < H; ++i) out[i] = c->p.t_ref + (double)t[(size_t)i];
        return TAM_OK;
    }
    case TAM_FIELD_MAT: {
        std::vector<uint8_t> m((size_t)V);
        HIPCHK(c, hipMemcpyAsync(m.data(), c->mat, (size_t)V, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (long i = 0; i < V; ++i) out[i] = m[(size_t)i];
        return TAM_OK;
    }
    case TAM_FIELD_JMEAN: src64 = c->jmean; n = V; break;
    case TAM_FIELD_TALLY: src64 = c->tally; n = V; break;
    case TAM_FIELD_RHOKAP: src64 = c->rhokap; n = H; break;
    case TAM_FIELD_TISSUE: src64 = c->omega; n = V; break;
    case TAM_FIELD_THRESTIME: src64 = c->thres; n = 3 * V; break;
    case TAM_FIELD_WATER: src64 = c->W; n = V; break;
    case TAM_FIELD_Q: src64 = c->Q; n = V; break;
    case TAM_FIELD_KAPPA: src64 = c->kappa; n = H; break;
    case TAM_FIELD_DENSITY: src64 = c->density; n = H; break;
    case TAM_FIELD_HEATCAP: src64 = c->heatcap; n = H; break;
    case TAM_FIELD_COEFF: src64 = c->coeff; n = H; break;
    default: return fail(c, TAM_ERR_ARG, "unknown field");
    }
    HIPCHK(c, hipMemcpyAsync(out, src64, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TAM_OK;
}

int tam_upload(tam_ctx* c, int32_t field, const double* in)
{
    if (!c) return TAM_ERR_ARG;
    cycle_reset(c);                 // captured cycles hold the old state / shapes
    if (!in) return fail(c, TAM_ERR_ARG, "null input");
    Dev d(c->dev);
    const long H = c->nhalo, V = c->nvox;
    double* dst = nullptr;
    long n = 0;
    switch (field) {
    case TAM_FIELD_TEMP: {
        std::vector<float> t((size_t)H);
        for (long i = 0; i < H; ++i) t[(size_t)i] = (float)(in[i] - c->p.t_ref);
        HIPCHK(c, hipMemcpyAsync(c->dT0, t.data(), sizeof(float) * (size_t)H, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->dT1, t.data(), sizeof(float) * (size_t)H, hipMemcpyHostToDevice, c->stream));
        if (c->fp64) {
            HIPCHK(c, hipMemcpyAsync(c->T64a, in, sizeof(double) * (size_t)H, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->T64b, in, sizeof(double) * (size_t)H, hipMemcpyHostToDevice, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->cur = c->dT0;
        c->nxt = c->dT1;
        c->cur64 = c->T64a;
        c->nxt64 = c->T64b;
        return TAM_OK;
    }
    case TAM_FIELD_RHOKAP: {
        std::vector<float> mut((size_t)V);
        for (int k = 1; k <= c->p.nz; ++k)
            for (int j = 1; j <= c->p.ny; ++j)
                for (int i = 1; i <= c->p.nx; ++i)
                    mut[(size_t)((i - 1) + (long)c->p.nx * ((j - 1) + (long)c->p.ny * (k - 1)))] =
                        (float)in[i + c->px * j + c->pxy * k];
        HIPCHK(c, hipMemcpyAsync(c->mut, mut.data(), sizeof(float) * (size_t)V, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->rhokap, in, sizeof(double) * (size_t)H, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->thermal_set) {                        // removed voxels follow rhokap
            HIPCHK(c, build_coeffs(c));
            return reclassify_from_device(c);
        }
        return TAM_OK;
    }
    case TAM_FIELD_JMEAN: dst = c->jmean; n = V; break;
    case TAM_FIELD_TALLY: dst = c->tally; n = V; break;
    case TAM_FIELD_TISSUE: dst = c->omega; n = V; break;
    case TAM_FIELD_THRESTIME: dst = c->thres; n = 3 * V; break;
    case TAM_FIELD_WATER: dst = c->W; n = V; break;
    case TAM_FIELD_Q: dst = c->Q; n = V; break;
    case TAM_FIELD_COEFF: dst = c->coeff; n = H; break;
    default: return fail(c, TAM_ERR_ARG, "field cannot be uploaded");
    }
    HIPCHK(c, hipMemcpyAsync(dst, in, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    if (field == TAM_FIELD_THRESTIME || field == TAM_FIELD_TISSUE) {
        // the mask's set bits and due chains follow the new times / Omega,
        // and the class bytes mirror the due chains
        HIPCHK(c, launch_tmask(c->thres, c->omega, c->p.arr_thr[0], c->p.arr_thr[1], c->p.arr_thr[2], c->tmask, V,
                               c->stream));
        if (c->lut)
            HIPCHK(c, launch_class_due(c->cls, c->tmask, c->p.nx, c->p.ny, c->p.nz, c->px, c->pxy, c->stream));
    }
    if (field == TAM_FIELD_JMEAN || field == TAM_FIELD_COEFF) {
        if (refresh_src(c) != TAM_OK) return TAM_ERR_HIP;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TAM_OK;
}

int tam_energy(tam_ctx* c, double* heat_content, double* boundary_out)
{
    if (!c) return TAM_ERR_ARG;
    Dev d(c->dev);
    HIPCHK(c, launch_fill_f64(c->energy + 1, 1, 0.0, c->stream));
    HIPCHK(c, launch_energy(c->p.nx, c->p.ny, c->p.nz, c->px, c->pxy, c->cur, c->fp64 ? c->cur64 : nullptr,
                            c->p.t_ref, c->heatcap, c->p.mass_voxel, c->energy + 1, c->stream));
    double e[2];
    std::vector<double> blk((size_t)c->neblk);
    HIPCHK(c, hipMemcpyAsync(e, c->energy, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(blk.data(), c->eblk, sizeof(double) * blk.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (double b : blk) e[0] += b;
    if (boundary_out) *boundary_out = e[0];
    if (heat_content) *heat_content = e[1];
    return TAM_OK;
}

int tam_get_timing(tam_ctx* c, tam_timing* out)
{
    if (!c || !out) return TAM_ERR_ARG;
    *out = c->timing;
    out->class_mode = c->lut ? 1.0 : 0.0;
    out->tile = c->tile;
    out->kz = c->kz;
    return TAM_OK;
}

int tam_copy_bandwidth(tam_ctx* c, int64_t bytes, int32_t reps, double* gbps)
{
    if (!c || !gbps) return TAM_ERR_ARG;
    if (bytes < 16 || reps < 1) return fail(c, TAM_ERR_ARG, "need bytes >= 16 and reps >= 1");
    Dev d(c->dev);
    bytes &= ~(int64_t)15;
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, (size_t)bytes));
    hipError_t e = hipMalloc(&b, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return hip_fail(c, "hipMalloc", e);
    }
    e = hipMemsetAsync(a, 0, (size_t)bytes, c->stream);
    if (e == hipSuccess) e = launch_copy(a, b, bytes, c->stream);          // warm-up
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(a, b, bytes, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
    const double ms = elapsed(c) / reps;
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(c, "copy", e);
    *gbps = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    return TAM_OK;
}

int tam_atomic_rate(tam_ctx* c, int64_t count, int64_t slots, double* per_s)
{
    if (!c || !per_s) return TAM_ERR_ARG;
    if (count < 1 || slots < 64 || slots > 0xffffffffLL)
        return fail(c, TAM_ERR_ARG, "need count >= 1 and 64 <= slots < 2^32");
    Dev d(c->dev);
    double* t = nullptr;
    HIPCHK(c, hipMalloc(&t, (size_t)slots * sizeof(double)));
    const long lanes = 2048L * 256L;
    const long per = count / lanes < 1 ? 1 : count / lanes;
    hipError_t e = launch_fill_f64(t, slots, 0.0, c->stream);
    if (e == hipSuccess) e = launch_atomic_rate(t, slots, lanes, c->stream);        // warm-up
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess) e = launch_atomic_rate(t, slots, per * lanes, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
    const double ms = elapsed(c);
    (void)hipFree(t);
    if (e != hipSuccess) return hip_fail(c, "atomic rate", e);
    *per_s = (double)(per * lanes) / (ms * 1e-3);
    return TAM_OK;
}

int tam_set_mc_launch(tam_ctx* c, int32_t blocks_per_cu, int32_t rng_batch)
{
    if (!c) return TAM_ERR_ARG;
    cycle_reset(c);                 // captured cycles hold the old state / shapes
    if (blocks_per_cu < 1 || blocks_per_cu > 64) return fail(c, TAM_ERR_ARG, "blocks_per_cu must be in [1, 64]");
    if (rng_batch < 1 || rng_batch > 64) return fail(c, TAM_ERR_ARG, "rng_batch must be in [1, 64]");
    c->mc_bpc = blocks_per_cu;
    c->mc_batch = rng_batch;
    return TAM_OK;
}

int tam_set_mc_box(tam_ctx* c, int32_t enable)
{
    if (!c) return TAM_ERR_ARG;
    if (enable < 0 || enable > 2) return fail(c, TAM_ERR_ARG, "box mode must be 0 (off), 1 (fixed point) or 2 (FP64)");
    cycle_reset(c);
    c->mc_box = enable;
    return TAM_OK;
}

int tam_get_mc_box(tam_ctx* c, int32_t* geom)
{
    if (!c || !geom) return TAM_ERR_ARG;
    Dev d(c->dev);
    int g[6];
    HIPCHK(c, hipMemcpyAsync(g, c->box_geom, sizeof(g), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < 6; ++i) geom[i] = g[i];
    return TAM_OK;
}

int tam_tune_mc(tam_ctx* c, int64_t sample, int32_t* blocks_per_cu, int32_t* rng_batch, double* packets_per_s)
{
    if (!c) return TAM_ERR_ARG;
    if (sample < 1) return fail(c, TAM_ERR_ARG, "sample must be positive");
    const int old_bpc = c->mc_bpc, old_batch = c->mc_batch, old_box = c->mc_box;
    const int bpcs[] = {1, 2, 3, 4, 6, 8};
    const int batches[] = {8, 16, 24, 32, 40, 48};
    double best = 0.0;
    int bb = c->mc_bpc, bt = c->mc_batch, bx = c->mc_box;
    const int box_on = old_box ? old_box : 1;           // the box's kind is kept
    struct Cand {
        int bpc, batch, box;
        double rate;
    };
    std::vector<Cand> rates;
    for (int x = 0; x < 2; ++x)
        for (int b : bpcs) {
            for (int t : batches) {
                c->mc_box = x ? box_on : 0;
                c->mc_bpc = b;
                c->mc_batch = t;
                tam_mc_stats st;
                const int rc = tam_mc_run(c, sample, 0x5eed5eedULL, 0xffffffffu, 0, &st);
                if (rc != TAM_OK) return rc;
                const double rate = st.kernel_ms > 0 ? sample / (st.kernel_ms * 1e-3) : 0.0;
                rates.push_back({b, t, c->mc_box, rate});
                if (rate > best) { best = rate; bb = b; bt = t; bx = c->mc_box; }
            }
        }
    // Ties go to the single CU-wide box (box on, 1024-lane blocks, one per
    // CU): candidates are timed on the tissue as it is, and once a crater
    // forms and the deposition moves below it, the largest box holds its lead
    // (c1 crater phase: 360 ms per 1e8 packets against 465 ms for three
    // boxes per CU, which can be a few percent ahead on fresh tissue).
    if (!(bx && bb == 4)) {
        double wide = 0.0;
        int wt = bt;
        for (size_t q = 0; q < rates.size(); ++q)
            if (rates[q].box && rates[q].bpc == 4 && rates[q].rate > wide) { wide = rates[q].rate; wt = rates[q].batch; }
        if (wide >= (1.0 - kWideBoxMargin) * best) {
            bb = 4;
            bt = wt;
            bx = box_on;
            best = wide;
        }
    }
    // With the CU-wide box, a random-phase batch of 32 is kept when it times
    // within kBatchMargin of the fastest batch: on fresh tissue the batches
    // 8-48 tie within the timing noise, while in the c1 crater phase 24-32
    // run fastest (318 ms per 1e8 packets against 348 ms at 8 and 353 at 48).
    if (bx && bb == 4 && bt != 32) {
        for (size_t q = 0; q < rates.size(); ++q)
            if (rates[q].box && rates[q].bpc == 4 && rates[q].batch == 32 &&
                rates[q].rate >= (1.0 - kBatchMargin) * best) {
                bt = 32;
                best = rates[q].rate;
            }
    }
    c->mc_bpc = bb;
    c->mc_batch = bt;
    c->mc_box = bx;
    if (bb != old_bpc || bt != old_batch || bx != old_box) cycle_reset(c);   // captured cycles use the old shape
    Dev d(c->dev);
    HIPCHK(c, launch_fill_f64(c->tally, c->nvox, 0.0, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (blocks_per_cu) *blocks_per_cu = bb;
    if (rng_batch) *rng_batch = bt;
    if (packets_per_s) *packets_per_s = best;
    return TAM_OK;
}

int tam_set_heat_launch(tam_ctx* c, int32_t tile, int32_t kz)
{
    if (!c) return TAM_ERR_ARG;
    cycle_reset(c);                 // captured cycles hold the old state / shapes
    if (tile < 0 || tile >= heat_num_tiles() || kz < 1) return fail(c, TAM_ERR_ARG, "bad stencil launch shape");
    c->tile = tile;
    c->kz = kz;
    return TAM_OK;
}

int tam_tune_heat(tam_ctx* c, int32_t reps, int32_t* tile, int32_t* kz, double* step_ms)
{
    if (!c) return TAM_ERR_ARG;
    if (!c->thermal_set) return fail(c, TAM_ERR_STATE, "tam_set_thermal must come first");
    if (reps < 1) reps = 1;
    Dev d(c->dev);
    double best = 1e300;
    int bt = c->tile, bk = c->kz;
    if (!c->fp64) {
        // Candidates run the real step (same flags as tam_heat_run) on the
        // current field, cur -> nxt without swapping, so the state does not
        // advance; the damage integral, threshold times, their mask, Q and
        // the energy slots they update are scratch copies of the live ones,
        // and the step time is the next step's, so the threshold chain does
        // the work it does in the run.
        HeatArgs a = heat_args(c);
        const long V = c->nvox;
        // scratch: allocated once per context (the tuner runs every few cycles)
        hipError_t e = hipSuccess;
        if (!c->tn_omega) e = dalloc(&c->tn_omega, V);
        if (e == hipSuccess && !c->tn_thres) e = dalloc(&c->tn_thres, 3 * V);
        if (e == hipSuccess && !c->tn_eblk) e = dalloc(&c->tn_eblk, c->neblk);
        if (e == hipSuccess && !c->tn_Q) e = dalloc(&c->tn_Q, V);
        if (e == hipSuccess && !c->tn_tmask) e = dalloc(&c->tn_tmask, V);
        if (e == hipSuccess && c->lut && !c->tn_cls) e = dalloc(&c->tn_cls, c->nhalo);
        double *om = c->tn_omega, *th = c->tn_thres, *eb = c->tn_eblk, *qs = c->tn_Q;
        uint8_t* tm = c->tn_tmask;
        if (e == hipSuccess) e = hipMemcpyAsync(om, c->omega, sizeof(double) * (size_t)V, hipMemcpyDeviceToDevice,
                                                c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(qs, c->Q, sizeof(double) * (size_t)V, hipMemcpyDeviceToDevice,
                                                c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(th, c->thres, sizeof(double) * 3 * (size_t)V,
                                                hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tm, c->tmask, (size_t)V, hipMemcpyDeviceToDevice, c->stream);
        // the steps keep the class bytes' due bits: candidates write a copy
        if (e == hipSuccess && c->lut)
            e = hipMemcpyAsync(c->tn_cls, c->cls, (size_t)c->nhalo, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = launch_fill_f64(eb, c->neblk, 0.0, c->stream);
        a.omega = om;
        a.thres = th;
        a.tmask = tm;
        if (c->lut) a.cls = c->tn_cls;
        a.eblk = eb;
        a.Q = qs;
        const int boil = c->p.boil != 0, robin = c->p.robin_top != 0, dmg = c->p.damage_every_step != 0;
        const int kzs[] = {4, 8, 12, 16, 32, 64};
        const bool two_ok = !boil && c->lut;
        const double ta = c->time + c->p.delt;
        for (int t = 0; t < heat_num_tiles() && e == hipSuccess; ++t) {
            const bool two = heat_tile_steps(t) == 2;
            if (two && !two_ok) continue;             // would run its one-step twin, timed already
            // class mode: the LDS-plane shapes (0-2) are slower than their
            // direct twins on every measured workload; they stay settable
            if (c->lut && t <= 2) continue;
            for (int kk : kzs) {
                if (kk > c->p.nz && kk != 4) continue;
                // a two-step candidate launches two steps, timed per step
                auto run = [&]() {
                    return two ? launch_heat_step2(a, c->cur, c->nxt, 1.f, 1.f, ta, ta + c->p.delt, robin, dmg, 1, t,
                                                   kk, c->stream)
                               : launch_heat_step(a, c->cur, c->nxt, 1.f, ta, boil, robin, dmg, 1, c->lut ? 1 : 0,
                                                  t, kk, c->stream);
                };
                e = run();
                if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
                for (int r = 0; r < reps && e == hipSuccess; ++r) e = run();
                if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
                if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
                if (e != hipSuccess) break;
                const double ms = elapsed(c) / (reps * (two ? 2 : 1));
                if (ms < best) { best = ms; bt = t; bk = kk; }
            }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, "heat tuning", e);
        if (bt != c->tile || bk != c->kz) cycle_reset(c);     // captured cycles use the old shape
        c->tile = bt;
        c->kz = bk;
    }
    if (tile) *tile = c->tile;
    if (kz) *kz = c->kz;
    if (step_ms) *step_ms = best < 1e300 ? best : 0.0;
    return TAM_OK;
}

int tam_synchronize(tam_ctx* c)
{
    if (!c) return TAM_ERR_ARG;
    Dev d(c->dev);
    return check_flags(c);           // waits for the stream, then reports T < 0 seen since
}

}  // extern "C"

// ---- batches of instances ------------------------------------------------

struct tam_batch {
    std::vector<tam_ctx*> ctx;
    int dev;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::string err;
    McArgs* dmc;
    HeatBatchEntry* dheat;
    int* dtile;
    int tile, kz, bpc, rng_batch;
    double last_ms;
};

namespace {

int bfail(tam_batch* b, int code, const std::string& m)
{
    b->err = m;
    return code;
}
#define BCHK(b, call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return bfail((b), TAM_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

thread_local std::string g_batch_err;

float bms(tam_batch* b)
{
    float ms = 0.f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    return ms;
}

int sync_ctxs(tam_batch* b)
{
    for (tam_ctx* c : b->ctx) BCHK(b, hipStreamSynchronize(c->stream));
    return TAM_OK;
}

// one batched transport launch (counters of every instance cleared first)
int batch_mc(tam_batch* b, const int64_t* npk, uint64_t seed, uint32_t stream0, int accumulate)
{
    const int n = (int)b->ctx.size();
    std::vector<McArgs> tab((size_t)n);
    long total = 0;
    for (int i = 0; i < n; ++i) {
        tam_ctx* c = b->ctx[(size_t)i];
        if (!accumulate) BCHK(b, launch_fill_f64(c->tally, c->nvox, 0.0, b->stream));
        BCHK(b, launch_fill_u64(c->counters, TAM_CNT_SLOTS, b->stream));
        tab[(size_t)i] = mc_args(c, npk[i], seed, stream0 + (uint32_t)i);
        tab[(size_t)i].blocks_per_cu = b->bpc;
        tab[(size_t)i].rng_batch = b->rng_batch;
        total += (long)npk[i];
        c->src_valid = false;
    }
    BCHK(b, hipMemcpyAsync(b->dmc, tab.data(), sizeof(McArgs) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    BCHK(b, launch_mc_batch(b->dmc, n, total, b->bpc, b->stream));
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    b->last_ms = bms(b);
    return TAM_OK;
}

// tile table for the batch's shape; returns the total tile count
int batch_tiles(tam_batch* b, int tile, int kz, std::vector<int>& start)
{
    int tx = 0, ty = 0;
    heat_tile_dims(tile, &tx, &ty);
    const int n = (int)b->ctx.size();
    start.assign((size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const tam_params& p = b->ctx[(size_t)i]->p;
        const long t = (long)((p.nx + tx - 1) / tx) * ((p.ny + ty - 1) / ty) * ((p.nz + kz - 1) / kz);
        start[(size_t)i + 1] = start[(size_t)i] + (int)t;
    }
    return start[(size_t)n];
}

// `steps` batched heat steps with the given shape; advance = 0 keeps the
// state (every step reads cur and writes nxt, no energy accounting)
int batch_heat(tam_batch* b, int steps, int tile, int kz, bool advance)
{
    const int n = (int)b->ctx.size();
    std::vector<HeatBatchEntry> tab((size_t)n);
    for (int i = 0; i < n; ++i) {
        tam_ctx* c = b->ctx[(size_t)i];
        tab[(size_t)i].a = heat_args(c);
        tab[(size_t)i].cur = c->cur;
        tab[(size_t)i].nxt = c->nxt;
    }
    const tam_params& p0 = b->ctx[0]->p;
    const int robin = p0.robin_top != 0, lut = b->ctx[0]->lut ? 1 : 0;
    // two-step shapes pair the steps (class mode); single steps run on the
    // one-step twin shape, direct 32x8 for 6 and 7, direct 64x8 for 8
    const bool two = heat_tile_steps(tile) == 2 && lut;
    const int one = tile == 8 ? 5 : (tile >= 6 ? 4 : tile);
    std::vector<int> start(2 * ((size_t)n + 1), 0), s1, s2;
    const int total1 = batch_tiles(b, one, kz, s1);
    const int total2 = two ? batch_tiles(b, tile, kz, s2) : 0;
    std::copy(s1.begin(), s1.end(), start.begin());
    if (two) std::copy(s2.begin(), s2.end(), start.begin() + (n + 1));
    BCHK(b, hipMemcpyAsync(b->dheat, tab.data(), sizeof(HeatBatchEntry) * (size_t)n, hipMemcpyHostToDevice,
                           b->stream));
    BCHK(b, hipMemcpyAsync(b->dtile, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    int par = 0;                       // launches so far, mod 2: which of cur / nxt holds the state
    for (int k = 0; k < steps; ++par) {
        const int pz = advance ? (par & 1) : 0;
        if (two && k + 1 < steps) {
            BCHK(b, launch_heat_batch2(b->dheat, n, b->dtile + (n + 1), total2, 1.f, 1.f, 0.0, 0.0, robin, 0,
                                       advance ? 1 : 0, tile, kz, pz, b->stream));
            k += 2;
        } else {
            BCHK(b, launch_heat_batch(b->dheat, n, b->dtile, total1, 1.f, 0.0, 0, robin, 0, advance ? 1 : 0, lut, one,
                                      kz, pz, b->stream));
            k += 1;
        }
    }
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    b->last_ms = bms(b);
    if (advance) {
        for (tam_ctx* c : b->ctx) {
            if (par & 1) std::swap(c->cur, c->nxt);
            c->time += steps * c->p.delt;
            c->timing.heat_steps = steps;
        }
    }
    return TAM_OK;
}

}  // namespace

extern "C" {

int tam_batch_create(tam_ctx* const* ctxs, int32_t n, tam_batch** out)
{
    if (!ctxs || !out || n < 1) return TAM_ERR_ARG;
    *out = nullptr;
    const tam_ctx* c0 = ctxs[0];
    for (int i = 0; i < n; ++i) {
        const tam_ctx* c = ctxs[i];
        if (!c || c->dev != c0->dev) { g_batch_err = "instances must share one device"; return TAM_ERR_ARG; }
        if (!c->optics_set || !c->thermal_set) { g_batch_err = "instances need optics and thermal state"; return TAM_ERR_STATE; }
        if (c->fp64 || c->p.boil || c->p.damage_every_step || c->p.robin_top != c0->p.robin_top || c->lut != c0->lut) {
            g_batch_err = "batched instances run the FP32 step without boiling or in-step damage, "
                          "with one convective-top setting and one coefficient mode";
            return TAM_ERR_ARG;
        }
    }
    tam_batch* b = new (std::nothrow) tam_batch();
    if (!b) return TAM_ERR_NOMEM;
    b->ctx.assign(ctxs, ctxs + n);
    b->dev = c0->dev;
    b->tile = 4;
    b->kz = 8;
    b->bpc = 4;
    b->rng_batch = 32;
    Dev d(b->dev);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipMalloc((void**)&b->dmc, sizeof(McArgs) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void**)&b->dheat, sizeof(HeatBatchEntry) * (size_t)n);
    // two tile tables: the launch shape's and its one-step twin's
    if (e == hipSuccess) e = hipMalloc((void**)&b->dtile, sizeof(int) * 2 * ((size_t)n + 1));
    if (e != hipSuccess) {
        g_batch_err = std::string("batch allocation: ") + hipGetErrorString(e);
        tam_batch_destroy(b);
        return TAM_ERR_HIP;
    }
    *out = b;
    return TAM_OK;
}

void tam_batch_destroy(tam_batch* b)
{
    if (!b) return;
    Dev d(b->dev);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->dmc) hipFree(b->dmc);
    if (b->dheat) hipFree(b->dheat);
    if (b->dtile) hipFree(b->dtile);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

const char* tam_batch_last_error(const tam_batch* b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int tam_batch_mc_run(tam_batch* b, const int64_t* npackets, uint64_t seed, uint32_t stream0, int32_t accumulate,
                     tam_mc_stats* stats)
{
    if (!b || !npackets) return TAM_ERR_ARG;
    const int n = (int)b->ctx.size();
    for (int i = 0; i < n; ++i)
        if (npackets[i] < 0) return bfail(b, TAM_ERR_ARG, "negative packet count");
    Dev d(b->dev);
    int rc = sync_ctxs(b);
    if (rc == TAM_OK) rc = batch_mc(b, npackets, seed, stream0, accumulate);
    if (rc != TAM_OK || !stats) return rc;
    for (int i = 0; i < n; ++i) {
        unsigned long long cnt[TAM_CNT_N];
        BCHK(b, hipMemcpy(cnt, b->ctx[(size_t)i]->counters, sizeof(cnt), hipMemcpyDeviceToHost));
        tam_mc_stats& s = stats[i];
        std::memset(&s, 0, sizeof(s));
        s.packets = (double)npackets[i];
        s.absorbed = (double)cnt[TAM_CNT_ABSORBED];
        s.specular = (double)cnt[TAM_CNT_SPECULAR];
        for (int f = 0; f < 6; ++f) s.escaped[f] = (double)cnt[TAM_CNT_ESC0 + f];
        s.missed = (double)cnt[TAM_CNT_MISSED];
        s.crossings = (double)cnt[TAM_CNT_CROSS];
        s.interactions = (double)cnt[TAM_CNT_INTER];
        s.killed = (double)cnt[TAM_CNT_KILLED];
        s.tally_adds = (double)cnt[TAM_CNT_TALLY];
        s.box_adds = (double)cnt[TAM_CNT_BOX];
        s.kernel_ms = b->last_ms;
    }
    return TAM_OK;
}

int tam_batch_heat_run(tam_batch* b, int32_t loops)
{
    if (!b || loops < 0) return TAM_ERR_ARG;
    Dev d(b->dev);
    int rc = sync_ctxs(b);
    if (rc == TAM_OK) rc = batch_heat(b, loops, b->tile, b->kz, true);
    if (rc != TAM_OK) return rc;
    for (tam_ctx* c : b->ctx) {
        c->timing.heat_ms = b->last_ms;
        const int r = check_flags(c);
        if (r != TAM_OK) return bfail(b, r, c->err);
    }
    return TAM_OK;
}

int tam_batch_set_launch(tam_batch* b, int32_t tile, int32_t kz, int32_t blocks_per_cu, int32_t rng_batch)
{
    if (!b) return TAM_ERR_ARG;
    if (tile < 3 || tile > 8 || kz < 1 || blocks_per_cu < 1 || blocks_per_cu > 64 || rng_batch < 1 || rng_batch > 64)
        return bfail(b, TAM_ERR_ARG, "batch launch: tile 3..8, kz >= 1, blocks_per_cu 1..64, rng_batch 1..64");
    b->tile = tile;
    b->kz = kz;
    b->bpc = blocks_per_cu;
    b->rng_batch = rng_batch;
    return TAM_OK;
}

int tam_batch_tune(tam_batch* b, const int64_t* sample_packets, int32_t heat_reps, int32_t* tile, int32_t* kz,
                   int32_t* blocks_per_cu, int32_t* rng_batch)
{
    if (!b || !sample_packets) return TAM_ERR_ARG;
    if (heat_reps < 1) heat_reps = 1;
    Dev d(b->dev);
    int rc = sync_ctxs(b);
    if (rc != TAM_OK) return rc;
    // transport: blocks per CU x random-phase batch on the sample counts
    double best = 1e300;
    int bb = b->bpc, bt = b->rng_batch;
    for (int bpc : {2, 4, 8})
        for (int rb : {16, 32, 48}) {
            b->bpc = bpc;
            b->rng_batch = rb;
            rc = batch_mc(b, sample_packets, 0x5eed5eedULL, 0xfffff000u, 0);
            if (rc != TAM_OK) return rc;
            if (b->last_ms < best) { best = b->last_ms; bb = bpc; bt = rb; }
        }
    b->bpc = bb;
    b->rng_batch = bt;
    for (tam_ctx* c : b->ctx) BCHK(b, launch_fill_f64(c->tally, c->nvox, 0.0, b->stream));
    // stencil: direct and two-step shapes x z depths, state not advanced
    // (heat_reps steps per candidate: two-step shapes run them in pairs)
    best = 1e300;
    int btl = b->tile, bk = b->kz;
    for (int t = 3; t <= 8; ++t)
        for (int k : {4, 8, 16, 32}) {
            rc = batch_heat(b, 1, t, k, false);            // warm-up
            if (rc == TAM_OK) rc = batch_heat(b, heat_reps, t, k, false);
            if (rc != TAM_OK) return rc;
            if (b->last_ms < best) { best = b->last_ms; btl = t; bk = k; }
        }
    b->tile = btl;
    b->kz = bk;
    BCHK(b, hipStreamSynchronize(b->stream));
    if (tile) *tile = btl;
    if (kz) *kz = bk;
    if (blocks_per_cu) *blocks_per_cu = bb;
    if (rng_batch) *rng_batch = bt;
    return TAM_OK;
}

double tam_batch_last_ms(const tam_batch* b) { return b ? b->last_ms : 0.0; }

}  // extern "C"
