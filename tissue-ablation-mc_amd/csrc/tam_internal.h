// tam_internal.h -- device data layout, kernel argument blocks and launcher
// prototypes shared by the HIP translation units of libtam.so.
//
// HBM layout (all arrays column-major, x fastest, like the reference):
//   halo grids   (nx+2)*(ny+2)*(nz+2): FP32 dT ping-pong, FP32 face
//                coefficients gx/gy/gz, FP32 heat source, FP64 kappa,
//                density, heatcap, coeff, rhokap
//   interior     nx*ny*nz: FP32 mu_t, uint8 material id, FP64 path-length
//                tally, FP64 jmean, FP64 Q, FP64 water, FP64 Omega,
//                FP64 threshold times (x3), uint8 mask of the threshold
//                times already set (bit k: time k != 0)
// Temperatures live on the device as dT = T - t_ref in FP32: the explicit
// update adds increments far below the FP32 ulp of an absolute Kelvin value,
// while relative to t_ref they keep ~1e-7 relative precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/tam.h"

#define TAM_MAX_MATERIALS 255

// Philox4x32-10 (Salmon et al., SC'11): counter c, key (k0, k1).
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

// First-crossing times of the damage thresholds (3dFD.f90:447-464): one call
// of the reference's chain "if (ThresTime(1) == 0 .and. Omega >= thr1) ...
// else if (ThresTime(2) == 0 ...) ... else if (ThresTime(3) == 0 ...)" at
// time t, with each "ThresTime(k) == 0" read from bit k of the voxel's mask m
// instead of from the FP64 time, so a step reads one byte rather than up to
// three doubles.  A time written as 0.0 leaves its bit clear, as ThresTime(k)
// stays 0 then.
//
// The reference runs the chain for every voxel on every call, inside the
// Arrhenius window or not, and writes at most one time per call.  A voxel
// whose Omega passed two thresholds in one step therefore gets the second
// time on the next call whatever its temperature: after the call, bit
// TAM_TMASK_DUE says that some threshold with an unset time is already
// reached, so the step kernels run the chain for that voxel in the next step
// even outside the window.  Returns the voxel's mask after the call (stored
// when it changed).
#define TAM_TMASK_DUE 8u
__device__ __forceinline__ unsigned thres_chain(double* thres, uint8_t* tmask, long v, long nv, double om,
                                                double thr0, double thr1, double thr2, double t, unsigned m)
{
    const unsigned m0 = m;
    int k = -1;
    if (!(m & 1u) && om >= thr0) k = 0;
    else if (!(m & 2u) && om >= thr1) k = 1;
    else if (!(m & 4u) && om >= thr2) k = 2;
    m &= ~TAM_TMASK_DUE;                   // nothing written: nothing is due either
    if (k >= 0) {
        thres[v + k * nv] = t;
        if (t != 0.0) m |= 1u << k;
        if ((!(m & 1u) && om >= thr0) || (!(m & 2u) && om >= thr1) || (!(m & 4u) && om >= thr2))
            m |= TAM_TMASK_DUE;
    }
    if (m != m0) tmask[v] = (uint8_t)m;
    return m;
}


struct MatEntry {          // per material, in device memory
    float albedo, g, n, mua;
    // Henyey-Greenstein constants and the specular reflectance of the top face
    float omg2;            // 1 - g^2
    float opg2;            // 1 + g^2
    float inv2g;           // 1 / (2g), 0 when g == 0 (isotropic)
    float rsp_top;         // ((n - n_top)/(n + n_top))^2, 0 when the top face is index matched
};

struct McArgs {
    int nx, ny, nz;
    float dx, dy, dz;          // voxel edge, cm
    float cx, cy;              // beam centre, cm (grid coordinates)
    float radius;              // beam radius, cm
    int beam;
    float n_zlo, n_zhi;        // outside indices (<= 0: matched)
    const float* mut;          // interior mu_t, cm^-1
    const uint8_t* mat;        // interior material ids
    const MatEntry* mtab;
    double* tally;             // interior path length, cm
    unsigned long long* counters;   // TAM_CNT_*
    uint64_t p0, npk;
    uint32_t key0, key1, stream;
    int blocks_per_cu;         // 256-lane blocks per CU (lanes = packets in flight)
    int rng_batch;             // pending lanes of a wave that start its random phase
    // graph replays: packet count and Philox stream read from device memory
    // at launch (nullptr: npk and stream above)
    const int64_t* dyn_npk;
    const uint32_t* dyn_stream;
    int box_enable;            // LDS tally box (single-grid launches): 0 off, 1 16-bit fixed-point sums, 2 FP64 sums
    float box_scale;           // fixed-point box: units per cm of path (set by launch_mc)
    double box_quantum;        // cm per unit, 1 / box_scale
    // block-private tally box in LDS (box_n[0] == 0: none): voxels
    // [box_o, box_o + box_n) are summed per block and added to the tally once
    int box_o[3], box_n[3];
    // box depth from the last run's path: before clearing the tally,
    // launch_mc picks the z range whose planes held the most of the tally's
    // path length inside the box's footprint (over voxels still with
    // mu_t > 0), through box_psum (nz doubles) into box_geom (o[3], n[3]),
    // which the box kernel reads (nullptr: box_o / box_n as given)
    double* box_psum;
    int* box_geom;
    int clear_tally;           // launch_mc zeroes the tally (after the placement) before the transport
};

// counters[TAM_CNT_N] is the transport's packet queue head
enum { TAM_CNT_ABSORBED = 0, TAM_CNT_SPECULAR, TAM_CNT_ESC0, TAM_CNT_MISSED = TAM_CNT_ESC0 + 6,
       TAM_CNT_CROSS, TAM_CNT_INTER, TAM_CNT_KILLED, TAM_CNT_TALLY, TAM_CNT_BOX, TAM_CNT_N, TAM_CNT_SLOTS };

struct HeatArgs {
    int nx, ny, nz;
    long px, pxy;              // halo pitches
    const float* gx;           // face coefficient between c and c+1 (includes delt)
    const float* gy;           // ... c and c+px
    const float* gz;           // ... c and c+pxy
    const float* src;          // per-step source delt/(rho c) * jmean, K
    const float2* robin;       // per column (beta, gamma'): dT_ghost = beta*dT_top + gamma'
    // boiling (3dFD.f90:167-183)
    double* Q;                 // interior
    const double* jmean;       // interior, W/m^3
    const double* heatcap;     // halo
    double delt, volume_voxel, mass_voxel, qvapor;
    float dT_boil;             // 373 - t_ref
    double t_ref;
    // Arrhenius inside the step
    double* omega;             // interior
    double* thres;             // interior x3
    uint8_t* tmask;            // interior: bit k set when thres[k] != 0, TAM_TMASK_DUE
    const double* rhokap;      // halo
    double arr_A, arr_dE, arr_R, arr_Tlo, arr_Thi, thr0, thr1, thr2;
    double lnA;                // log(arr_A), for the in-step damage rate
    double thr_min;            // smallest threshold of the chain
    float dT_win_pad;          // extra margin of the Omega prefetch predicate, K
    // energy accounting
    double* energy;            // [0] boundary outflow, J
    double* eblk;              // per-block boundary outflow slots (FP32 stencil)
    int* flags;                // [0] negative temperature seen
    // class mode (see heat.hip)
    uint8_t* cls;              // halo: thermal class per cell (the step keeps TAM_CLASS_DUE)
    const float* glut;         // [3][16][16] face coefficient per (axis, minus class, plus class)
    const float2* robin_lut;   // [16] convective ghost (beta, gamma') per top-cell class
    // FP32 pre-tests (conservative) before the exact FP64 tests
    float dT_neg, dT_boil_test, dT_lo_test, dT_hi_test;
    // graph replays: this step's laser state and time after the step, and the
    // Arrhenius (delt, time), read from device memory (nullptr: by value)
    const float* dyn_lon;
    const double* dyn_tafter;
    const double* dyn_arr;
};

// one instance of a batched heat step: its arguments and its ping-pong pair
struct HeatBatchEntry {
    HeatArgs a;
    float* cur;                // input at even parity (output at odd)
    float* nxt;
};

#define TAM_NCLASS 16          // LUT classes; the last one is "removed" (no conduction)
#define TAM_CLASS_REMOVED (TAM_NCLASS - 1)
// Bit 7 of an interior class byte: all six neighbours have the same class,
// so the face coefficients are the class's own x / y / z entries.  Readers
// take the class as (byte & TAM_CLASS_MASK).
#define TAM_CLASS_HOM 0x80u
#define TAM_CLASS_MASK 0x0fu
// bit 6 of an interior class byte mirrors its mask's TAM_TMASK_DUE: the step
// kernels read the class byte of every voxel anyway, so the rare voxel with a
// threshold time due outside the Arrhenius window costs the others nothing
#define TAM_CLASS_DUE 0x40u

struct Heat64Args {            // FP64 reference-order step (heat64.hip)
    int nx, ny, nz;
    long px, pxy;
    const double* ax;          // a_f (no delt) between c and c+1
    const double* ay;
    const double* az;
    const double* kappa;       // halo (Robin ghost)
    const double* heatcap;     // halo
    const double* coeff;       // halo, delt/(rho c)
    const double* jmean;       // interior, W/m^3
    double* Q;
    double delt, volume_voxel, mass_voxel, qvapor;
    double h_top, t_amb, dz;
    double* omega; double* thres;
    uint8_t* tmask;
    double arr_A, arr_dE, arr_R, arr_Tlo, arr_Thi, thr0, thr1, thr2;
    double* energy;
    int* flags;
    const float* dyn_lon;      // graph replays, as in HeatArgs
    const double* dyn_tafter;
};

struct Coeff64Args {
    int nx, ny, nz;
    long px, pxy;
    double dscale, idx2, idy2, idz2;
    const double* kappa; const double* density; const double* heatcap;
    double* ax; double* ay; double* az;
    const double* rhokap;
    int removed_inert;
};

struct RebuildArgs {
    int nx, ny, nz;
    long px, pxy;
    int mode;                  // TAM_THERMAL_*
    double ablate_K, delt, mu_water, mu_protein, qvapor, t_ref;
    const float* dT;           // current temperature (halo), FP32 path
    const double* T64;         // current temperature (halo), FP64 path (nullptr otherwise)
    double* rhokap;            // halo
    double* kappa; double* density; double* heatcap; double* coeff;
    double* W; const double* Q;
    float* mut; uint8_t* mat;
    uint8_t* mask;             // interior scratch
    int air_id;                // material id given to removed voxels
    uint8_t* cls;              // halo thermal classes (class mode), nullptr otherwise
};

struct CoeffArgs {
    int nx, ny, nz;
    long px, pxy;
    double delt, dscale, idx2, idy2, idz2;
    const double* kappa; const double* density; const double* heatcap;
    float* gx; float* gy; float* gz;
    int robin; double h_top, t_amb, t_ref, dz;
    float2* robin_coef;
    const double* rhokap;      // removed voxels: interior cells with rhokap == 0
    int removed_inert;         // zero conductance across faces of removed voxels
};

struct tam_ctx {
    tam_params p;
    int dev;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::string err;
    long nvox, nhalo, px, pxy;
    int nmat, air_id;
    bool optics_set, thermal_set, src_valid;
    // device arrays
    float *dT0, *dT1, *gx, *gy, *gz, *src;
    float *cur, *nxt;           // current / next temperature buffer (ping-pong)
    bool fp64;                  // FP64 reference-order heat path
    double *T64a, *T64b, *cur64, *nxt64, *ax, *ay, *az;
    bool lut;                   // class mode active (FP32 path)
    uint8_t* cls;               // halo classes
    float* glut;                // [3][16][16]
    float2* robin_lut;          // [16]
    int tile, kz;               // stencil launch shape (autotuned)
    double* eblk;               // per-block energy slots
    long neblk;                 // slots: enough for the largest launch grid
    int mc_bpc;                 // transport launch: blocks per CU (autotuned)
    int mc_batch;               // transport launch: random-phase batch (autotuned)
    int mc_box;                 // transport launch: LDS tally box off / 16-bit fixed point / FP64 (0 / 1 / 2; on-off autotuned)
    float2* robin;
    double *kappa, *density, *heatcap, *coeff, *rhokap;
    float* mut;
    uint8_t *mat, *mask;
    MatEntry* mtab;
    double* mua_frac;           // per material mu_a / mu_t (FP64)
    double *tally, *jmean, *Q, *W, *omega, *thres;
    uint8_t* tmask;             // bit k: thres[k] != 0 (kept with every write of thres), TAM_TMASK_DUE
    // heat tuner scratch copies (allocated on the first tam_tune_heat, kept)
    double *tn_omega, *tn_thres, *tn_Q, *tn_eblk;
    uint8_t* tn_tmask;
    uint8_t* tn_cls;
    double* box_psum;           // LDS tally box placement: per-plane sums (nz) and the chosen
    int* box_geom;              // geometry (o[3], n[3])
    double* energy;             // [0] boundary outflow, [1..] reduction scratch
    int* flags;
    unsigned long long* counters;
    double time;                // heat time when no pulse state is passed
    tam_material host_tab[TAM_MAX_MATERIALS + 1];
    tam_timing timing;
    // captured coupling cycle (tam_cycle): one executable graph per starting
    // buffer of the temperature ping-pong (an odd step count swaps them)
    hipGraphExec_t cyc_exec[2];
    int cyc_swap[2];            // ping-pong swaps (mod 2) one replay of cyc_exec[p] makes
    long cyc_npk;               // capture key: packet capacity, steps, launch shapes, flags
    int cyc_loops, cyc_arr;
    uint64_t cyc_seed;
    char* cyc_dev;              // device parameter block of the replays
    size_t cyc_bytes;
    char* cyc_host;             // pinned staging ring of parameter blocks
    int cyc_slot;
    hipEvent_t cyc_ev[4];       // staging slot i is free once cyc_ev[i] completed
    bool cyc_ev_used[4];
    unsigned long long* cnt_acc;   // transport counters summed over replays
};

// parameter block of one graph replay (device layout, 64-byte header, then
// float lon[loops] padded to 8 bytes, then double t_after[loops])
struct CycleHdr {
    int64_t npk;               // packets this cycle (0: laser off, no transport)
    uint32_t stream;           // Philox stream id
    int32_t mc_on;             // 1: run transport + normalise
    double scale;              // normalisation scale
    double arr[2];             // Arrhenius delt, time (reference cadence)
    double pad[3];
};
static_assert(sizeof(CycleHdr) == 64, "cycle header layout");

// ---- launchers (each enqueues on `s`, returns hipError_t) --------------
hipError_t launch_mc(const McArgs& a, hipStream_t s);
// once per context, before any capture: lets the LDS-box kernel take the
// device's whole per-block LDS
hipError_t mc_prepare(int dev);
// the fixed-point LDS box's unit for a's voxel size (box_scale, box_quantum)
void mc_box_units(McArgs& a);
// all instances of a device table in one launch (see mc_batch_kernel)
hipError_t launch_mc_batch(const McArgs* dtab, int n, long total_npk, int blocks_per_cu, hipStream_t s);
// `count` FP64 atomic adds (whole waves) to pseudo-random slots of t[0, slots)
hipError_t launch_atomic_rate(double* t, long slots, long count, hipStream_t s);
hipError_t launch_heat_step(const HeatArgs& a, const float* Tin, float* Tout, float lon, double time_after,
                            int boil, int robin, int damage, int energy, int lut, int tile, int kz,
                            hipStream_t s);
int heat_num_tiles();
int heat_tile_dims(int tile, int* tx, int* ty);
// 2 for the two-step shapes (launch_heat_step2), 1 otherwise
int heat_tile_steps(int tile);
// two steps Tin -> Tout in one launch (two-step shapes; class mode without
// boiling); step 1 uses (lon0, time_after0), step 2 (lon1, time_after1), or
// a.dyn_lon[0..1] / a.dyn_tafter[0..1] when a.dyn_lon is set
hipError_t launch_heat_step2(const HeatArgs& a, const float* Tin, float* Tout, float lon0, float lon1,
                             double time_after0, double time_after1, int robin, int damage, int energy, int tile,
                             int kz, hipStream_t s);
// upper bound of the block count of any heat launch shape (energy slots)
long heat_max_blocks(int nx, int ny, int nz);
// one step of every instance of a table (direct shapes 3-5 only); parity 0
// reads cur and writes nxt, parity 1 the reverse
// two steps of every instance of a table (two-step shapes 6-8, class mode)
hipError_t launch_heat_batch2(const HeatBatchEntry* tab, int n, const int* tile_start, int total_tiles, float lon0,
                              float lon1, double time_after0, double time_after1, int robin, int damage, int energy,
                              int tile, int kz, int parity, hipStream_t s);
hipError_t launch_heat_batch(const HeatBatchEntry* tab, int n, const int* tile_start, int total_tiles, float lon,
                             double time_after, int boil, int robin, int damage, int energy, int lut, int tile,
                             int kz, int parity, hipStream_t s);
const char* heat_tile_name(int tile);
hipError_t launch_heat64_step(const Heat64Args& a, const double* Tin, double* Tout, double lon,
                              double time_after, int boil, int robin, int damage, int energy, hipStream_t s);
hipError_t launch_coeffs64(const Coeff64Args& a, hipStream_t s);
hipError_t launch_acc_u64(unsigned long long* acc, const unsigned long long* src, long n, hipStream_t s);
hipError_t launch_arrhenius(const HeatArgs& a, const float* T, const double* T64, double delt, double time,
                            hipStream_t s);
hipError_t launch_removal_mask(const RebuildArgs& a, hipStream_t s);
hipError_t launch_apply_props(const RebuildArgs& a, hipStream_t s);
hipError_t launch_water_update(const RebuildArgs& a, hipStream_t s);
hipError_t launch_coeffs(const CoeffArgs& a, hipStream_t s);
// TAM_CLASS_HOM bits of the interior class bytes from the current classes
hipError_t launch_class_flags(uint8_t* cls, int nx, int ny, int nz, long px, long pxy, hipStream_t s);
// dyn (graph replays, else nullptr): scale from dyn->scale; nothing is
// written when dyn->mc_on == 0 (laser off: jmean keeps its last value)
hipError_t launch_normalise(int nx, int ny, int nz, long px, long pxy, const double* tally,
                            const double* rhokap, const uint8_t* mat, const double* mua_tab,
                            const CycleHdr* dyn, double scale, const double* coeff, double* jmean,
                            float* src, hipStream_t s);
hipError_t launch_fill_f64(double* p, long n, double v, hipStream_t s);
// error of a call that has no context (read back by tam_last_error(NULL))
int tam_fail_noctx(int code, const std::string& msg);
// tmask from the threshold times and Omega: bit k = (thres[k * nvox + v] != 0),
// TAM_TMASK_DUE where a threshold with an unset time is already reached
hipError_t launch_tmask(const double* thres, const double* omega, double thr0, double thr1, double thr2,
                        uint8_t* tmask, long nvox, hipStream_t s);
// TAM_CLASS_DUE bits of the interior class bytes from the masks' TAM_TMASK_DUE
hipError_t launch_class_due(uint8_t* cls, const uint8_t* tmask, int nx, int ny, int nz, long px, long pxy,
                            hipStream_t s);
hipError_t launch_fill_f32(float* p, long n, float v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, long n, hipStream_t s);
hipError_t launch_copy(const void* in, void* out, long bytes, hipStream_t s);
// k^3 block sums of an interior field into out[(nz/k)(ny/k)(nx/k)] (k divides every edge)
hipError_t launch_block_sums(const double* f, int nx, int ny, int nz, int k, double* out, hipStream_t s);
hipError_t launch_energy(int nx, int ny, int nz, long px, long pxy, const float* dT, const double* T64,
                         double t_ref, const double* heatcap, double mass_voxel, double* out, hipStream_t s);
