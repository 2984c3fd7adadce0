/*
 * tam.h -- C-ABI of the MI355X Tissue-Ablation-MC hot path.
 *
 * Plain C, no torch or HIP types.  The Fortran host binds these through
 * ISO_C_BINDING (tissue-ablation-mc_amd/fortran/tam_bind.f90); Python binds
 * them through ctypes (tissue-ablation-mc_amd/tam_amd/lib.py).
 *
 * Granularity is one call per photon batch / heat call / cycle, not per
 * photon: the reference's per-packet subroutines (sourcephCO2, tauint1) are
 * replaced by tam_mc_run, which runs the whole photon loop on the GPU.
 *
 * Conventions (same as the reference's arrays, src/iarray.f90:8-12 and
 * src/3dFD.f90:10): FP64, column-major (x fastest).  "Halo" arrays are
 * (nx+2)*(ny+2)*(nz+2) with the Fortran lower bound 0 (src/subs.f90:66,72);
 * "interior" arrays are nx*ny*nz.  Every function returns TAM_OK (0) or a
 * negative TAM_ERR_* code; tam_last_error() gives the message.  A context is
 * owned by one host thread and is not thread safe.
 */
#ifndef TAM_H
#define TAM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TAM_ABI_VERSION 1

enum {
    TAM_OK = 0,
    TAM_ERR_ARG = -1,            /* bad argument / shape */
    TAM_ERR_HIP = -2,            /* HIP runtime error */
    TAM_ERR_NEGATIVE_TEMP = -3,  /* src/3dFD.f90:179-182 (mpi_abort on T < 0) */
    TAM_ERR_STATE = -4,          /* call out of order (e.g. heat before thermal state) */
    TAM_ERR_NOMEM = -5,
    TAM_ERR_NODEVICE = -6
};

/* beam profiles (sourceph.f90:28-31 is the top-hat disk) */
enum { TAM_BEAM_TOPHAT = 0, TAM_BEAM_GAUSSIAN = 1 };
/* thermal property models for the rebuild */
enum { TAM_THERMAL_WATER = 0,   /* reference: water-content fits, 3dFD.f90:328-345 */
       TAM_THERMAL_TABLE = 1 }; /* fixed per-material properties */
/* faces, in the order of tam_mc_stats.escaped */
enum { TAM_FACE_XLO = 0, TAM_FACE_XHI, TAM_FACE_YLO, TAM_FACE_YHI, TAM_FACE_ZLO, TAM_FACE_ZHI };
/* fields for tam_download / tam_upload */
enum {
    TAM_FIELD_TEMP = 0,      /* halo, K (heat_sim_3D's temp) */
    TAM_FIELD_JMEAN = 1,     /* interior, W/m^3 after tam_normalise (jmeanGLOBAL) */
    TAM_FIELD_TALLY = 2,     /* interior, path length per voxel in cm (raw MC tally) */
    TAM_FIELD_RHOKAP = 3,    /* halo, cm^-1 (iarray rhokap) */
    TAM_FIELD_TISSUE = 4,    /* interior, Arrhenius integral Omega (iarray tissue) */
    TAM_FIELD_THRESTIME = 5, /* interior x 3, s (iarray ThresTime) */
    TAM_FIELD_WATER = 6,     /* interior (Heat WaterContent) */
    TAM_FIELD_Q = 7,         /* interior, J (Heat Q) */
    TAM_FIELD_KAPPA = 8,     /* halo, W/m/K */
    TAM_FIELD_DENSITY = 9,   /* halo, kg/m^3 */
    TAM_FIELD_HEATCAP = 10,  /* halo, J/kg/K */
    TAM_FIELD_COEFF = 11,    /* halo, K m^3/J (delt/(rho c)) */
    TAM_FIELD_MAT = 12       /* interior material ids, as doubles */
};

typedef struct tam_ctx tam_ctx;

typedef struct {
    double mua, mus, g, n;       /* cm^-1, cm^-1, HG anisotropy, refractive index */
} tam_material;

/* Everything the reference keeps in module globals for this path
 * (constants.f90:12-14, opt_prop.f90:5, 3dFD.f90:5-12). */
typedef struct {
    int32_t nx, ny, nz;          /* voxels (constants.f90:12 nxg,nyg,nzg) */
    double  xmax, ymax, zmax;    /* half extents, cm: the grid spans [0, 2*xmax] (gridset.f90:23-31) */
    /* heat (initThermalCoeff, 3dFD.f90:249-291) */
    double  dx_m, dy_m, dz_m;    /* FD node spacing, m */
    double  delt;                /* time step, s */
    double  dscale;              /* 0.5 reproduces the factor of 3dFD.f90:128-129; 1.0 = textbook */
    double  volume_voxel;        /* m^3, MC voxel volume (3dFD.f90:289) */
    double  mass_voxel;          /* kg (3dFD.f90:290) */
    double  qvapor;              /* J (3dFD.f90:291) */
    double  t_ref;               /* K: the device holds T - t_ref in FP32 */
    int32_t boil;                /* boiling branch (3dFD.f90:169-183) on/off */
    int32_t robin_top;           /* convective top face (extension) */
    double  h_top, t_amb;        /* W/m^2/K, K */
    /* rebuild (setupThermalCoeff, 3dFD.f90:312-361) */
    int32_t thermal_mode;        /* TAM_THERMAL_* */
    double  mu_water, mu_protein;/* ch_opt.f90:171-172 */
    double  ablate_K;            /* ablateTemp + 273 */
    /* Arrhenius (3dFD.f90:439-446) */
    double  arr_A, arr_dE, arr_R, arr_Tlo, arr_Thi, arr_thr[3];
    int32_t damage_every_step;   /* 1: damage integrated inside every heat step; 0: tam_arrhenius only */
    /* photon source and boundaries (sourceph.f90:7-49) */
    int32_t beam;                /* TAM_BEAM_* */
    double  beam_radius;         /* cm: top-hat radius or Gaussian 1/e^2 radius */
    double  n_ext_zlo, n_ext_zhi;/* index outside the bottom/top faces; <= 0: index matched */
    int32_t device;              /* HIP device ordinal */
    int32_t heat_fp64;           /* 1: FP64 temperatures in the reference's operation order
                                    (needed for the boiling clamp); 0: FP32 T - t_ref */
    int32_t removed_inert;       /* 1: no conduction across faces of removed (ablated) voxels.
                                    The reference gives them air properties (3dFD.f90:350-357),
                                    whose diffusivity is ~140x tissue's: explicit steps at the
                                    tissue's stability limit then diverge once two removed
                                    voxels touch.  0 keeps the reference rule. */
} tam_params;

typedef struct {
    double packets;
    double absorbed, specular, escaped[6], missed;   /* packet counts by fate */
    double crossings, interactions;                  /* work counters */
    double kernel_ms;                                /* device time of the transport kernel */
    double killed;                                   /* abandoned after 2^24 events (exit guarantee) */
    double tally_adds;                               /* FP64 path deposits into the tally, one per
                                                        voxel visit with mu_t > 0 (device atomics,
                                                        or LDS atomics inside the tally box) */
    double box_adds;                                 /* of these, the ones into the LDS tally box */
} tam_mc_stats;

/* Pulse state (3dFD.f90:7-12, :199-214); advanced by tam_heat_run. */
typedef struct {
    double time, pulseCount, repetitionCount, repetitionRate_1, realPulseLength, laserOn;
    int32_t laser_flag, pulsesDone, pulsesToDo, use_pulse;  /* use_pulse 0: laserOn fixed */
} tam_pulse;

typedef struct {
    double mc_ms, normalise_ms, heat_ms, arrhenius_ms, rebuild_ms;   /* last call of each */
    double heat_steps;           /* steps in the last tam_heat_run */
    double heat_step_ms;         /* mean device time per step in the last tam_heat_run (a two-step
                                    launch counts as two steps) */
    double class_mode;           /* 1: stencil runs on thermal classes + LDS coefficient table */
    double tile, kz;             /* stencil launch shape in use */
    double cycle_ms;             /* device time of the last timed tam_cycle (transfer + graph) */
} tam_timing;

/* --- lifetime -------------------------------------------------------- */
/* Replaces alloc_array/zarray (subs.f90:27-77) and the device side of
 * initThermalCoeff (3dFD.f90:233-309).  Allocates every device array. */
int  tam_create(const tam_params* p, tam_ctx** out);
void tam_destroy(tam_ctx* ctx);
const char* tam_last_error(const tam_ctx* ctx);
int  tam_abi_version(void);
/* sizeof of the ABI structs, for FFI layout checks: 0 tam_params,
 * 1 tam_material, 2 tam_mc_stats, 3 tam_pulse, 4 tam_timing; -1 otherwise. */
int64_t tam_struct_size(int32_t which);

/* One step of the laser pulse state machine (3dFD.f90:199-214): turn the
 * laser off after realPulseLength, on again after repetitionRate_1 while
 * pulses remain, then advance pulseCount, repetitionCount and time by delt.
 * Host-only; tam_heat_run applies it after every step. */
void tam_pulse_advance(tam_pulse* pulse, double delt);

/* --- optics: init_opt1 + gridset (ch_opt.f90:7-25, gridset.f90:6-46) ----
 * mat: interior material ids (< nmat <= 255).  rhokap: optional halo array
 * of mu_t in cm^-1 (the reference's rhokap, used in TAM_THERMAL_WATER mode);
 * NULL derives mu_t = mua + mus from the table.  Material ids index the
 * table for albedo, g and n. */
int tam_set_optics(tam_ctx* ctx, const uint8_t* mat, const tam_material* table, int32_t nmat,
                   const double* rhokap);

/* --- thermal state: temp/kappa/density/heatcap halo arrays (3dFD.f90:10,
 * mcpolar.f90:123-130); Q, W interior or NULL (zero / 0.75). */
int tam_set_thermal(tam_ctx* ctx, const double* temp, const double* kappa, const double* density,
                    const double* heatcap, const double* Q, const double* W);

/* --- photon loop: sourcephCO2 + tauint1 (+ stokes) for npackets packets,
 * replacing mcpolar.f90:151-170.  Tallies path length per voxel (FP64) into
 * the device tally; accumulate=0 clears it first.  Voxels with mu_t = 0
 * (ablated: mu_a = 0, so jmean is 0 there as in the reference) are crossed
 * without a tally.  Packet p draws from the Philox4x32-10 stream key=seed,
 * counter=(p, event, stream). */
int tam_mc_run(tam_ctx* ctx, int64_t npackets, uint64_t seed, uint32_t stream, int32_t accumulate,
               tam_mc_stats* stats);

/* --- transport launch shape (no reference counterpart: the reference runs
 * one packet at a time per rank).  blocks_per_cu 256-lane blocks per CU set
 * the packets in flight; rng_batch (1..64) is how many lanes of a 64-lane
 * wave must wait for random numbers before the wave draws them together.
 * tam_tune_mc times candidate pairs on a sample of `sample` packets and keeps
 * the fastest (the tally is cleared afterwards), with the LDS tally box on
 * and off; near-ties go to the CU-wide box (4 blocks per CU, within 10%) and
 * to a random-phase batch of 32 (within 3%), the fastest once a crater
 * forms.  None of them changes results beyond FP64 summation order: packet p
 * always draws from its own stream.
 * tam_set_mc_box: each block sums the path it lays in a cube of voxels on
 * the beam axis in LDS (the CU's LDS shared by the blocks per CU, at most the
 * device's per-block limit) and adds the box to the FP64 tally once at the
 * end, instead of one device atomic per voxel visit.  enable: 0 = off;
 * 1 = on with fixed-point sums in 16-bit integers, two voxels per word (the
 * default: four times the voxels of FP64; each deposit rounded to 2^-13 of
 * dx + dy + dz, and the sums independent of the order of the adds; a slot
 * that wraps hands 2^16 units to the tally, so nothing is lost); 2 = on with
 * FP64 sums.  With the box on, blocks_per_cu 4 and 8 run as 1 and 2 blocks of
 * 1024 lanes per CU: the same lanes, one box per 1024 (42^2 x 46 voxels fixed
 * point, 27^3 FP64, at 1 per CU on MI355X).  Each transport launch places the cube's z range where the
 * tally it is about to clear (the last run's path), over voxels with
 * mu_t > 0, sums highest inside the cube's x-y footprint (ties: the higher
 * range); with no path yet the cube starts at the top face.
 * tam_get_mc_box: the box of the last launch that placed one (no reference
 * counterpart; for tests and tuning): geom = {x0, y0, z0, nx, ny, nz}, 0-based
 * interior voxel indices; waits for the context's stream. */
int tam_set_mc_launch(tam_ctx* ctx, int32_t blocks_per_cu, int32_t rng_batch);
int tam_set_mc_box(tam_ctx* ctx, int32_t enable);
int tam_get_mc_box(tam_ctx* ctx, int32_t* geom);
int tam_tune_mc(tam_ctx* ctx, int64_t sample, int32_t* blocks_per_cu, int32_t* rng_batch,
                double* packets_per_s);

/* --- normalisation: jmeanGLOBAL = mu_a * L * scale (mcpolar.f90:173-174);
 * scale = (getPwr()/81)/(nphotons*volumeVoxel) in the reference.  Also
 * forms the per-step heat source delt/(rho c) * jmean for the stencil. */
int tam_normalise(tam_ctx* ctx, double scale);

/* --- heat_sim_3D (3dFD.f90:21-230): `loops` explicit steps.  pulse (may be
 * NULL: laser on) is advanced exactly as 3dFD.f90:199-214.  With
 * damage_every_step the Arrhenius integral runs inside each step. */
int tam_heat_run(tam_ctx* ctx, int32_t loops, tam_pulse* pulse);

/* --- stencil launch shape (no reference counterpart: the reference's z-slab
 * split is fixed by the rank count, 3dFD.f90:45-72).  tile: 0 = 64x4,
 * 1 = 32x8, 2 = 64x8 (x by y) voxels per block with x-y planes staged in
 * LDS; 3, 4, 5 = the same tiles loading neighbours directly (no LDS, no
 * barriers); 6, 7, 8 = two steps per launch on 32x8, 32x16, 64x8 footprints
 * (the inner 30x6, 30x14, 62x6 columns owned; the intermediate field stays
 * on chip), used in class mode without boiling, with an odd last step and
 * every other step on direct 32x8 (6, 7) or direct 64x8 (8); kz: z planes
 * per block.  Every shape gives bit-identical fields, damage and threshold
 * times.
 * tam_tune_heat times every candidate (in class mode the direct and two-step
 * shapes) x kz on the current field and damage state (results go to scratch
 * copies, the state is not advanced) and keeps the fastest per step. */
int tam_set_heat_launch(tam_ctx* ctx, int32_t tile, int32_t kz);
int tam_tune_heat(tam_ctx* ctx, int32_t reps, int32_t* tile, int32_t* kz, double* step_ms);

/* --- Arrhenius (3dFD.f90:424-466), once, with the given delt and time. */
int tam_arrhenius(tam_ctx* ctx, double delt, double time);

/* --- setupThermalCoeff (3dFD.f90:312-361): ablation mask, property and
 * optics rebuild ("ch_opt rebuild"), then the stencil coefficients. */
int tam_rebuild(tam_ctx* ctx);

/* --- one whole coupling cycle: the body of mcpolar's time loop
 * (mcpolar.f90:148-186) as one replay of a HIP graph.  In order:
 *   transport of npackets (0: the laser is off, no transport, jmean keeps
 *   its last value as the reference's jmeanGLOBAL does) on Philox stream
 *   `stream` of key `seed`; normalisation with `scale`; `loops` heat steps
 *   with the pulse state advanced per step as in tam_heat_run; Arrhenius
 *   at the reference cadence with (arr_delt, time after the steps) when
 *   arr_delt > 0; the rebuild of tam_rebuild.
 * The graph is captured at the first call and replayed while loops, the
 * packet capacity, seed and cadence stay the same (any tam_set_* /
 * tam_upload / launch-shape call drops it).  Per-cycle values go to the
 * device through a pinned staging ring, so with stats == NULL the call
 * returns without waiting for the GPU.  With stats it waits, reports the
 * transport counters summed since the previous timed call, checks for
 * negative temperatures (TAM_ERR_NEGATIVE_TEMP) and sets
 * tam_timing.cycle_ms. */
int tam_cycle(tam_ctx* ctx, int64_t npackets, uint64_t seed, uint32_t stream, double scale, int32_t loops,
              double arr_delt, tam_pulse* pulse, tam_mc_stats* stats);

/* --- data movement ---------------------------------------------------- */
/* Sums of the transport tally over k^3 blocks of voxels (k divides every
 * edge), x-block fastest: a cheap device-side summary for checks that
 * compare two transport runs (no reference counterpart). */
int tam_tally_blocks(tam_ctx* ctx, int32_t k, double* sums);
/* tam_tally_blocks with no device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The sums go through a fixed
 * device buffer of TAM_BLOCKSUM_SCRATCH_MAX doubles (2 MiB) that each GPU
 * the library runs on holds until the process ends; more blocks than that
 * is TAM_ERR_ARG.  Every call on one GPU shares that GPU's buffer: calls
 * that overlap on one GPU (from several threads or contexts) race. */
#define TAM_BLOCKSUM_SCRATCH_MAX 262144
int tam_tally_blocks_noalloc(tam_ctx* ctx, int32_t k, double* sums);
int tam_download(tam_ctx* ctx, int32_t field, double* out);
int tam_upload(tam_ctx* ctx, int32_t field, const double* in);

/* --- diagnostics ------------------------------------------------------ */
/* FP64 heat content sum(heatcap * mass_voxel * (T - t_ref)) over the interior, J,
 * and the energy that left through the boundary faces since tam_set_thermal. */
int tam_energy(tam_ctx* ctx, double* heat_content, double* boundary_out);
int tam_get_timing(tam_ctx* ctx, tam_timing* out);
/* Device-to-device streaming copy of `bytes` (rounded down to 16), `reps`
 * times; *gbps = (read + written bytes) / mean kernel time.  A measured
 * bandwidth reference for the stencil's roofline (no reference counterpart). */
int tam_copy_bandwidth(tam_ctx* ctx, int64_t bytes, int32_t reps, double* gbps);
/* FP64 atomic adds without return, `count` of them (rounded to whole waves)
 * to random addresses over a scratch array of `slots` doubles (at least
 * 64); *per_s = adds / kernel time.  The measured ceiling of the transport's
 * tally (no reference counterpart). */
int tam_atomic_rate(tam_ctx* ctx, int64_t count, int64_t slots, double* per_s);
/* Wait for the context's stream; TAM_ERR_NEGATIVE_TEMP if any heat step
 * since tam_set_thermal saw T < 0 (the reference's mpi_abort,
 * 3dFD.f90:179-182), which asynchronous tam_cycle calls cannot report. */
int tam_synchronize(tam_ctx* ctx);

/* --- batches of independent instances (no reference counterpart: the
 * reference runs one tissue per job).  A batch refers to existing contexts
 * on one device (it does not own them); every instance keeps its own
 * arrays.  Instances run the FP32 step without boiling or in-step damage,
 * with one convective-top setting and one coefficient mode.
 * tam_batch_mc_run: the transport of all instances in one launch (instance
 *   i: npackets[i] packets on Philox stream stream0 + i of key seed, as
 *   tam_mc_run would give); stats[i] per instance when stats is not NULL.
 *   Instances with the LDS tally box on (tam_set_mc_box, either kind) get a
 *   cube under their top face on the beam axis per wave (a quarter of the
 *   block's LDS, 16-bit fixed-point sums), not placed from the last run's
 *   path.
 * tam_batch_heat_run: `loops` steps of every instance, one launch per step
 *   or, with a two-step shape in class mode, per pair of steps (laser on;
 *   the source from each instance's last tam_normalise).
 * tam_batch_tune: times transport shapes (blocks per CU x random-phase
 *   batch) on sample_packets[i] per instance and stencil shapes (direct
 *   tiles 3-5 and two-step tiles 6-8 x z depths) over heat_reps steps
 *   without advancing the state, and keeps the fastest.
 * tam_batch_set_launch: tile 3..8 as tam_set_heat_launch. */
typedef struct tam_batch tam_batch;
int  tam_batch_create(tam_ctx* const* ctxs, int32_t n, tam_batch** out);
void tam_batch_destroy(tam_batch* b);
const char* tam_batch_last_error(const tam_batch* b);
int  tam_batch_mc_run(tam_batch* b, const int64_t* npackets, uint64_t seed, uint32_t stream0, int32_t accumulate,
                      tam_mc_stats* stats);
int  tam_batch_heat_run(tam_batch* b, int32_t loops);
int  tam_batch_set_launch(tam_batch* b, int32_t tile, int32_t kz, int32_t blocks_per_cu, int32_t rng_batch);
int  tam_batch_tune(tam_batch* b, const int64_t* sample_packets, int32_t heat_reps, int32_t* tile, int32_t* kz,
                    int32_t* blocks_per_cu, int32_t* rng_batch);
/* device time of the last batched launch sequence, ms */
double tam_batch_last_ms(const tam_batch* b);

/* --- spectral sampling helpers (ch_opt.f90:28-152; compiled but never
 * called in the reference, which runs one wavelength).  A spectrum is the
 * reference's table array(len, 2), column-major: table[i] the wavelength of
 * row i+1 (strictly ascending), table[len + i] its value.  No context:
 * errors are read with tam_last_error(NULL).
 * tam_spectrum_cdf replaces mk_cdf (ch_opt.f90:138-152): cdf[j-1] = the
 *   trapezoid area of rows 1..j+1 over the whole area, j = 1..len-1;
 *   cdf[len-1] is divided like the rest (the reference leaves that entry to
 *   its caller).  Host arithmetic, the reference's operation order.
 * tam_spectrum_sample replaces n calls of sample (ch_opt.f90:28-45, with
 *   search_1D :97-116 and lin_inter_1D :63-76) on `device`: draw i takes
 *   its uniform from Philox4x32-10 (key = seed, counter = (i, 0, stream),
 *   first word, ((x >> 9) + 0.5) * 2^-23) in place of ran2, and is inverted
 *   through the CDF in FP64 as the reference does (len >= 3).
 * tam_spectrum_interp replaces search_2D + lin_inter_2D (ch_opt.f90:79-135):
 *   the table's value at each of n wavelengths, on `device`. */
int tam_spectrum_cdf(const double* table, int32_t len, double* cdf);
int tam_spectrum_sample(int32_t device, const double* table, const double* cdf, int32_t len, int64_t n,
                        uint64_t seed, uint32_t stream, double* wave);
int tam_spectrum_interp(int32_t device, const double* table, int32_t len, const double* x, int64_t n, double* y);

#ifdef __cplusplus
}
#endif
#endif
