/*
 * tam_oracle.h -- FP64 CPU restatement of the Tissue-Ablation-MC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * Fortran host, the Python wrapper) links, loads or calls this code.  It is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, always as the checker / the timed CPU baseline, never as the thing
 * measured or shipped.
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arrays follow the reference layout:
 * column-major (x fastest); halo arrays have one extra plane on every side
 * (lower bound 0 in the Fortran, src/subs.f90:66,72 and src/3dFD.f90:254-258).
 *
 * Pinning of each function (how a plausible mistake is caught) is listed in
 * tests/test_oracle.py and tests/test_oracle_pins.py:
 *   - or_ran2, or_ref_mc (sourcephCO2 + tauint1), or_stokes and the
 *     thermal-property fits are pinned bit-for-bit against the reference
 *     Fortran compiled from its own sources (oracle/Makefile target `ref`).
 *   - or_philox4x32_10 is pinned by the Random123 known-answer vectors and by
 *     rocRAND's Philox4x32-10 (a different implementation).
 *   - or_mc_run (extended physics) is pinned by Beer-Lambert, the HG moment
 *     <cos theta> = g, the beam radial CDFs, A+R+T=1, normal-incidence Fresnel,
 *     the multiple-reflection closed form of a pure-absorber slab between two
 *     mismatched faces (the internal Fresnel reflection of the walk), and
 *     statistical agreement with the reference-formula variant.
 *   - or_heat_steps is pinned by the discrete Fourier-mode decay factor,
 *     a steady linear profile, the steady profile of a convective (Robin) top
 *     face, the series-resistance steady state of random kappa/rho/c, the
 *     reciprocity of the update with random properties, and the exact heat
 *     balance of an insulated box with random conductivity.
 *   - or_arrhenius by the constant-temperature closed form; or_rebuild by a
 *     hand-computed crater and by brute force over tiny grids.
 *   - the spectral helpers (or_mk_cdf .. or_sample) are pinned bit-for-bit
 *     against the reference's ch_opt module compiled from its sources, and by
 *     closed forms (a uniform spectrum samples x1 + u (xK - x1); a linear
 *     density samples its nodes at u = CDF(node)).
 */
#ifndef TAM_ORACLE_H
#define TAM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- RNGs -- */

/* Numerical-Recipes ran2 (src/ran2.f:1-33).  The Fortran keeps its shuffle
 * table in SAVE variables; here the state is explicit so threads can each own
 * one, but the arithmetic and the first-call initialisation are the same. */
typedef struct {
    int32_t idum;
    int32_t idum2;
    int32_t iy;
    int32_t iv[32];
} or_ran2_t;

void   or_ran2_init(or_ran2_t* s, int32_t iseed);   /* iseed <= 0 as in mcpolar.f90:97-98 */
double or_ran2(or_ran2_t* s);

/* Philox4x32-10 (Salmon et al., SC'11). */
void   or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 23-bit uniform in (0,1): ((x >> 9) + 0.5) * 2^-23, exact in FP32 -- identical on the GPU. */
double or_u01(uint32_t x);

/* ------------------------------------------------ reference MC (mu_s = 0) */

/* Geometry of the reference grid: faces xface(i) = (i-1)*2*xmax/nxg
 * (src/gridset.f90:23-31); coordinates run over [0, 2*xmax]. */
typedef struct {
    int    nxg, nyg, nzg;
    double xmax, ymax, zmax;        /* half extents, cm */
} or_grid_t;

/* sourcephCO2 + tauint1 for nphotons packets with one ran2 stream
 * (mcpolar.f90:151-170, sourceph.f90:7-49, inttau2.f90:7-239).
 * rhokap: (nxg+2)(nyg+2)(nzg+2) halo array, cm^-1.  jmean accumulates
 * dcell*rhokap per voxel (nxg*nyg*nzg, col-major).  The ran2 state is
 * advanced exactly as in the reference.  Returns the number of escapes. */
long or_ref_mc(const or_grid_t* g, const double* rhokap, long nphotons,
               or_ran2_t* rng, double* jmean);

/* stokes (src/stokes.f90:6-153): HG scattering of the direction held in
 * d[0..4] = (cost, sint, cosp, sinp, phi); writes n[0..2] = (nxp,nyp,nzp).
 * hgg = g, g2 = g*g. Uses two ran2 draws (or one/two for g == 0). */
void or_stokes(double d[5], double n[3], double hgg, or_ran2_t* rng);

/* --------------------------------------------- extended-physics MC ------ */

enum { OR_BEAM_TOPHAT = 0, OR_BEAM_GAUSSIAN = 1 };
enum { OR_RNG_PHILOX = 0, OR_RNG_RAN2 = 1 };
enum { OR_SCATTER_ROTATION = 0, OR_SCATTER_STOKES = 1 };
/* exit faces, in the order used by the stats counters */
enum { OR_FACE_XLO = 0, OR_FACE_XHI, OR_FACE_YLO, OR_FACE_YHI, OR_FACE_ZLO, OR_FACE_ZHI };

typedef struct {
    double mua, mus, g, n;          /* cm^-1, cm^-1, -, - */
} or_material_t;

typedef struct {
    or_grid_t        grid;
    const double*    mut;           /* nxg*nyg*nzg, cm^-1: mu_t per voxel ("rhokap") */
    const uint8_t*   mat;           /* nxg*nyg*nzg material ids (albedo, g, n) */
    const or_material_t* table;     /* material table */
    int              nmat;
    double           n_ext[6];      /* refractive index outside each face; <= 0: index matched */
    int              beam;          /* OR_BEAM_* */
    double           beam_radius;   /* top-hat radius or Gaussian 1/e^2 radius, cm */
    int              rng;           /* OR_RNG_* */
    int              scatter;       /* OR_SCATTER_* */
    uint64_t         seed;          /* Philox key */
    uint32_t         stream;        /* Philox counter word 3 (instance id) */
    int32_t          ran2_seed;     /* ran2 seed for OR_RNG_RAN2 (<= 0) */
} or_mc_t;

typedef struct {
    double absorbed;                /* packets ended by absorption */
    double specular;                /* packets reflected at launch */
    double escaped[6];              /* packets escaped through each face */
    double missed;                  /* launched outside the grid */
    double crossings;               /* voxel crossings */
    double interactions;            /* interaction events */
} or_mc_stats_t;

/* Runs packets [p0, p0+npk).  tally_L (nxg*nyg*nzg) accumulates path length
 * per voxel in cm (the absorbed fraction per voxel is mu_a * L); voxels with
 * mu_t = 0 (ablated, mu_a = 0) are crossed without a tally.  With
 * OR_RNG_PHILOX each packet has its own counter stream, so the result does
 * not depend on the thread count; nthreads <= 0 uses OpenMP's default. */
int or_mc_run(const or_mc_t* m, uint64_t p0, uint64_t npk, double* tally_L,
              or_mc_stats_t* stats, int nthreads);

/* Fresnel reflectance, unpolarised (textbook), n_i -> n_t, cos of incidence. */
double or_fresnel(double n_i, double n_t, double cos_i);

/* The samplers or_mc_run uses, exposed for distribution tests:
 * HG deflection cosine for uniform xi; beam launch offset (x, y) from the
 * beam axis for uniforms (u0 radius, u1 azimuth). */
double or_hg_cos(double g, double xi);
void   or_beam_sample(int beam, double radius, double u0, double u1, double xy[2]);

/* ------------------------------------------------------------ heat ------ */

typedef struct {
    int     nx, ny, nz;             /* interior points (reference: numpoints each) */
    double  dx, dy, dz;             /* m (3dFD.f90:249-251) */
    double  delt;                   /* s */
    double  dscale;                 /* 0.5 reproduces the factor at 3dFD.f90:128-129 */
    double* temp;                   /* halo array, K */
    const double* kappa;            /* halo arrays, SI */
    const double* density;
    const double* heatcap;
    const double* coeff;            /* delt/(rho c) */
    double* Q;                      /* nx*ny*nz, J (boiling energy) */
    const double* jmean;            /* nx*ny*nz, W/m^3 */
    double  volumeVoxel, massVoxel, QVapor;
    int     boil;                   /* boiling branch, 3dFD.f90:169-183 */
    int     robin_top;              /* extension: convective top face */
    double  h_top, T_amb;           /* W/m^2/K, K */
    double  flux_out;               /* accumulates energy through the boundary, J (extension) */
    const double* rhokap;           /* halo; with removed_inert, interior cells with rhokap == 0 */
    int     removed_inert;          /* extension: no conduction across faces of removed voxels */
} or_heat_t;

/* `loops` explicit steps of heat_sim_3D (3dFD.f90:113-215, single rank).
 * laser_on[p] is the laserOn factor used in step p (the host evaluates the
 * pulse state machine 3dFD.f90:199-214).  Returns -1 if a temperature went
 * negative (the mpi_abort at 3dFD.f90:179-182). */
int or_heat_steps(or_heat_t* h, int loops, const double* laser_on);

/* Arrhenius (3dFD.f90:424-466), extended with A, dE, R and the window. */
typedef struct {
    double A, dE, R;
    double Tlo, Thi;                /* 316, 373 in the reference */
    double thr[3];                  /* 0.53, 1, 1e4 */
} or_arrhenius_t;
void or_arrhenius(const or_arrhenius_t* a, int nx, int ny, int nz, const double* temp,
                  const double* rhokap, double delt, double time, double* tissue,
                  double* threstime /* nx*ny*nz*3 */);

/* setupThermalCoeff (3dFD.f90:312-361), sequential in-place sweep.
 * mode 0: reference (water-content model, rhokap = W*mu_water + mu_protein).
 * mode 1: material table -- surviving tissue keeps its properties and mu_t,
 *         removed voxels get the same air properties as the reference.
 * Arrays are halo arrays except Q, W (interior).  ablate_K = ablateTemp+273. */
typedef struct {
    int     nx, ny, nz;
    int     mode;
    double  ablate_K;
    double  delt;
    double  mu_water, mu_protein;   /* ch_opt.f90:171-172 */
    double  QVapor;
    double* rhokap; double* density; double* heatcap; double* kappa;
    double* coeff;  double* alpha;
    double* W;      const double* Q;
    const double* temp;
} or_rebuild_t;
void or_rebuild(or_rebuild_t* r);

/* Single-pass parallel form of the rebuild's removal rule:
 * Z(v) = B(v) | (U(v) & B(v-x) & B(v-y) & B(v-z)), B = hot | empty,
 * U = the +x,+y,+z neighbours were empty before the sweep.  Writes a 0/1
 * mask (interior).  Used by tests to show the closed form equals the sweep. */
void or_removal_mask(int nx, int ny, int nz, const double* temp, const double* rhokap,
                     double ablate_K, uint8_t* mask);

/* ------------------------------------------------- spectral helpers --- */

/* ch_opt.f90:28-152 (compiled in the reference, never called there): a
 * tabulated spectrum array(len, 2), column-major -- array[i] the wavelength
 * of row i+1, array[len + i] its value -- its trapezoid CDF, and sampling by
 * inverse interpolation.  Indices are the reference's 1-based ones. */
void   or_mk_cdf(const double* array, double* cdf, int length);                 /* :138-152 */
int    or_search_1d(int length, const double* array, double value);             /* :97-116 */
int    or_search_2d(int length, const double* array, double value);             /* :119-135 */
double or_lin_inter_1d(const double* array, const double* cdf, double value, int length,
                       int nlow);                                               /* :63-76 */
double or_lin_inter_2d(const double* array, double value, int length, int nlow);   /* :79-94 */
double or_sample(const double* array, int size_of, const double* cdf, or_ran2_t* rng);   /* :28-45 */
/* sample from a given uniform (the draw the GPU takes from Philox) */
double or_sample_u(const double* array, int size_of, const double* cdf, double value);

/* ---------------------------------------- thermal fits, pulse, setup --- */

double or_airThermalCond(double T);          /* thermalConst_mod.f90:11-25 */
double or_airDensity(double T);              /* :28-38 */
double or_getWaterContent(double Q, double W, double QVapor);   /* :41-50 */
double or_getSkinDensity(double W);          /* :54-62 */
double or_getSkinHeatCap(double W);          /* :65-73 */
double or_getSkinThermalCond(double W, double rho);   /* :76-85 */

typedef struct {
    double power, pulseLength, realPulseLength, time;
    double pulseCount, repetitionCount, repetitionRate_1;
    double laserOn;
    int    laser_flag, pulseFlag, pulsesDone, pulsesToDo;
    int    type;                    /* 0 tophat, 1 gaussian, 2 triangular */
} or_pulse_t;
double or_getPwr(or_pulse_t* p);             /* 3dFD.f90:365-421 */
void   or_pulse_advance(or_pulse_t* p, double delt);   /* 3dFD.f90:199-214 */

/* initThermalCoeff scalars (3dFD.f90:233-309). */
typedef struct {
    double dx, dy, dz, delt, alpha, kappa_skin, density, heatcap, coeff;
    double kappa_air, alpha_air_top, volumeVoxel, massVoxel, QVapor, pulseLength;
} or_thermal_init_t;
void or_init_thermal(int n, double xmax, double ymax, double zmax, int nxg, int nyg, int nzg,
                     double energyPerPixel, double power, or_thermal_init_t* out);

#ifdef __cplusplus
}
#endif
#endif
