/*
 * tam_oracle.c -- FP64 CPU restatement of the Tissue-Ablation-MC hot path.
 * TEST INFRASTRUCTURE ONLY (see tam_oracle.h): the checker for the HIP path
 * and the CPU baseline that bench.py times; never part of the product.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  FP
 * contraction is off so that the arithmetic is the operation-by-operation
 * restatement of the Fortran expressions, which lets the pure-absorber walk
 * and the property fits match the reference compiled by flang bit for bit.
 */
#include "tam_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.f90:13 -- the reference's truncated values */
#define REF_PI    3.141592
#define REF_TWOPI 6.283185
#define TWO_PI    6.283185307179586476925286766559

/* ===================================================================== ran2 */
/* src/ran2.f:1-33 */
#define IM1 2147483563
#define IM2 2147483399
#define IMM1 (IM1 - 1)
#define IA1 40014
#define IA2 40692
#define IQ1 53668
#define IQ2 52774
#define IR1 12211
#define IR2 3791
#define NTAB 32
#define NDIV (1 + IMM1 / NTAB)

void or_ran2_init(or_ran2_t* s, int32_t iseed)
{
    memset(s, 0, sizeof(*s));
    s->idum = iseed;
    s->idum2 = 123456789;          /* DATA idum2/123456789/ (ran2.f:10) */
}

double or_ran2(or_ran2_t* s)
{
    const double AM = 1.0 / (double)IM1;
    const double EPS = 1.2e-7;
    const double RNMX = 1.0 - EPS;
    int32_t k, j;
    if (s->idum <= 0) {                                   /* ran2.f:11-21 */
        s->idum = (-s->idum > 1) ? -s->idum : 1;
        s->idum2 = s->idum;
        for (j = NTAB + 8; j >= 1; --j) {
            k = s->idum / IQ1;
            s->idum = IA1 * (s->idum - k * IQ1) - k * IR1;
            if (s->idum < 0) s->idum += IM1;
            if (j <= NTAB) s->iv[j - 1] = s->idum;
        }
        s->iy = s->iv[0];
    }
    k = s->idum / IQ1;                                    /* ran2.f:22-32 */
    s->idum = IA1 * (s->idum - k * IQ1) - k * IR1;
    if (s->idum < 0) s->idum += IM1;
    k = s->idum2 / IQ2;
    s->idum2 = IA2 * (s->idum2 - k * IQ2) - k * IR2;
    if (s->idum2 < 0) s->idum2 += IM2;
    j = 1 + s->iy / NDIV;
    s->iy = s->iv[j - 1] - s->idum2;
    s->iv[j - 1] = s->idum;
    if (s->iy < 1) s->iy += IMM1;
    {
        double r = AM * (double)s->iy;
        return r < RNMX ? r : RNMX;
    }
}

/* =================================================================== Philox */
/* Philox4x32 with 10 rounds: multipliers 0xD2511F53/0xCD9E8D57, Weyl key
 * increments 0x9E3779B9/0xBB67AE85 (Salmon, Moraes, Dror, Shaw, SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double or_u01(uint32_t x)
{
    return ((double)(x >> 9) + 0.5) * (1.0 / 8388608.0);
}

/* ================================================= reference MC (mu_s = 0) */

static inline long hidx(const or_grid_t* g, int i, int j, int k)   /* halo array, 0-based halo */
{
    return (long)i + (long)(g->nxg + 2) * ((long)j + (long)(g->nyg + 2) * (long)k);
}
static inline long vidx(const or_grid_t* g, int i, int j, int k)   /* interior, 1-based */
{
    return (long)(i - 1) + (long)g->nxg * ((long)(j - 1) + (long)g->nyg * (long)(k - 1));
}

/* find (inttau2.f90:208-239): bracketing index of val in a(1..n). */
static int ref_find(double val, const double* a /* a[0..n-1] = a(1..n) */, int n)
{
    int lo = 0, hi = n + 1, mid;
    if (val == a[0]) return 1;
    if (val == a[n - 1]) return n - 1;
    if (val > a[n - 1] || val < a[0]) return -1;
    for (;;) {
        if (hi - lo <= 1) break;
        mid = (hi + lo) / 2;
        if (val >= a[mid - 1]) lo = mid;
        else hi = mid;
    }
    return lo;
}

typedef struct {
    double xcur, ycur, zcur;
    double nxp, nyp, nzp;
    int celli, cellj, cellk;
} ref_walk_t;

/* wall_dist (inttau2.f90:75-121); dir = 1,2,3 for x,y,z (0 = error). */
static double ref_wall_dist(const ref_walk_t* w, const double* xf, const double* yf,
                            const double* zf, int* dir)
{
    double dx, dy, dz, wd;
    if (w->nxp > 0.) dx = (xf[w->celli] - w->xcur) / w->nxp;
    else if (w->nxp < 0.) dx = (xf[w->celli - 1] - w->xcur) / w->nxp;
    else dx = 100000.;
    if (w->nyp > 0.) dy = (yf[w->cellj] - w->ycur) / w->nyp;
    else if (w->nyp < 0.) dy = (yf[w->cellj - 1] - w->ycur) / w->nyp;
    else dy = 100000.;
    if (w->nzp > 0.) dz = (zf[w->cellk] - w->zcur) / w->nzp;
    else if (w->nzp < 0.) dz = (zf[w->cellk - 1] - w->zcur) / w->nzp;
    else dz = 100000.;
    wd = dx;                                  /* min(dx, dy, dz) */
    if (dy < wd) wd = dy;
    if (dz < wd) wd = dz;
    *dir = 0;
    if (wd == dx) *dir = 1;
    if (wd == dy) *dir = 2;
    if (wd == dz) *dir = 3;
    return wd;
}

/* update_pos + update_voxels (inttau2.f90:124-205). */
static void ref_update_pos(ref_walk_t* w, const or_grid_t* g, const double* xf,
                           const double* yf, const double* zf, double dcell, int wall,
                           int dir, double delta)
{
    if (wall) {
        if (dir == 1) {
            if (w->nxp > 0.) w->xcur = xf[w->celli] + delta;
            else if (w->nxp < 0.) w->xcur = xf[w->celli - 1] - delta;
            w->ycur = w->ycur + w->nyp * dcell;
            w->zcur = w->zcur + w->nzp * dcell;
        } else if (dir == 2) {
            w->xcur = w->xcur + w->nxp * dcell;
            if (w->nyp > 0.) w->ycur = yf[w->cellj] + delta;
            else if (w->nyp < 0.) w->ycur = yf[w->cellj - 1] - delta;
            w->zcur = w->zcur + w->nzp * dcell;
        } else {
            w->xcur = w->xcur + w->nxp * dcell;
            w->ycur = w->ycur + w->nyp * dcell;
            if (w->nzp > 0.) w->zcur = zf[w->cellk] + delta;
            else if (w->nzp < 0.) w->zcur = zf[w->cellk - 1] - delta;
        }
        w->celli = ref_find(w->xcur, xf, g->nxg + 1);
        w->cellj = ref_find(w->ycur, yf, g->nyg + 1);
        w->cellk = ref_find(w->zcur, zf, g->nzg + 1);
    } else {
        w->xcur = w->xcur + w->nxp * dcell;
        w->ycur = w->ycur + w->nyp * dcell;
        w->zcur = w->zcur + w->nzp * dcell;
    }
}

/* gridset faces (gridset.f90:23-31) */
static void ref_faces(const or_grid_t* g, double** xf, double** yf, double** zf)
{
    *xf = (double*)malloc(sizeof(double) * (g->nxg + 1));
    *yf = (double*)malloc(sizeof(double) * (g->nyg + 1));
    *zf = (double*)malloc(sizeof(double) * (g->nzg + 1));
    for (int i = 1; i <= g->nxg + 1; ++i) (*xf)[i - 1] = (double)(i - 1) * 2. * g->xmax / (double)g->nxg;
    for (int i = 1; i <= g->nyg + 1; ++i) (*yf)[i - 1] = (double)(i - 1) * 2. * g->ymax / (double)g->nyg;
    for (int i = 1; i <= g->nzg + 1; ++i) (*zf)[i - 1] = (double)(i - 1) * 2. * g->zmax / (double)g->nzg;
}

long or_ref_mc(const or_grid_t* g, const double* rhokap, long nphotons, or_ran2_t* rng,
               double* jmean)
{
    double *xf, *yf, *zf;
    long escapes = 0;
    const double delta = 1.e-8 * (2. * g->zmax / (double)g->nzg);        /* mcpolar.f90:112 */
    const double spot = 250e-4;                                         /* sourceph.f90:23 */
    ref_faces(g, &xf, &yf, &zf);
    for (long j = 0; j < nphotons; ++j) {
        /* sourcephCO2 (sourceph.f90:28-47) */
        double h = spot / 2.;
        double r = or_ran2(rng) * (h * h);
        double theta = or_ran2(rng) * REF_TWOPI;
        double xp = sqrt(r) * cos(theta);
        double yp = sqrt(r) * sin(theta);
        double zp = g->zmax - (1.e-8 * (2. * g->zmax / (double)g->nzg));
        double phi = REF_TWOPI * or_ran2(rng);
        double cosp = cos(phi), sinp = sin(phi);
        double sint = 0., cost = -1.;
        ref_walk_t w;
        w.nxp = sint * cosp;
        w.nyp = sint * sinp;
        w.nzp = cost;
        w.celli = (int)((double)g->nxg * (xp + g->xmax) / (2. * g->xmax)) + 1;
        w.cellj = (int)((double)g->nyg * (yp + g->ymax) / (2. * g->ymax)) + 1;
        w.cellk = (int)((double)g->nzg * (zp + g->zmax) / (2. * g->zmax)) + 1;
        /* tauint1 (inttau2.f90:24-63) */
        w.xcur = xp + g->xmax;
        w.ycur = yp + g->ymax;
        w.zcur = zp + g->zmax;
        double taurun = 0., tau = -log(or_ran2(rng));
        for (;;) {
            int dir;
            double dcell = ref_wall_dist(&w, xf, yf, zf, &dir);
            double rk = rhokap[hidx(g, w.celli, w.cellj, w.cellk)];
            double taucell = dcell * rk;
            if (taurun + taucell < tau) {
                taurun = taurun + taucell;
                jmean[vidx(g, w.celli, w.cellj, w.cellk)] += dcell * rk;
                ref_update_pos(&w, g, xf, yf, zf, dcell, 1, dir, delta);
            } else {
                dcell = (tau - taurun) / rk;
                jmean[vidx(g, w.celli, w.cellj, w.cellk)] += dcell * rk;
                ref_update_pos(&w, g, xf, yf, zf, dcell, 0, dir, delta);
                break;
            }
            if (w.celli == -1 || w.cellj == -1 || w.cellk == -1) {
                ++escapes;
                break;
            }
        }
    }
    free(xf); free(yf); free(zf);
    return escapes;
}

/* stokes (stokes.f90:6-153). d = (cost, sint, cosp, sinp, phi). */
void or_stokes(double d[5], double n[3], double hgg, or_ran2_t* rng)
{
    double cost = d[0], sint = d[1], cosp = d[2], sinp = d[3], phi = d[4];
    double g2 = hgg * hgg;
    if (hgg == 0.0) {
        cost = 2. * or_ran2(rng) - 1.;
        sint = (1. - cost * cost);
        if (sint <= 0.) sint = 0.;
        else sint = sqrt(sint);
        phi = REF_TWOPI * or_ran2(rng);
        sinp = sin(phi);
        cosp = cos(phi);
        n[0] = sint * cosp; n[1] = sint * sinp; n[2] = cost;
    } else {
        double costp = cost, sintp = sint, phip = phi;
        double q = (1. - g2) / (1. - hgg + 2. * hgg * or_ran2(rng));
        double bmu = ((1. + g2) - q * q) / (2. * hgg);
        double cosb2 = bmu * bmu, sinbt, ri1, cosdph, sini2, cosi2 = 0., bott;
        if (fabs(bmu) > 1.) {
            if (bmu > 1.) { bmu = 1.; cosb2 = 1.; }
            else { bmu = -1.; cosb2 = 1.; }
        }
        sinbt = sqrt(1. - cosb2);
        ri1 = REF_TWOPI * or_ran2(rng);
        if (ri1 > REF_PI) {
            double ri3 = REF_TWOPI - ri1;
            double cosi3 = cos(ri3), sini3 = sin(ri3);
            if (bmu == 1. || bmu == -1.) goto done;        /* goto 100 */
            cost = costp * bmu + sintp * sinbt * cosi3;
            if (fabs(cost) < 1.) {
                sint = fabs(sqrt(1. - cost * cost));
                sini2 = sini3 * sintp / sint;
                bott = sint * sinbt;
                cosi2 = costp / bott - cost * bmu / bott;
            } else {
                sint = 0.;
                sini2 = 0.;
                if (cost >= 1.) cosi2 = -1.;
                if (cost <= -1.) cosi2 = 1.;
            }
            cosdph = -cosi2 * cosi3 + sini2 * sini3 * bmu;
            if (fabs(cosdph) > 1.) cosdph = cosdph > 1. ? 1. : -1.;
            phi = phip + acos(cosdph);
            if (phi > REF_TWOPI) phi = phi - REF_TWOPI;
            if (phi < 0.) phi = phi + REF_TWOPI;
        } else {
            double cosi1 = cos(ri1), sini1 = sin(ri1);
            if (bmu == 1. || bmu == -1.) goto done;
            cost = costp * bmu + sintp * sinbt * cosi1;
            if (fabs(cost) < 1.) {
                sint = fabs(sqrt(1. - cost * cost));
                sini2 = sini1 * sintp / sint;
                bott = sint * sinbt;
                cosi2 = costp / bott - cost * bmu / bott;
            } else {
                sint = 0.;
                sini2 = 0.;
                if (cost >= 1.) cosi2 = -1.;
                if (cost <= -1.) cosi2 = 1.;
            }
            cosdph = -cosi1 * cosi2 + sini1 * sini2 * bmu;
            if (fabs(cosdph) > 1.) cosdph = cosdph > 1. ? 1. : -1.;
            phi = phip - acos(cosdph);
            if (phi > REF_TWOPI) phi = phi - REF_TWOPI;
            if (phi < 0.) phi = phi + REF_TWOPI;
        }
        cosp = cos(phi);
        sinp = sin(phi);
        n[0] = sint * cosp; n[1] = sint * sinp; n[2] = cost;
    }
done:
    d[0] = cost; d[1] = sint; d[2] = cosp; d[3] = sinp; d[4] = phi;
}

/* ============================================== extended-physics MC ====== */

double or_fresnel(double n_i, double n_t, double cos_i)
{
    if (n_i == n_t) return 0.0;
    if (cos_i > 1.0) cos_i = 1.0;
    double sin_i = sqrt(fmax(0.0, 1.0 - cos_i * cos_i));
    double sin_t = n_i / n_t * sin_i;
    if (sin_t >= 1.0) return 1.0;
    double cos_t = sqrt(1.0 - sin_t * sin_t);
    double rs = (n_i * cos_i - n_t * cos_t) / (n_i * cos_i + n_t * cos_t);
    double rp = (n_i * cos_t - n_t * cos_i) / (n_i * cos_t + n_t * cos_i);
    return 0.5 * (rs * rs + rp * rp);
}

typedef struct {
    int kind;
    or_ran2_t r2;
    uint32_t key[2];
    uint64_t pid;
    uint32_t stream, event;
    double u[4];
} rng_t;

static void rng_block(rng_t* r)   /* Philox: next block of 4 uniforms; ran2: nothing */
{
    if (r->kind != OR_RNG_PHILOX) return;
    uint32_t ctr[4] = {(uint32_t)r->pid, (uint32_t)(r->pid >> 32), r->event, r->stream};
    uint32_t out[4];
    or_philox4x32_10(ctr, r->key, out);
    for (int i = 0; i < 4; ++i) r->u[i] = or_u01(out[i]);
    r->event++;
}
static double rng_get(rng_t* r, int slot)
{
    return r->kind == OR_RNG_PHILOX ? r->u[slot] : or_ran2(&r->r2);
}

/* HG deflection with the standard direction-cosine rotation (Wang, Jacques &
 * Zheng, MCML 1995, eqs. 3.22-3.27).  Same formula and 0.99999 threshold as
 * the GPU kernel. */
double or_hg_cos(double g, double xi)
{
    double ct;
    if (g == 0.0) {
        ct = 2.0 * xi - 1.0;
    } else {
        double t = (1.0 - g * g) / (1.0 - g + 2.0 * g * xi);
        ct = (1.0 + g * g - t * t) / (2.0 * g);
    }
    if (ct > 1.0) ct = 1.0;
    if (ct < -1.0) ct = -1.0;
    return ct;
}

void or_beam_sample(int beam, double radius, double u0, double u1, double xy[2])
{
    double r = beam == OR_BEAM_TOPHAT ? sqrt(u0 * (radius * radius)) : radius * sqrt(-log(u0) * 0.5);
    xy[0] = r * cos(TWO_PI * u1);
    xy[1] = r * sin(TWO_PI * u1);
}

static void rotate_hg(double* ux, double* uy, double* uz, double g, double xi_c, double xi_p)
{
    const double ct = or_hg_cos(g, xi_c);
    double st = sqrt(1.0 - ct * ct);
    double psi = TWO_PI * xi_p;
    double cp = cos(psi), sp = sin(psi);
    double x = *ux, y = *uy, z = *uz, nx, ny, nz;
    if (fabs(z) > 0.99999) {
        nx = st * cp;
        ny = st * sp;
        nz = z >= 0.0 ? ct : -ct;
    } else {
        double tmp = sqrt(1.0 - z * z);
        nx = st * (x * z * cp - y * sp) / tmp + x * ct;
        ny = st * (y * z * cp + x * sp) / tmp + y * ct;
        nz = -st * cp * tmp + z * ct;
    }
    double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
    *ux = nx * inv; *uy = ny * inv; *uz = nz * inv;
}

static int mc_packet(const or_mc_t* m, const double* xf, const double* yf, const double* zf,
                     rng_t* rng, double* tally, or_mc_stats_t* st)
{
    const or_grid_t* g = &m->grid;
    const double delta = 1.e-8 * (2. * g->zmax / (double)g->nzg);
    const int ref_consts = (m->rng == OR_RNG_RAN2);
    const double twopi = ref_consts ? REF_TWOPI : TWO_PI;
    double xp, yp, zp, r, theta, u;
    ref_walk_t w;
    double ang[5] = {-1., 0., 1., 0., 0.};        /* cost, sint, cosp, sinp, phi */

    /* ---- launch: sourcephCO2 (sourceph.f90:28-47) with a beam profile */
    rng_block(rng);
    u = rng_get(rng, 0);
    if (m->beam == OR_BEAM_TOPHAT) {
        r = sqrt(u * (m->beam_radius * m->beam_radius));
    } else {
        r = m->beam_radius * sqrt(-log(u) * 0.5);
    }
    theta = rng_get(rng, 1) * twopi;
    xp = r * cos(theta);
    yp = r * sin(theta);
    zp = g->zmax - (1.e-8 * (2. * g->zmax / (double)g->nzg));
    if (m->rng == OR_RNG_RAN2) {
        double phi = REF_TWOPI * or_ran2(&rng->r2);     /* direction azimuth draw */
        ang[2] = cos(phi); ang[3] = sin(phi); ang[4] = phi;
    }
    w.nxp = 0. * ang[2];
    w.nyp = 0. * ang[3];
    w.nzp = -1.;
    w.celli = (int)((double)g->nxg * (xp + g->xmax) / (2. * g->xmax)) + 1;
    w.cellj = (int)((double)g->nyg * (yp + g->ymax) / (2. * g->ymax)) + 1;
    w.cellk = (int)((double)g->nzg * (zp + g->zmax) / (2. * g->zmax)) + 1;
    if (xp + g->xmax < 0. || yp + g->ymax < 0. || w.celli < 1 || w.celli > g->nxg ||
        w.cellj < 1 || w.cellj > g->nyg) {
        st->missed += 1.0;
        return 0;
    }
    w.xcur = xp + g->xmax;
    w.ycur = yp + g->ymax;
    w.zcur = zp + g->zmax;
    {
        long v = vidx(g, w.celli, w.cellj, w.cellk);
        double n_in = m->table[m->mat[v]].n, n_out = m->n_ext[OR_FACE_ZHI];
        int mismatch = n_out > 0. && n_out != n_in;
        /* ran2 draws the specular test only where a mismatch exists, so the
         * reference's own sequence (r, theta, phi, tau) is kept otherwise */
        if (mismatch) {
            double rsp = (n_in - n_out) / (n_in + n_out);
            u = rng_get(rng, 2);
            if (u < rsp * rsp) { st->specular += 1.0; return 0; }
        }
    }
    double tau = -log(rng_get(rng, 3));

    /* ---- walk: tauint1 (inttau2.f90:24-63), extended with per-voxel media,
     *      Fresnel on the z faces and a path-length tally */
    for (;;) {
        double taurun = 0.;
        for (;;) {
            int dir;
            double dcell = ref_wall_dist(&w, xf, yf, zf, &dir);
            long v = vidx(g, w.celli, w.cellj, w.cellk);
            double rk = m->mut[v];
            double taucell = dcell * rk;
            if (taurun + taucell < tau) {
                taurun = taurun + taucell;
                if (rk > 0.) tally[v] += dcell;     /* mu_t = 0: nothing absorbed */
                st->crossings += 1.0;
                ref_update_pos(&w, g, xf, yf, zf, dcell, 1, dir, delta);
            } else {
                dcell = (tau - taurun) / rk;
                tally[v] += dcell;
                ref_update_pos(&w, g, xf, yf, zf, dcell, 0, dir, delta);
                break;
            }
            if (w.celli == -1 || w.cellj == -1) {       /* lateral faces: index matched */
                int face = (w.celli == -1) ? (w.xcur < xf[0] ? OR_FACE_XLO : OR_FACE_XHI)
                                           : (w.ycur < yf[0] ? OR_FACE_YLO : OR_FACE_YHI);
                st->escaped[face] += 1.0;
                return 0;
            }
            if (w.cellk == -1) {
                int top = w.zcur > zf[g->nzg];
                int face = top ? OR_FACE_ZHI : OR_FACE_ZLO;
                int kin = top ? g->nzg : 1;
                long vin = vidx(g, w.celli, w.cellj, kin);
                double n_in = m->table[m->mat[vin]].n, n_out = m->n_ext[face];
                int reflect = 0;
                if (n_out > 0. && n_out != n_in) {
                    rng_block(rng);
                    double R = or_fresnel(n_in, n_out, fabs(w.nzp));
                    reflect = rng_get(rng, 0) < R;
                }
                if (!reflect) {
                    st->escaped[face] += 1.0;
                    return 0;
                }
                /* specular bounce back into voxel layer kin */
                w.zcur = top ? (zf[g->nzg] - (w.zcur - zf[g->nzg])) : (zf[0] + (zf[0] - w.zcur));
                w.nzp = -w.nzp;
                ang[0] = -ang[0];
                w.cellk = kin;
            }
        }
        /* ---- interaction: absorb or scatter (albedo test, then HG) */
        st->interactions += 1.0;
        rng_block(rng);
        long v = vidx(g, w.celli, w.cellj, w.cellk);
        const or_material_t* mt = &m->table[m->mat[v]];
        double mt_t = mt->mua + mt->mus;
        double albedo = mt_t > 0. ? mt->mus / mt_t : 0.;
        if (albedo == 0.0 && m->rng == OR_RNG_RAN2) {
            /* pure absorber: the reference's scattering loop exits without a
             * draw (mcpolar.f90:166-169) */
            st->absorbed += 1.0;
            return 0;
        }
        if (!(rng_get(rng, 0) < albedo)) {
            st->absorbed += 1.0;
            return 0;
        }
        if (m->scatter == OR_SCATTER_STOKES) {
            double nn[3];
            if (m->rng == OR_RNG_PHILOX) return -1;    /* stokes draws from ran2 only */
            or_stokes(ang, nn, mt->g, &rng->r2);
            w.nxp = nn[0]; w.nyp = nn[1]; w.nzp = nn[2];
            /* stokes leaves (nxp,nyp,nzp) untouched when bmu = +-1 (goto 100) */
        } else {
            double ux = w.nxp, uy = w.nyp, uz = w.nzp;
            rotate_hg(&ux, &uy, &uz, mt->g, rng_get(rng, 1), rng_get(rng, 2));
            w.nxp = ux; w.nyp = uy; w.nzp = uz;
        }
        tau = -log(rng_get(rng, 3));
    }
}

int or_mc_run(const or_mc_t* m, uint64_t p0, uint64_t npk, double* tally_L,
              or_mc_stats_t* stats, int nthreads)
{
    const or_grid_t* g = &m->grid;
    long nvox = (long)g->nxg * g->nyg * g->nzg;
    double *xf, *yf, *zf;
    int nt = 1;
    if (m->scatter == OR_SCATTER_STOKES && m->rng == OR_RNG_PHILOX) return -1;
    ref_faces(g, &xf, &yf, &zf);
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    if (m->rng == OR_RNG_RAN2) nt = 1;           /* one sequential stream, as one MPI rank */
    double** part = (double**)calloc((size_t)nt, sizeof(double*));
    or_mc_stats_t* pst = (or_mc_stats_t*)calloc((size_t)nt, sizeof(or_mc_stats_t));
    for (int t = 0; t < nt; ++t) part[t] = (t == 0) ? tally_L : (double*)calloc((size_t)nvox, sizeof(double));
    int err = 0;
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        rng_t rng;
        memset(&rng, 0, sizeof(rng));
        rng.kind = m->rng;
        rng.key[0] = (uint32_t)m->seed;
        rng.key[1] = (uint32_t)(m->seed >> 32);
        rng.stream = m->stream;
        if (m->rng == OR_RNG_RAN2) or_ran2_init(&rng.r2, m->ran2_seed);
#pragma omp for schedule(dynamic, 256)
        for (long long p = 0; p < (long long)npk; ++p) {
            rng.pid = p0 + (uint64_t)p;
            rng.event = 0;
            if (mc_packet(m, xf, yf, zf, &rng, part[t], &pst[t]) < 0) err = 1;
        }
    }
    memset(stats, 0, sizeof(*stats));
    for (int t = 0; t < nt; ++t) {
        if (t > 0) {
            for (long v = 0; v < nvox; ++v) tally_L[v] += part[t][v];
            free(part[t]);
        }
        stats->absorbed += pst[t].absorbed;
        stats->specular += pst[t].specular;
        for (int f = 0; f < 6; ++f) stats->escaped[f] += pst[t].escaped[f];
        stats->missed += pst[t].missed;
        stats->crossings += pst[t].crossings;
        stats->interactions += pst[t].interactions;
    }
    free(part); free(pst); free(xf); free(yf); free(zf);
    return err ? -1 : 0;
}

/* ================================================================= heat === */

int or_heat_steps(or_heat_t* h, int loops, const double* laser_on)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const long px = nx + 2, pxy = (long)(nx + 2) * (ny + 2);
    const long ntot = pxy * (nz + 2);
    double* t0 = (double*)malloc(sizeof(double) * ntot);
    double* tn = (double*)malloc(sizeof(double) * ntot);
    const double *kp = h->kappa, *de = h->density, *hc = h->heatcap;
    const double s = h->dscale;
    /* removed voxel: an interior cell with rhokap == 0 (extension) */
#define RM(cc, ii, jj, kk) (h->removed_inert && (ii) >= 1 && (ii) <= nx && (jj) >= 1 && (jj) <= ny && \
                            (kk) >= 1 && (kk) <= nz && h->rhokap[(cc)] == 0.0)
    const double idz2 = 1.0 / (h->dz * h->dz), idy2 = 1.0 / (h->dy * h->dy),
                 idx2 = 1.0 / (h->dx * h->dx);
    int bad = 0;
    memcpy(t0, h->temp, sizeof(double) * ntot);
    memcpy(tn, t0, sizeof(double) * ntot);
    /* the negative-temperature test reads the input array (3dFD.f90:179) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                if (h->temp[i + px * j + pxy * k] < 0.0) bad = 1;
    for (int p = 0; p < loops; ++p) {
        const double lon = laser_on ? laser_on[p] : 1.0;
        if (h->robin_top) {          /* extension: ghost node of a convective face */
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    long c = i + px * j + pxy * nz;
                    double kk = kp[c] / h->dz;
                    double den = kk + 0.5 * h->h_top;
                    t0[c + pxy] = ((kk - 0.5 * h->h_top) * t0[c] + h->h_top * h->T_amb) / den;
                }
        }
        double flux = 0.0;
        /* z-slabs in parallel, as the reference's ranks split z (3dFD.f90:45-72);
         * every voxel's update is independent (Jacobi), so the result does not
         * depend on the thread count */
#pragma omp parallel for reduction(+ : flux) schedule(static)
        for (int k = 1; k <= nz; ++k)          /* 3dFD.f90:115-186 */
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    long c = i + px * j + pxy * k;
                    long v = (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1));
                    double kPH, kMH, dPH, dMH, hPH, hMH, a, d, b, u_xx, u_yy, u_zz;
                    kPH = .5 * (kp[c] + kp[c + pxy]); kMH = .5 * (kp[c] + kp[c - pxy]);
                    dPH = .5 * (de[c] + de[c + pxy]); dMH = .5 * (de[c] + de[c - pxy]);
                    hPH = .5 * (hc[c] + hc[c + pxy]); hMH = .5 * (hc[c] + hc[c - pxy]);
                    a = s * (kMH / (dMH * hMH)) * idz2;
                    d = s * (kPH / (dPH * hPH)) * idz2;
                    if (RM(c, i, j, k) || RM(c - pxy, i, j, k - 1)) a = 0.0;
                    if (RM(c, i, j, k) || RM(c + pxy, i, j, k + 1)) d = 0.0;
                    b = 0.5 * (a + d);
                    u_zz = a * t0[c - pxy] - 2.0 * b * t0[c] + d * t0[c + pxy];
                    kPH = 0.5 * (kp[c] + kp[c + px]); kMH = 0.5 * (kp[c] + kp[c - px]);
                    dPH = 0.5 * (de[c] + de[c + px]); dMH = 0.5 * (de[c] + de[c - px]);
                    hPH = 0.5 * (hc[c] + hc[c + px]); hMH = 0.5 * (hc[c] + hc[c - px]);
                    a = s * (kMH / (dMH * hMH)) * idy2;
                    d = s * (kPH / (dPH * hPH)) * idy2;
                    if (RM(c, i, j, k) || RM(c - px, i, j - 1, k)) a = 0.0;
                    if (RM(c, i, j, k) || RM(c + px, i, j + 1, k)) d = 0.0;
                    b = 0.5 * (a + d);
                    u_yy = a * t0[c - px] - 2.0 * b * t0[c] + d * t0[c + px];
                    kPH = .5 * (kp[c] + kp[c + 1]); kMH = .5 * (kp[c] + kp[c - 1]);
                    dPH = .5 * (de[c] + de[c + 1]); dMH = .5 * (de[c] + de[c - 1]);
                    hPH = .5 * (hc[c] + hc[c + 1]); hMH = .5 * (hc[c] + hc[c - 1]);
                    a = s * (kMH / (dMH * hMH)) * idx2;
                    d = s * (kPH / (dPH * hPH)) * idx2;
                    if (RM(c, i, j, k) || RM(c - 1, i - 1, j, k)) a = 0.0;
                    if (RM(c, i, j, k) || RM(c + 1, i + 1, j, k)) d = 0.0;
                    b = 0.5 * (a + d);
                    u_xx = a * t0[c - 1] - 2.0 * b * t0[c] + d * t0[c + 1];

                    /* boundary energy flow (extension): halo-neighbour terms */
                    if ((i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) && !RM(c, i, j, k)) {
                        double e = 0.0, cc = hc[c] * h->massVoxel * h->delt;
                        if (k == 1) e += s * ((.5 * (kp[c] + kp[c - pxy])) / ((.5 * (de[c] + de[c - pxy])) * (.5 * (hc[c] + hc[c - pxy])))) * idz2 * (t0[c - pxy] - t0[c]);
                        if (k == nz) e += s * ((.5 * (kp[c] + kp[c + pxy])) / ((.5 * (de[c] + de[c + pxy])) * (.5 * (hc[c] + hc[c + pxy])))) * idz2 * (t0[c + pxy] - t0[c]);
                        if (j == 1) e += s * ((.5 * (kp[c] + kp[c - px])) / ((.5 * (de[c] + de[c - px])) * (.5 * (hc[c] + hc[c - px])))) * idy2 * (t0[c - px] - t0[c]);
                        if (j == ny) e += s * ((.5 * (kp[c] + kp[c + px])) / ((.5 * (de[c] + de[c + px])) * (.5 * (hc[c] + hc[c + px])))) * idy2 * (t0[c + px] - t0[c]);
                        if (i == 1) e += s * ((.5 * (kp[c] + kp[c - 1])) / ((.5 * (de[c] + de[c - 1])) * (.5 * (hc[c] + hc[c - 1])))) * idx2 * (t0[c - 1] - t0[c]);
                        if (i == nx) e += s * ((.5 * (kp[c] + kp[c + 1])) / ((.5 * (de[c] + de[c + 1])) * (.5 * (hc[c] + hc[c + 1])))) * idx2 * (t0[c + 1] - t0[c]);
                        flux -= cc * e;
                    }

                    double tempIncrease = h->delt * (u_xx + u_yy + u_zz);
                    double jt = h->jmean ? h->jmean[v] : 0.0;
                    double energyIncrease = lon * jt * h->delt * h->volumeVoxel +
                                            hc[c] * h->massVoxel * tempIncrease;
                    if (h->boil && tn[c] >= 100. + 273. && h->Q[v] < h->QVapor) {
                        if (energyIncrease > 0.0) {
                            double q = h->Q[v] + energyIncrease;
                            h->Q[v] = q < h->QVapor ? q : h->QVapor;
                            tn[c] = 100.0 + 273.0;
                        } else {
                            tn[c] = tn[c] + tempIncrease + lon * h->coeff[c] * jt;
                        }
                    } else {
                        tn[c] = tn[c] + tempIncrease + lon * h->coeff[c] * jt;
                    }
                }
        h->flux_out += flux;
        memcpy(t0, tn, sizeof(double) * ntot);  /* t0 = tn (3dFD.f90:187) */
    }
#undef RM
    /* collate: interior z planes come back (3dFD.f90:218-219) */
    memcpy(h->temp + pxy, t0 + pxy, sizeof(double) * pxy * nz);
    free(t0); free(tn);
    return bad ? -1 : 0;
}

void or_arrhenius(const or_arrhenius_t* a, int nx, int ny, int nz, const double* temp,
                  const double* rhokap, double delt, double time, double* tissue,
                  double* thr)
{
    const long px = nx + 2, pxy = (long)(nx + 2) * (ny + 2);
    const long nv = (long)nx * ny * nz;
    for (int z = 1; z <= nz; ++z)             /* 3dFD.f90:447-464 */
        for (int y = 1; y <= ny; ++y)
            for (int x = 1; x <= nx; ++x) {
                long c = x + px * y + pxy * z;
                long v = (x - 1) + (long)nx * ((y - 1) + (long)ny * (z - 1));
                double T = temp[c];
                if (T >= a->Tlo && T < a->Thi && rhokap[c] >= 0.)
                    tissue[v] = tissue[v] + delt * a->A * exp(-a->dE / (a->R * T));
                if (thr[v] == 0.0 && tissue[v] >= a->thr[0]) thr[v] = time;
                else if (thr[v + nv] == 0.0 && tissue[v] >= a->thr[1]) thr[v + nv] = time;
                else if (thr[v + 2 * nv] == 0.0 && tissue[v] >= a->thr[2]) thr[v + 2 * nv] = time;
            }
}

/* ================================================== thermal fits ======== */
#define W0 0.75
#define PROTEIN (1.0 - W0)

double or_airThermalCond(double T)
{
    const double a = -0.188521, b = 0.000367259, c = 0.212453;
    return a * exp(-b * (T - 273.15)) + c;
}
double or_airDensity(double T) { return 101.325e3 / (287.058 * T); }
double or_getWaterContent(double Q, double W, double QVapor)
{
    double x = W0 - W0 * (Q / QVapor);
    if (x > W0) x = W0;
    if (x > W) x = W;
    return x > 0.0 ? x : 0.0;
}
double or_getSkinDensity(double W) { return 1000.0 / (W + 0.649 * PROTEIN); }
double or_getSkinHeatCap(double W) { return 1000.0 * (4.2 * W + 1.09 * PROTEIN); }
double or_getSkinThermalCond(double W, double rho) { return rho * (6.28e-4 * W + 1.17e-4 * PROTEIN); }

void or_rebuild(or_rebuild_t* r)
{
    const int nx = r->nx, ny = r->ny, nz = r->nz;
    const long px = nx + 2, pxy = (long)(nx + 2) * (ny + 2);
    double* rk = r->rhokap;
    if (r->mode == 0) {                                    /* 3dFD.f90:328 */
        for (long v = 0; v < (long)nx * ny * nz; ++v)
            r->W[v] = or_getWaterContent(r->Q[v], r->W[v], r->QVapor);
    }
    for (int k = 1; k <= nz; ++k)                         /* 3dFD.f90:330-360 */
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                long c = i + px * j + pxy * k;
                long v = (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1));
                double T = r->temp[c];
                if (T >= r->ablate_K) {
                    rk[c] = 0.0;
                } else if (rk[c] > 0.) {
                    if (r->mode == 0) {
                        double W = r->W[v];
                        r->density[c] = or_getSkinDensity(W);
                        rk[c] = W * r->mu_water + r->mu_protein;
                        r->heatcap[c] = or_getSkinHeatCap(W);
                        r->kappa[c] = or_getSkinThermalCond(W, r->density[c]);
                        r->coeff[c] = r->delt / (r->density[c] * r->heatcap[c]);
                    }
                }
                double summ = rk[c + pxy] + rk[c + px] + rk[c + 1] + rk[c - pxy] + rk[c - px] + rk[c - 1];
                if (summ == 0.0) rk[c] = 0.0;
                if (rk[c] <= 0.01) {
                    r->density[c] = or_airDensity(T);
                    r->heatcap[c] = 1.006e3;
                    rk[c] = 0.;
                    r->kappa[c] = or_airThermalCond(T);
                    r->alpha[c] = r->kappa[c] / (r->density[c] * r->heatcap[c]);
                    r->coeff[c] = r->delt / (or_airDensity(T) * r->heatcap[c]);
                }
            }
}

void or_removal_mask(int nx, int ny, int nz, const double* temp, const double* rk,
                     double ablate_K, uint8_t* mask)
{
    const long px = nx + 2, pxy = (long)(nx + 2) * (ny + 2);
#define B_(c) ((c) >= 0 && (temp[(c)] >= ablate_K || rk[(c)] == 0.0))
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                long c = i + px * j + pxy * k;
                long v = (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1));
                int b = temp[c] >= ablate_K || rk[c] == 0.0;
                int u = rk[c + 1] == 0.0 && rk[c + px] == 0.0 && rk[c + pxy] == 0.0;
                /* halo neighbours are empty and never hot-tested: B(halo) = 1 */
                int bx = (i == 1) ? 1 : (temp[c - 1] >= ablate_K || rk[c - 1] == 0.0);
                int by = (j == 1) ? 1 : (temp[c - px] >= ablate_K || rk[c - px] == 0.0);
                int bz = (k == 1) ? 1 : (temp[c - pxy] >= ablate_K || rk[c - pxy] == 0.0);
                mask[v] = (uint8_t)(b || (u && bx && by && bz));
            }
#undef B_
}

/* ====================================================== pulse / init ==== */

double or_getPwr(or_pulse_t* p)
{
    if (p->type == 1) {                                   /* 3dFD.f90:365-377 */
        double fact = (2. * sqrt(2. * log(2.)));
        double mu = fact * p->pulseLength;
        double sig = p->pulseLength / fact;
        double x = p->time - mu;
        return p->power * exp(-(x * x) / (2.0 * (sig * sig)));
    } else if (p->type == 0) {                            /* :380-392 */
        return p->laser_flag ? p->power : 0.0;
    } else {                                              /* :395-421 */
        double m = p->power / p->pulseLength, c = 2.0 * p->power, P;
        if (!p->laser_flag) return 0.0;
        if (p->pulseFlag) {
            P = -m * p->time + c;
            if (P < 0.0) P = 0.0;
        } else if (p->time >= p->pulseLength) {
            p->pulseFlag = 1;
            P = -m * p->time + c;
            if (P < 0.0) P = 0.0;
        } else {
            P = m * p->time;
        }
        return P;
    }
}

void or_pulse_advance(or_pulse_t* p, double delt)
{
    if (p->pulseCount >= p->realPulseLength && p->laser_flag) {    /* 3dFD.f90:199-214 */
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
    p->pulseCount = p->pulseCount + delt;
    p->repetitionCount = p->repetitionCount + delt;
    p->time = p->time + delt;
}

void or_init_thermal(int n, double xmax, double ymax, double zmax, int nxg, int nyg, int nzg,
                     double energyPerPixel, double power, or_thermal_init_t* o)
{
    /* 3dFD.f90:249-291 */
    o->dx = (2.0 * xmax * 1.e-2) / ((double)n + 2.0);
    o->dy = (2.0 * ymax * 1.e-2) / ((double)n + 2.0);
    o->dz = (2.0 * zmax * 1.e-2) / ((double)n + 2.0);
    o->heatcap = or_getSkinHeatCap(W0);
    o->density = or_getSkinDensity(W0);
    o->kappa_skin = or_getSkinThermalCond(W0, o->density);
    o->alpha = o->kappa_skin / (o->density * or_getSkinHeatCap(W0));
    o->kappa_air = or_airThermalCond(25. + 273.);
    o->alpha_air_top = or_airThermalCond(25.0 + 273.0) / (or_airDensity(25.0 + 273.0) * 1.006e3);
    {
        double constd = (1.0 / (o->dx * o->dx)) + (1.0 / (o->dy * o->dy)) + (1.0 / (o->dz * o->dz));
        o->delt = 1.0 / (1.0 * o->alpha * constd);
    }
    o->coeff = o->alpha * o->delt / o->kappa_skin;
    o->pulseLength = (energyPerPixel * 1.e-3 * (double)(9 * 9)) / power;
    o->volumeVoxel = (2.0 * xmax * 1.e-2 / (double)nxg) * (2.0 * ymax * 1.e-2 / (double)nyg) *
                     (2.0 * zmax * 1.e-2 / (double)nzg);
    o->massVoxel = o->density * o->volumeVoxel;
    o->QVapor = 2256.e3 * o->massVoxel;
}

/* ========================================================= spectral === */

/* mk_cdf (ch_opt.f90:138-152): cdf(j) = sum_{i<=j} trapezoid(i, i+1) for
 * j = 1..length-1, each sum restarted from zero and added in order; then
 * every entry, cdf(length) included (left as the caller passed it), divided
 * by cdf(length-1). */
void or_mk_cdf(const double* array, double* cdf, int length)
{
    const double* x = array;
    const double* y = array + length;
    for (int j = 1; j <= length - 1; ++j) {
        double summ = 0.0;
        for (int i = 1; i <= j; ++i)
            summ = summ + 0.5 * (y[i] + y[i - 1]) * (x[i] - x[i - 1]);
        cdf[j - 1] = summ;
    }
    const double norm = cdf[length - 2];
    for (int j = 0; j < length; ++j) cdf[j] = cdf[j] / norm;
}

/* search_1D (ch_opt.f90:97-116): bisection, value > array(middle) moves the
 * lower end; array(length) itself is never compared. */
int or_search_1d(int length, const double* array, double value)
{
    int nup = length, nlow = 1;
    while (nup - nlow > 1) {
        const int middle = (int)((double)(nup + nlow) / 2.0);
        if (value > array[middle - 1]) nlow = middle;
        else nup = middle;
    }
    return nlow;
}

/* search_2D (ch_opt.f90:119-135): the same on the first column. */
int or_search_2d(int length, const double* array, double value)
{
    return or_search_1d(length, array, value);
}

/* lin_inter_1D (ch_opt.f90:63-76): cdf(j) is the CDF at row j+1, so the
 * bracket [cdf(nlow), cdf(nlow+1)] maps onto rows nlow+1 .. nlow+2. */
double or_lin_inter_1d(const double* array, const double* cdf, double value, int length, int nlow)
{
    (void)length;
    const double* x = array;
    return x[nlow] + (x[nlow + 1] - x[nlow]) * (value - cdf[nlow - 1]) / (cdf[nlow] - cdf[nlow - 1]);
}

/* lin_inter_2D (ch_opt.f90:79-94): the second column at `value` of the first. */
double or_lin_inter_2d(const double* array, double value, int length, int nlow)
{
    const double* x = array;
    const double* y = array + length;
    return y[nlow - 1] + (y[nlow] - y[nlow - 1]) * (value - x[nlow - 1]) / (x[nlow] - x[nlow - 1]);
}

double or_sample_u(const double* array, int size_of, const double* cdf, double value)
{
    const int nlow = or_search_1d(size_of, cdf, value);
    return or_lin_inter_1d(array, cdf, value, size_of, nlow);
}

/* sample (ch_opt.f90:28-45): one ran2 draw, inverted through the CDF. */
double or_sample(const double* array, int size_of, const double* cdf, or_ran2_t* rng)
{
    return or_sample_u(array, size_of, cdf, or_ran2(rng));
}
