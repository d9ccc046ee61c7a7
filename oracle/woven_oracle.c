/*
 * FP64 CPU oracle of the woven-turbulence generator. TEST INFRASTRUCTURE ONLY: the tests,
 * the smoke check and the CPU-baseline leg of the benchmark use it as the checker; the
 * product library never links it.
 *
 * A plain-C restatement of the reference Fortran (YYgroup/FastWeavTurb tube.f90 and
 * subroutine.f90), compiled with -ffp-contract=off so that every FP64 operation is one IEEE
 * rounding, in the reference's evaluation order.
 *
 * Threads (OpenMP; OMP_NUM_THREADS=1 gives the serial program) never change a result bit:
 *  - geometry runs one tube per thread (tubes are independent, tube.f90:528-751);
 *  - the deposition splits the grid into z-slabs as the reference's MPI ranks do
 *    (tube.f90:803-804): every thread walks every sub-segment but only the planes k of its own
 *    slab, so each cell receives its contributions in the serial order (sub-segment, then kp,
 *    jp, ip);
 *  - each 1-D line of a 3-D FFT is an independent transform;
 *  - sums over the field (Etot, enstrophy, spectra, structure functions) stay serial.
 *
 * Conventions fixed here (the reference leaves them to its compiler):
 *  - real-valued integral exponents are multiplication chains: x**2 = x*x,
 *    x**3 = (x*x)*x, x**4 = (x*x)*(x*x), x**5 = ((x*x)*(x*x))*x (binary powering);
 *  - Fortran left-to-right evaluation of + - * / chains is kept literally.
 *
 * Deviation forced by undefined behaviour in the reference:
 *  - T1: dsigmadzeta (subroutine.f90:748-761) calls sigma_func without its n_T argument.
 *    Here the tube's own n_T is passed. With sigma_ratio = 1 the result is exactly 0 anyway.
 */
#define _POSIX_C_SOURCE 199309L
#include "woven_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.141592653589793115997963468544185161590576171875  /* 4*atan(1d0) */
#define PI2 6.28318530717958623199592693708837032318115234375  /* 8*atan(1d0) */

static char g_err[512];
const char *or_last_error(void) { return g_err; }

int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
static int fail(const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return -1;
}

static inline double p2(double x) { return x * x; }
static inline double p3(double x) { return (x * x) * x; }
static inline double p4(double x) { double y = x * x; return y * y; }
static inline double p5(double x) { double y = x * x; return (y * y) * x; }

/* norm_zeta, subroutine.f90:793-797 */
static inline double norm3(const double *v) { return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }
/* cross_zeta, subroutine.f90:799-804 */
static inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
/* dot_zeta, subroutine.f90:806-810 */
static inline double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
/* cminus, subroutine.f90:695-712: minimum-image difference r - l */
static inline void cminus(const double *r, const double *l, double *d) {
    for (int c = 0; c < 3; c++) {
        double v = r[c] - l[c];
        d[c] = v;
    }
    for (int c = 0; c < 3; c++) if (d[c] > PI) d[c] = d[c] - PI2;
    for (int c = 0; c < 3; c++) if (d[c] < -PI) d[c] = d[c] + PI2;
}

/* sigma_func, subroutine.f90:164-173. omega = floor(length/pi + 0.5) * n_T is an integer. */
double or_sigma(double zeta, double length, double smax, double smin, int n_T) {
    double omega = (double)((long long)floor(length / PI + 0.5) * (long long)n_T);
    return smin + (smax - smin) * (sin(omega * zeta) + 1.0) / 2.0;
}

/* get_spline_5th_coe, subroutine.f90:440-461 */
static void spline5_coe(double *coe, double zr, double xl, double xr, double vl, double vr,
                        double al, double ar) {
    double xpl = xl, xpr = xr;
    if (xl - xr > PI) xpr = xr + PI2;
    else if (xl - xr < -PI) xpl = xl + PI2;
    double d = xpr - xpl;
    coe[0] = xl;
    coe[1] = vl;
    coe[2] = al / 2.0;
    coe[3] = ((20.0 * d - (12.0 * vl + 8.0 * vr) * zr) - (3.0 * al - ar) * p2(zr)) / (2.0 * p3(zr));
    coe[4] = ((-30.0 * d + (16.0 * vl + 14.0 * vr) * zr) + (3.0 * al - 2.0 * ar) * p2(zr)) / (2.0 * p4(zr));
    coe[5] = ((12.0 * d - 6.0 * (vl + vr) * zr) - (al - ar) * p2(zr)) / (2.0 * p5(zr));
}

/* coo_curve_spline_5th, subroutine.f90:490-516; coe: [c*6 + j] */
static void spline_pos(double z, const double *coe, double *c3) {
    double z2 = p2(z), z3 = p3(z), z4 = p4(z), z5 = p5(z);
    for (int c = 0; c < 3; c++) {
        const double *a = coe + 6 * c;
        c3[c] = ((((a[5] * z5 + a[4] * z4) + a[3] * z3) + a[2] * z2) + a[1] * z) + a[0];
    }
    for (int c = 0; c < 3; c++) if (c3[c] >= PI) c3[c] = c3[c] - PI2;
    for (int c = 0; c < 3; c++) if (c3[c] < -PI) c3[c] = c3[c] + PI2;
}
/* d1_coo_curve_spline_5th, subroutine.f90:518-537 */
static void spline_d1(double z, const double *coe, double *d) {
    double z2 = p2(z), z3 = p3(z), z4 = p4(z);
    for (int c = 0; c < 3; c++) {
        const double *a = coe + 6 * c;
        d[c] = ((((5.0 * a[5]) * z4 + (4.0 * a[4]) * z3) + (3.0 * a[3]) * z2) + (2.0 * a[2]) * z) + a[1];
    }
}
/* d2_coo_curve_spline_5th, subroutine.f90:539-558 */
static void spline_d2(double z, const double *coe, double *d) {
    double z2 = p2(z), z3 = p3(z);
    for (int c = 0; c < 3; c++) {
        const double *a = coe + 6 * c;
        d[c] = (((20.0 * a[5]) * z3 + (12.0 * a[4]) * z2) + (6.0 * a[3]) * z) + 2.0 * a[2];
    }
}

void or_geom_free(or_geom *g) {
    void **ptrs[] = {(void **)&g->pt_base, (void **)&g->sec_base, (void **)&g->seg_base,
                     (void **)&g->Len_Disc, (void **)&g->length, (void **)&g->cze, (void **)&g->ck,
                     (void **)&g->ckk, (void **)&g->coe, (void **)&g->cze_l, (void **)&g->cze_r,
                     (void **)&g->ntf, (void **)&g->dtf, (void **)&g->theta, (void **)&g->sp_base,
                     (void **)&g->ci, (void **)&g->s, (void **)&g->tci, (void **)&g->ndci,
                     (void **)&g->seg_sec, (void **)&g->seg_j};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) {
        free(*ptrs[i]);
        *ptrs[i] = NULL;
    }
}

/* tube.f90:392-515: cze (normalised cumulative min-image chord length), Len_Disc, ck, ckk. */
int or_preprocess(const or_params *p, int nline, const int32_t *npoint, const double *cx,
                  const double *cy, const double *cz, or_geom *g) {
    (void)p;
    memset(g, 0, sizeof *g);
    if (nline <= 0) return fail("nline must be positive");
    g->nline = nline;
    g->pt_base = (int64_t *)malloc(sizeof(int64_t) * (nline + 1));
    g->pt_base[0] = 0;
    for (int l = 0; l < nline; l++) {
        if (npoint[l] < 3) return fail("every tube needs at least 3 points");
        g->pt_base[l + 1] = g->pt_base[l] + npoint[l];
    }
    int64_t P = g->pt_base[nline];
    g->npoint_all = P;
    g->Len_Disc = (double *)malloc(sizeof(double) * nline);
    g->length = (double *)malloc(sizeof(double) * nline);
    g->cze = (double *)malloc(sizeof(double) * P);
    g->ck = (double *)malloc(sizeof(double) * 3 * P);
    g->ckk = (double *)malloc(sizeof(double) * 3 * P);
    for (int l = 0; l < nline; l++) {
        int64_t b = g->pt_base[l];
        int np = npoint[l];
        const double *x = cx + b, *y = cy + b, *z = cz + b;
        double *cze = g->cze + b;
        double cdx, cdy, cdz, dist;
        cze[0] = 0.0;
        for (int i = 1; i < np; i++) { /* tube.f90:419-428 */
            cdx = fabs(x[i] - x[i - 1]);
            cdy = fabs(y[i] - y[i - 1]);
            cdz = fabs(z[i] - z[i - 1]);
            if (cdx > PI) cdx = cdx - PI2;
            if (cdy > PI) cdy = cdy - PI2;
            if (cdz > PI) cdz = cdz - PI2;
            dist = sqrt((p2(cdx) + p2(cdy)) + p2(cdz));
            if (!(dist > 0.0)) return fail("zero-length chord between consecutive centerline points");
            cze[i] = cze[i - 1] + dist;
        }
        cdx = fabs(x[0] - x[np - 1]); /* tube.f90:429-435 */
        cdy = fabs(y[0] - y[np - 1]);
        cdz = fabs(z[0] - z[np - 1]);
        if (cdx > PI) cdx = cdx - PI2;
        if (cdy > PI) cdy = cdy - PI2;
        if (cdz > PI) cdz = cdz - PI2;
        dist = sqrt((p2(cdx) + p2(cdy)) + p2(cdz));
        if (!(dist > 0.0)) return fail("zero-length closing chord");
        double tot = cze[np - 1] + dist;
        double LD = tot / PI2;
        g->Len_Disc[l] = LD;
        for (int i = 0; i < np; i++) cze[i] = cze[i] * PI2 / tot; /* tube.f90:437 */
        /* tube.f90:439-466: ck by central differences with min-image wrap */
        double *ck = g->ck + 3 * b;
        for (int i = 0; i < np; i++) {
            int ip = (i == np - 1) ? 0 : i + 1;
            int im = (i == 0) ? np - 1 : i - 1;
            double d[3] = {x[ip] - x[im], y[ip] - y[im], z[ip] - z[im]};
            double ds;
            if (i == 0) ds = cze[1] - cze[np - 1] + PI2;
            else if (i == np - 1) ds = cze[0] - cze[np - 2] + PI2;
            else ds = cze[i + 1] - cze[i - 1];
            for (int c = 0; c < 3; c++) if (d[c] > PI) d[c] = d[c] - PI2;
            for (int c = 0; c < 3; c++) if (d[c] < -PI) d[c] = d[c] + PI2;
            for (int c = 0; c < 3; c++) ck[3 * i + c] = d[c] / (ds * LD);
        }
        /* tube.f90:468-489: ckk by central differences of ck (no wrap) */
        double *ckk = g->ckk + 3 * b;
        for (int i = 0; i < np; i++) {
            int ip = (i == np - 1) ? 0 : i + 1;
            int im = (i == 0) ? np - 1 : i - 1;
            double ds;
            if (i == 0) ds = cze[1] - cze[np - 1] + PI2;
            else if (i == np - 1) ds = cze[0] - cze[np - 2] + PI2;
            else ds = cze[i + 1] - cze[i - 1];
            for (int c = 0; c < 3; c++) ckk[3 * i + c] = (ck[3 * ip + c] - ck[3 * im + c]) / (ds * LD);
        }
        g->length[l] = LD * PI2; /* tube.f90:501 */
    }
    return 0;
}

/* tube.f90:528-751: ghost points, quintic spline, turning angle -> ntf, sub-points, chords. */
int or_segments(const or_params *p, const double *cx, const double *cy, const double *cz,
                or_geom *g) {
    int nline = g->nline;
    g->sec_base = (int64_t *)malloc(sizeof(int64_t) * (nline + 1));
    g->sec_base[0] = 0;
    for (int l = 0; l < nline; l++)
        g->sec_base[l + 1] = g->sec_base[l] + (g->pt_base[l + 1] - g->pt_base[l]) + 1;
    int64_t S = g->sec_base[nline];
    g->nsec_all = S;
    g->coe = (double *)malloc(sizeof(double) * 18 * S);
    g->cze_l = (double *)malloc(sizeof(double) * S);
    g->cze_r = (double *)malloc(sizeof(double) * S);
    g->ntf = (int32_t *)malloc(sizeof(int32_t) * S);
    g->dtf = (double *)malloc(sizeof(double) * S);
    g->theta = (double *)malloc(sizeof(double) * S);
    g->sp_base = (int64_t *)malloc(sizeof(int64_t) * (S + 1));
    const int ntf_pre = p->ntf_pre;
    if (ntf_pre < 1 || p->ntf_times < 1) return fail("ntf_pre and ntf_times must be positive");

    /* pass 1: coefficients, turning angle, ntf, dtf (one tube per thread) */
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int l = 0; l < nline; l++) {
        int64_t b = g->pt_base[l];
        int n_sec = (int)(g->pt_base[l + 1] - b);
        double LD = g->Len_Disc[l];
        /* *_pl arrays with the two periodic ghost points, tube.f90:553-597 (0-based here) */
        int n = n_sec + 2;
        double *pl = (double *)malloc(sizeof(double) * 10 * n);
        double *czepl = pl, *xpl = pl + n, *vpl = pl + 4 * n, *apl = pl + 7 * n;
        for (int i = 0; i < n; i++) {
            int src = (i < n_sec) ? i : (i - n_sec);
            int64_t q = b + src;
            xpl[i] = cx[q];
            xpl[n + i] = cy[q];
            xpl[2 * n + i] = cz[q];
            for (int c = 0; c < 3; c++) {
                vpl[c * n + i] = g->ck[3 * q + c];
                apl[c * n + i] = g->ckk[3 * q + c];
            }
            if (i < n_sec) czepl[i] = g->cze[q];
            else if (i == n_sec) czepl[i] = PI2;
            else czepl[i] = PI2 + g->cze[b + 1];
        }
        /* curve_spline_5th receives cze_pl*Len_Disc (tube.f90:598; subroutine.f90:463-488) */
        for (int i = 0; i < n_sec + 1; i++) {
            int64_t s = g->sec_base[l] + i;
            double zr = czepl[i + 1] * LD - czepl[i] * LD;
            for (int c = 0; c < 3; c++)
                spline5_coe(g->coe + 18 * s + 6 * c, zr, xpl[c * n + i], xpl[c * n + i + 1],
                            vpl[c * n + i], vpl[c * n + i + 1], apl[c * n + i], apl[c * n + i + 1]);
            g->cze_l[s] = czepl[i];
            g->cze_r[s] = czepl[i + 1];
        }
        /* tube.f90:625-642: turning angle with ntf_pre+1 samples, then ntf and dtf */
        for (int i = 0; i < n_sec + 1; i++) {
            int64_t s = g->sec_base[l] + i;
            const double *coe = g->coe + 18 * s;
            double dtfp = (czepl[i + 1] - czepl[i]) / (1.0 * ntf_pre);
            double theta = 0.0;
            for (int tj = 1; tj <= ntf_pre + 1; tj++) {
                double sp = czepl[i] + dtfp * (double)(tj - 1);
                double zz = (sp - czepl[i]) * LD;
                double dc[3], ddc[3], cr[3];
                spline_d1(zz, coe, dc);
                double ndc = norm3(dc);
                spline_d2(zz, coe, ddc);
                cross3(dc, ddc, cr);
                double ncr = norm3(cr);
                theta = theta + ncr / p2(ndc) * dtfp * LD;
            }
            g->theta[s] = theta;
            double q = ceil((double)p->ntf_times * theta / PI2);
            if (!(q >= 0.0 && q < 1.0e8)) { bad = 1; q = 0.0; }
            g->ntf[s] = (int32_t)q + 1;
            g->dtf[s] = (czepl[i + 1] - czepl[i]) / (1.0 * g->ntf[s]);
        }
        free(pl);
    }
    if (bad) return fail("turning angle is not finite");
    /* sub-point layout */
    g->sp_base[0] = 0;
    for (int64_t s = 0; s < S; s++) g->sp_base[s + 1] = g->sp_base[s] + g->ntf[s] + 1;
    int64_t NSP = g->sp_base[S];
    g->nsp_all = NSP;
    g->ci = (double *)malloc(sizeof(double) * 3 * NSP);
    g->s = (double *)malloc(sizeof(double) * NSP);
    g->tci = (double *)calloc(3 * NSP, sizeof(double));
    g->ndci = (double *)calloc(NSP, sizeof(double));
    /* tube.f90:706-713: sub-points of every section (ghost included) */
#pragma omp parallel for schedule(dynamic, 16)
    for (int l = 0; l < nline; l++) {
        double LD = g->Len_Disc[l];
        for (int64_t s = g->sec_base[l]; s < g->sec_base[l + 1]; s++) {
            for (int tj = 1; tj <= g->ntf[s] + 1; tj++) {
                int64_t q = g->sp_base[s] + tj - 1;
                double sv = g->cze_l[s] + g->dtf[s] * (double)(tj - 1);
                g->s[q] = sv;
                spline_pos((sv - g->cze_l[s]) * LD, g->coe + 18 * s, g->ci + 3 * q);
            }
        }
    }
    /* tube.f90:719-723: seam fix; tube.f90:740-751: chords */
    g->seg_base = (int64_t *)malloc(sizeof(int64_t) * (nline + 1));
    g->seg_base[0] = 0;
    for (int l = 0; l < nline; l++) {
        int64_t nseg = 0;
        for (int64_t s = g->sec_base[l]; s < g->sec_base[l + 1] - 1; s++) nseg += g->ntf[s];
        g->seg_base[l + 1] = g->seg_base[l] + nseg;
        for (int64_t s = g->sec_base[l]; s < g->sec_base[l + 1] - 1; s++) {
            int64_t last = g->sp_base[s] + g->ntf[s];
            int64_t first_next = g->sp_base[s + 1];
            for (int c = 0; c < 3; c++) g->ci[3 * last + c] = g->ci[3 * first_next + c];
        }
        for (int64_t s = g->sec_base[l]; s < g->sec_base[l + 1] - 1; s++) {
            for (int tj = 1; tj <= g->ntf[s]; tj++) {
                int64_t q = g->sp_base[s] + tj - 1;
                double *d = g->tci + 3 * q;
                cminus(g->ci + 3 * (q + 1), g->ci + 3 * q, d);
                double nd = norm3(d);
                g->ndci[q] = nd;
                for (int c = 0; c < 3; c++) d[c] = d[c] / nd;
            }
            int64_t q = g->sp_base[s] + g->ntf[s];
            int64_t f = g->sp_base[s + 1];
            double *d = g->tci + 3 * q;
            cminus(g->ci + 3 * (f + 1), g->ci + 3 * f, d);
            double nd = norm3(d);
            g->ndci[q] = nd;
            for (int c = 0; c < 3; c++) d[c] = d[c] / nd;
        }
    }
    int64_t NSEG = g->seg_base[nline];
    g->nseg_all = NSEG;
    g->seg_sec = (int32_t *)malloc(sizeof(int32_t) * (NSEG > 0 ? NSEG : 1));
    g->seg_j = (int32_t *)malloc(sizeof(int32_t) * (NSEG > 0 ? NSEG : 1));
    int64_t k = 0;
    for (int l = 0; l < nline; l++)
        for (int64_t s = g->sec_base[l]; s < g->sec_base[l + 1] - 1; s++)
            for (int tj = 0; tj < g->ntf[s]; tj++) {
                g->seg_sec[k] = (int32_t)s;
                g->seg_j[k] = tj;
                k++;
            }
    return 0;
}

/* tube.f90:766-798 for one sub-segment */
static void bin_one(const or_params *p, double smin, int n_T, const or_geom *g, int l,
                    int64_t seg, or_bin_rec *r) {
    const int N = p->N;
    const double dx = PI2 / N, start = -PI;
    int64_t s = g->seg_sec[seg];
    int64_t q = g->sp_base[s] + g->seg_j[seg];
    double s0 = g->s[q], s1 = g->s[q + 1];
    double ci0[3], ci1[3];
    for (int c = 0; c < 3; c++) { ci0[c] = g->ci[3 * q + c]; ci1[c] = g->ci[3 * (q + 1) + c]; }
    for (int c = 0; c < 3; c++) { /* mark1, tube.f90:779-785 */
        if (ci1[c] - ci0[c] > PI) ci0[c] = ci0[c] + PI2;
        else if (ci1[c] - ci0[c] < -PI) ci1[c] = ci1[c] + PI2;
    }
    double smax = p->sigma_ratio * smin;
    double sig = or_sigma(s0, g->length[l], smax, smin, n_T);
    double R = p->Rtube_ratio * sig;
    sig = or_sigma(s1, g->length[l], smax, smin, n_T);
    double R2 = p->Rtube_ratio * sig;
    R = (R2 > R) ? R2 : R;  /* max(Rtube, Rtube_ratio*sigma) */
    for (int c = 0; c < 3; c++) {
        double mn = ci0[c] < ci1[c] ? ci0[c] : ci1[c];
        double mx = ci0[c] > ci1[c] ? ci0[c] : ci1[c];
        r->lo[c] = (int32_t)floor((mn - start - R) / dx + 1.0);
        r->hi[c] = (int32_t)ceil((mx - start + R) / dx + 1.0);
        r->ci0[c] = ci0[c];
        r->ci1[c] = ci1[c];
    }
    r->R = R;
}

int or_bin(const or_params *p, const double *sigma_line, const int32_t *n_T_line,
           const or_geom *g, or_bin_rec *out) {
    for (int l = 0; l < g->nline; l++) {
        int active = (l + 1 >= p->line_s && l + 1 <= p->line_e);
        for (int64_t seg = g->seg_base[l]; seg < g->seg_base[l + 1]; seg++) {
            if (active) bin_one(p, sigma_line[l], n_T_line[l], g, l, seg, out + seg);
            else {
                memset(out + seg, 0, sizeof(or_bin_rec));
                out[seg].R = -1.0;
            }
        }
    }
    return 0;
}

static inline double mesh(int32_t ip, double d, double start) { return ((double)ip - 1.0) * d + start; }

/* Fortran MOD (truncating) wrap of an extended index to 1..N; returns <= 0 when the
 * reference would address outside the array. */
static inline int wrap_idx(int32_t ip, int N) { return (ip + N - 1) % N + 1; }

int or_pair_hits(const or_params *p, const or_geom *g, const or_bin_rec *b, int64_t seg,
                 int32_t ip, int32_t jp, int32_t kp) {
    (void)seg;
    const int N = p->N;
    const double dx = PI2 / N, start = -PI;
    if (ip < b->lo[0] || ip > b->hi[0] || jp < b->lo[1] || jp > b->hi[1] || kp < b->lo[2] || kp > b->hi[2])
        return 0;
    int k = wrap_idx(kp, N);
    if (!(k > 0 && k <= N)) return 0;
    int64_t q = g->sp_base[g->seg_sec[seg]] + g->seg_j[seg];
    const double *tci0 = g->tci + 3 * q, *tci1 = g->tci + 3 * (q + 1);
    double X[3] = {mesh(ip, dx, start), mesh(jp, dx, start), mesh(kp, dx, start)};
    double a0 = X[0] - b->ci0[0], a1 = X[1] - b->ci0[1], a2 = X[2] - b->ci0[2];
    double t0 = (a0 * tci0[0] + a1 * tci0[1]) + a2 * tci0[2];
    if (!(t0 >= 0.0)) return 0;
    double t1 = ((X[0] - b->ci1[0]) * tci1[0] + (X[1] - b->ci1[1]) * tci1[1]) + (X[2] - b->ci1[2]) * tci1[2];
    if (!(t1 < 0.0)) return 0;
    double rho = sqrt(((p2(a0) + p2(a1)) + p2(a2)) - p2(t0));
    return rho < b->R;
}

/* tube.f90:800-892 for one sub-segment, restricted to the grid planes k in [klo, khi] (1-based,
 * wrapped); returns the number of deposits. */
static int64_t deposit_seg(const or_params *p, const or_geom *g, int l, int64_t seg, double Gamma, double smin,
                           double smax, int n_T, double *omega, int32_t klo, int32_t khi) {
    const int N = p->N;
    const double dx = PI2 / N, start = -PI;
    const int64_t N3 = (int64_t)N * N * N;
    double *wx = omega, *wy = omega + N3, *wz = omega + 2 * N3;
    const double eta = p->eta_t;
    const double LD = g->Len_Disc[l], len = g->length[l];
    int64_t hits = 0;
    or_bin_rec b;
    bin_one(p, smin, n_T, g, l, seg, &b);
    int64_t sec = g->seg_sec[seg];
    int64_t q = g->sp_base[sec] + g->seg_j[seg];
    const double *coe = g->coe + 18 * sec;
    double s0 = g->s[q], s1 = g->s[q + 1];
    double sfirst = g->s[g->sp_base[sec]];
    const double *tci0 = g->tci + 3 * q, *tci1 = g->tci + 3 * (q + 1);
    double ndci0 = g->ndci[q];
    double R = b.R;
    double dzeta = p->dzeta_sup < g->dtf[sec] ? p->dzeta_sup : g->dtf[sec];
    for (int32_t kp = b.lo[2]; kp <= b.hi[2]; kp++) {
        int k = wrap_idx(kp, N);
        if (!(k > 0 && k <= N)) continue;
        if (k < klo || k > khi) continue;   /* another thread's slab */
        for (int32_t jp = b.lo[1]; jp <= b.hi[1]; jp++)
            for (int32_t ip = b.lo[0]; ip <= b.hi[0]; ip++) {
                double X[3] = {mesh(ip, dx, start), mesh(jp, dx, start), mesh(kp, dx, start)};
                double a0 = X[0] - b.ci0[0], a1 = X[1] - b.ci0[1], a2 = X[2] - b.ci0[2];
                double t0 = (a0 * tci0[0] + a1 * tci0[1]) + a2 * tci0[2];
                if (!(t0 >= 0.0)) continue;
                double t1 = ((X[0] - b.ci1[0]) * tci1[0] + (X[1] - b.ci1[1]) * tci1[1]) +
                            (X[2] - b.ci1[2]) * tci1[2];
                if (!(t1 < 0.0)) continue;
                double rho = sqrt(((p2(a0) + p2(a1)) + p2(a2)) - p2(t0));
                if (!(rho < R)) continue;
                /* tube.f90:819-892 */
                t1 = ((X[0] - b.ci1[0]) * tci0[0] + (X[1] - b.ci1[1]) * tci0[1]) +
                     (X[2] - b.ci1[2]) * tci0[2];
                double zeta = (s1 * t0 - s0 * t1) / ndci0;
                double zsec = zeta - sfirst;
                double cz3[3], dc[3], ddc[3], cr[3], xc[3], T[3];
                spline_pos(zsec * LD, coe, cz3);
                spline_d1(zsec * LD, coe, dc);
                double ndc = norm3(dc);
                for (int c = 0; c < 3; c++) T[c] = dc[c] / ndc;
                spline_d2(zsec * LD, coe, ddc);
                cross3(dc, ddc, cr);
                double ncr = norm3(cr);
                cminus(X, cz3, xc);
                double nxc = norm3(xc);
                rho = nxc;
                double sigma = or_sigma(zeta, len, smax, smin, n_T);
                double tp = exp(-(p2(rho - 0.0) / 2.0 / p2(sigma))) / (PI2 * p2(sigma));
                double ftd = -exp(-(p2(R) / 2.0 / p2(sigma))) / (PI2 * p2(sigma));
                tp = tp + ftd;
                /* dsigmadzeta (subroutine.f90:748-761) with the tube's n_T (T1) */
                double sl = or_sigma(zeta - dzeta, len, smax, smin, n_T);
                double sr = or_sigma(zeta + dzeta, len, smax, smin, n_T);
                double dsds = (sr - sl) / (2.0 * dzeta);
                dsds = dsds / ndc / LD;
                int i = wrap_idx(ip, N), j = wrap_idx(jp, N);
                int64_t idx = (int64_t)(i - 1) + (int64_t)N * (j - 1) + (int64_t)N * N * (k - 1);
                double tg = tp * Gamma;
                double v[3];
                if (ncr < 1.0e-8) { /* tube.f90:854-867 */
                    double er[3], eth[3];
                    for (int c = 0; c < 3; c++) er[c] = xc[c] / (nxc + 1.0e-15);
                    cross3(er, T, eth);
                    for (int c = 0; c < 3; c++)
                        v[c] = tg * ((T[c] + eta * rho * eth[c]) + (rho * dsds / sigma) * er[c]);
                } else { /* tube.f90:868-892 */
                    double B[3], Nv[3];
                    for (int c = 0; c < 3; c++) B[c] = cr[c] / ncr;
                    cross3(B, T, Nv);
                    double kap = ncr / p3(ndc);
                    double costh = dot3(xc, Nv) / (nxc + 1.0e-15);
                    double sinth = dot3(xc, B) / (nxc + 1.0e-15);
                    double D = 1.0 - kap * rho * costh;
                    double t5 = -(sinth * rho / D);
                    double t6 = costh * rho / D;
                    double t7 = costh * rho * dsds / sigma / D;
                    double t8 = sinth * rho * dsds / sigma / D;
                    for (int c = 0; c < 3; c++)
                        v[c] = tg * ((T[c] + eta * (t5 * Nv[c] + t6 * B[c])) + (t7 * Nv[c] + t8 * B[c]));
                }
                wx[idx] += v[0];
                wy[idx] += v[1];
                wz[idx] += v[2];
                hits++;
            }
    }
    return hits;
}

static int deposit_planes(const or_params *p, const double *sigma_line, const double *Gamma_line,
                          const int32_t *n_T_line, const or_geom *g, int64_t seg_lo, int64_t seg_hi,
                          int32_t kfirst, int32_t klast, double *omega, int64_t *hits_out) {
    const int N = p->N;
    int l0 = p->line_s - 1, l1 = p->line_e - 1;
    if (l0 < 0) l0 = 0;
    if (l1 > g->nline - 1) l1 = g->nline - 1;
    /* Fortran MOD of a negative index would address outside the field */
    int bad = 0;
    for (int l = l0; l <= l1 && !bad; l++) {
        int64_t sa = g->seg_base[l] > seg_lo ? g->seg_base[l] : seg_lo;
        int64_t sb = g->seg_base[l + 1] < seg_hi ? g->seg_base[l + 1] : seg_hi;
#pragma omp parallel for reduction(| : bad)
        for (int64_t seg = sa; seg < sb; seg++) {
            or_bin_rec b;
            bin_one(p, sigma_line[l], n_T_line[l], g, l, seg, &b);
            for (int c = 0; c < 3; c++)
                if (b.lo[c] < 1 - N || b.hi[c] > 2 * N) bad = 1;
        }
    }
    if (bad) return fail("a tube's cut-off radius spans more than the periodic box");
    int64_t hits = 0;
#pragma omp parallel reduction(+ : hits)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        T = omp_get_num_threads();
#endif
        const int64_t K = (int64_t)klast - kfirst + 1;
        const int32_t klo = kfirst + (int32_t)(K * t / T), khi = kfirst + (int32_t)(K * (t + 1) / T) - 1;
        for (int l = l0; l <= l1; l++) {
            const double smin = sigma_line[l], smax = p->sigma_ratio * smin;
            int64_t sa = g->seg_base[l] > seg_lo ? g->seg_base[l] : seg_lo;
            int64_t sb = g->seg_base[l + 1] < seg_hi ? g->seg_base[l + 1] : seg_hi;
            if (klo <= khi)
                for (int64_t seg = sa; seg < sb; seg++)
                    hits += deposit_seg(p, g, l, seg, Gamma_line[l], smin, smax, n_T_line[l], omega, klo, khi);
        }
    }
    if (hits_out) *hits_out = hits;
    return 0;
}

int or_deposit_range(const or_params *p, const double *sigma_line, const double *Gamma_line,
                     const int32_t *n_T_line, const or_geom *g, int64_t seg_lo, int64_t seg_hi,
                     double *omega, int64_t *hits_out) {
    return deposit_planes(p, sigma_line, Gamma_line, n_T_line, g, seg_lo, seg_hi, 1, p->N, omega, hits_out);
}

int or_deposit_planes(const or_params *p, const double *sigma_line, const double *Gamma_line,
                      const int32_t *n_T_line, const or_geom *g, int32_t kfirst, int32_t klast,
                      double *omega, int64_t *hits_out) {
    if (kfirst < 1 || klast > p->N || kfirst > klast) return fail("plane range outside 1..N");
    return deposit_planes(p, sigma_line, Gamma_line, n_T_line, g, 0, g->nseg_all, kfirst, klast, omega, hits_out);
}

int or_deposit(const or_params *p, const double *sigma_line, const double *Gamma_line,
               const int32_t *n_T_line, const or_geom *g, double *omega, int64_t *hits_out) {
    return or_deposit_range(p, sigma_line, Gamma_line, n_T_line, g, 0, g->nseg_all, omega, hits_out);
}

/* ---------------------------------------------------------------------------------------
 * Spectral part. A DFT is fixed by its definition; the reference gets it from FFTW 3.3.8
 * (legacy dfftw_* API, subroutine.f90:860-874). Here: recursive mixed-radix decimation in
 * time over the smallest prime factor, naive DFT for prime lengths.
 * ------------------------------------------------------------------------------------- */
static void dft_rec(int n, const double *in, int stride, double *out, int sign, double *tmp) {
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int r = 2;
    while (r * r <= n && n % r) r++;
    if (n % r) r = n;
    int m = n / r;
    if (m == 1) { /* prime length: direct DFT */
        for (int k = 0; k < n; k++) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                double ang = sign * PI2 * (double)((long long)j * k % n) / n;
                double c = cos(ang), s = sin(ang);
                double xr = in[2 * j * stride], xi = in[2 * j * stride + 1];
                sr += xr * c - xi * s;
                si += xr * s + xi * c;
            }
            out[2 * k] = sr;
            out[2 * k + 1] = si;
        }
        return;
    }
    /* r sub-transforms of length m over inputs j*r + q */
    for (int q = 0; q < r; q++) dft_rec(m, in + 2 * q * stride, stride * r, tmp + 2 * q * m, sign, out);
    for (int k = 0; k < m; k++)
        for (int t = 0; t < r; t++) {
            int kk = k + t * m;
            double sr = 0, si = 0;
            for (int q = 0; q < r; q++) {
                double ang = sign * PI2 * (double)((long long)q * kk % n) / n;
                double c = cos(ang), s = sin(ang);
                double xr = tmp[2 * (q * m + k)], xi = tmp[2 * (q * m + k) + 1];
                sr += xr * c - xi * s;
                si += xr * s + xi * c;
            }
            out[2 * kk] = sr;
            out[2 * kk + 1] = si;
        }
}

int or_fft1d(int n, double *data, int forward) {
    double *in = (double *)malloc(sizeof(double) * 2 * n);
    double *tmp = (double *)malloc(sizeof(double) * 2 * n);
    double *out = (double *)malloc(sizeof(double) * 2 * n);
    memcpy(in, data, sizeof(double) * 2 * n);
    /* dft_rec uses `out` as scratch for its children; give it a separate buffer */
    double *scratch = (double *)malloc(sizeof(double) * 2 * n);
    (void)scratch;
    dft_rec(n, in, 1, out, forward ? -1 : 1, tmp);
    memcpy(data, out, sizeof(double) * 2 * n);
    free(in); free(tmp); free(out); free(scratch);
    return 0;
}

int or_fft3d(int N, double *data, int forward) {
    const int64_t sx = 1, sy = N, sz = (int64_t)N * N;
    const int64_t strides[3] = {sx, sy, sz};
    for (int ax = 0; ax < 3; ax++) {
        int64_t st = strides[ax];
        int64_t o1 = strides[(ax + 1) % 3], o2 = strides[(ax + 2) % 3];
#pragma omp parallel
        {
        double *line = (double *)malloc(sizeof(double) * 2 * N);
#pragma omp for schedule(static)
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                int64_t base = a * o1 + b * o2;
                for (int i = 0; i < N; i++) {
                    line[2 * i] = data[2 * (base + i * st)];
                    line[2 * i + 1] = data[2 * (base + i * st) + 1];
                    if (forward) { line[2 * i] /= N; line[2 * i + 1] /= N; } /* subroutine.f90:1002,1010,1018 */
                }
                or_fft1d(N, line, forward);
                for (int i = 0; i < N; i++) {
                    data[2 * (base + i * st)] = line[2 * i];
                    data[2 * (base + i * st) + 1] = line[2 * i + 1];
                }
            }
        free(line);
        }
    }
    return 0;
}

/* wavenumber, subroutine.f90:888-919 (box length 2*pi: integer wavenumbers, Nyquist = -N/2) */
static inline int wnum(int i, int N) { return ((i + N / 2) % N) - N / 2; }

int or_filter_spec(int N, int kc, double *omega) {
    const int64_t N3 = (int64_t)N * N * N;
    const double kc2 = (double)kc * (double)kc; /* k_c**2, compared as real*8 (subroutine.f90:1244) */
    double *w = (double *)malloc(sizeof(double) * 2 * N3);
    if (!w) return fail("out of memory");
    for (int c = 0; c < 3; c++) {
        double *f = omega + c * N3;
        for (int64_t i = 0; i < N3; i++) { w[2 * i] = f[i]; w[2 * i + 1] = 0.0; }
        or_fft3d(N, w, 1);
        for (int k = 0; k < N; k++)
            for (int j = 0; j < N; j++)
                for (int i = 0; i < N; i++) {
                    double kx = wnum(i, N), ky = wnum(j, N), kz = wnum(k, N);
                    double k2 = (kx * kx + ky * ky) + kz * kz; /* subroutine.f90:914 */
                    if (k2 > kc2) {
                        int64_t id = i + (int64_t)N * j + (int64_t)N * N * k;
                        w[2 * id] = 0.0;
                        w[2 * id + 1] = 0.0;
                    }
                }
        or_fft3d(N, w, 0);
        for (int64_t i = 0; i < N3; i++) f[i] = w[2 * i]; /* dreal, subroutine.f90:1083 */
    }
    free(w);
    return 0;
}

int or_biot_savart(int N, const double *omega, double *u) {
    const int64_t N3 = (int64_t)N * N * N;
    double *w[3], *uh[3];
    for (int c = 0; c < 3; c++) {
        w[c] = (double *)malloc(sizeof(double) * 2 * N3);
        uh[c] = (double *)malloc(sizeof(double) * 2 * N3);
        for (int64_t i = 0; i < N3; i++) { w[c][2 * i] = omega[c * N3 + i]; w[c][2 * i + 1] = 0.0; }
        or_fft3d(N, w[c], 1);
    }
    /* tube.f90:994-1002: u_hat = i (k x w_hat) / k2 */
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                int64_t id = i + (int64_t)N * j + (int64_t)N * N * k;
                double kx = wnum(i, N), ky = wnum(j, N), kz = wnum(k, N);
                double k2 = (kx * kx + ky * ky) + kz * kz;
                double wr[3], wi[3];
                for (int c = 0; c < 3; c++) { wr[c] = w[c][2 * id]; wi[c] = w[c][2 * id + 1]; }
                double cr[3] = {ky * wr[2] - kz * wr[1], kz * wr[0] - kx * wr[2], kx * wr[1] - ky * wr[0]};
                double ci[3] = {ky * wi[2] - kz * wi[1], kz * wi[0] - kx * wi[2], kx * wi[1] - ky * wi[0]};
                for (int c = 0; c < 3; c++) { /* i*(a + ib) = -b + ia */
                    uh[c][2 * id] = -ci[c] / k2;
                    uh[c][2 * id + 1] = cr[c] / k2;
                }
            }
    for (int c = 0; c < 3; c++) { /* tube.f90:1003-1007 */
        uh[c][0] = 0.0;
        uh[c][1] = 0.0;
        or_fft3d(N, uh[c], 0);
        for (int64_t i = 0; i < N3; i++) u[c * N3 + i] = uh[c][2 * i]; /* dreal, subroutine.f90:1083 */
        free(w[c]);
        free(uh[c]);
    }
    return 0;
}

/* cal_helicity, subroutine.f90:1550-1569: sequential sum in the reference loop order */
static double helicity(int N, const double *a, const double *b) {
    const int64_t N3 = (int64_t)N * N * N;
    double h = 0.0;
    for (int64_t i = 0; i < N3; i++)
        h = ((h + a[i] * b[i]) + a[N3 + i] * b[N3 + i]) + a[2 * N3 + i] * b[2 * N3 + i];
    double dx = PI2 / N;
    return h * (dx * dx * dx);
}

int or_normalize(int N, double *omega, double *u, double *out) {
    const int64_t N3 = (int64_t)N * N * N;
    double L = PI2;
    double Etot = helicity(N, u, u);
    Etot = Etot / 2.0 / L / L / L; /* tube.f90:1029 */
    double up = sqrt(2 * Etot / 3);
    for (int64_t i = 0; i < 3 * N3; i++) { omega[i] = omega[i] / up; u[i] = u[i] / up; }
    double Diss = helicity(N, omega, omega);
    double Ens = Diss / 2.0 / L / L / L;
    out[0] = Etot;
    out[1] = up;
    out[2] = Ens;
    out[3] = Etot / up / up;
    return 0;
}

/* get_spectrum, subroutine.f90:1629: nek = int(sqrt((nx^2+ny^2+nz^2)/12.0)) in FP32 */
int or_nek(int N) {
    float v = (float)(3 * N * N) / 12.0f;
    return (int)sqrtf(v);
}

int or_spectrum(int N, const double *f, double *E, int nek_cap) {
    const int64_t N3 = (int64_t)N * N * N;
    int nek = or_nek(N);
    if (nek > nek_cap) return fail("spectrum buffer too small");
    for (int i = 0; i < nek; i++) E[i] = 0.0;
    double *buf = (double *)malloc(sizeof(double) * 2 * N3);
    double *spec = (double *)calloc(N3, sizeof(double));
    for (int c = 0; c < 3; c++) { /* subroutine.f90:1613-1619 */
        for (int64_t i = 0; i < N3; i++) { buf[2 * i] = f[c * N3 + i]; buf[2 * i + 1] = 0.0; }
        or_fft3d(N, buf, 1);
        for (int64_t i = 0; i < N3; i++) spec[i] += (buf[2 * i] * buf[2 * i] + buf[2 * i + 1] * buf[2 * i + 1]) / 2.0;
    }
    for (int k = 0; k < N; k++) /* subroutine.f90:1630-1636: shells int(sqrt(k2+1e-8)+0.5) */
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                double kx = wnum(i, N), ky = wnum(j, N), kz = wnum(k, N);
                double k2 = kx * kx + ky * ky + kz * kz;
                int ik = (int)(sqrt(k2 + 0.00000001) + 0.5);
                if (ik >= 1 && ik <= nek) E[ik - 1] += spec[i + (int64_t)N * j + (int64_t)N * N * k];
            }
    free(buf);
    free(spec);
    return nek;
}

/* subroutine.f90:78, 97-98: lg2L and r = floor(sqrt(2.0)**lg2r + 0.5) in FP32 */
int or_sp_separations(int N, int32_t *r_list, int cap) {
    int lg2L = (int)floorf(logf((float)N) / logf(2.0f) + 0.01f);
    int nr = 2 * lg2L + 1;
    if (nr > cap) return fail("separation buffer too small");
    float s2 = sqrtf(2.0f);
    for (int lr = 0; lr < nr; lr++) {
        float v = 1.0f; /* integer power by repeated multiplication */
        for (int t = 0; t < lr; t++) v = v * s2;
        r_list[lr] = (int32_t)floorf(v + 0.5f);
    }
    return nr;
}

static inline double ipow(double x, int p) {
    double r = 1.0;
    for (int i = 0; i < p; i++) r *= x;
    return r;
}

int or_sp_calculate(int N, int order, const double *u, double *Sp, int32_t *r_list, int cap) {
    int nr = or_sp_separations(N, r_list, cap);
    if (nr < 0) return nr;
    const int64_t N3 = (int64_t)N * N * N;
    const double *ux = u, *uy = u + N3, *uz = u + 2 * N3;
    for (int pp = 1; pp <= order; pp++)
        for (int lr = 0; lr < nr; lr++) {
            int r = r_list[lr];
            double acc = 0.0;
            for (int k = 0; k < N; k++)
                for (int j = 0; j < N; j++)
                    for (int i = 0; i < N; i++) {
                        int64_t id = i + (int64_t)N * j + (int64_t)N * N * k;
                        double dux = ux[(i + r) % N + (int64_t)N * j + (int64_t)N * N * k] - ux[id];
                        double duy = uy[i + (int64_t)N * ((j + r) % N) + (int64_t)N * N * k] - uy[id];
                        double duz = uz[i + (int64_t)N * j + (int64_t)N * N * ((k + r) % N)] - uz[id];
                        double sp = ipow(dux, pp) + ipow(duy, pp) + ipow(duz, pp);
                        sp = sp / N / N / N / 3.0;
                        acc = acc + sp;
                    }
            Sp[(pp - 1) * nr + lr] = acc;
        }
    return nr;
}

int or_sf_x(int N, const double *ux, int pmin, int pmax, int nsep, const int32_t *seps, double *S) {
    const int64_t N3 = (int64_t)N * N * N;
    for (int pp = pmin; pp <= pmax; pp++)
        for (int m = 0; m < nsep; m++) {
            int r = ((seps[m] % N) + N) % N;
            double acc = 0.0;
            for (int64_t line = 0; line < (int64_t)N * N; line++)
                for (int i = 0; i < N; i++) {
                    double d = ux[line * N + (i + r) % N] - ux[line * N + i];
                    acc += ipow(d, pp);
                }
            S[(pp - pmin) * nsep + m] = acc / (double)N3;
        }
    return 0;
}

int or_deposit_filtered(const or_params *p, const double *sigma_line, const double *Gamma_line,
                        const int32_t *n_T_line, const int32_t *if_filter, const or_geom *g,
                        double *omega, int64_t *hits_out) {
    int64_t hits = 0;
    int l0 = p->line_s - 1, l1 = p->line_e - 1;
    if (l0 < 0) l0 = 0;
    if (l1 > g->nline - 1) l1 = g->nline - 1;
    for (int l = l0; l <= l1; l++) {
        int64_t h = 0;
        int rc = or_deposit_range(p, sigma_line, Gamma_line, n_T_line, g, g->seg_base[l], g->seg_base[l + 1],
                                  omega, &h);
        if (rc != 0) return rc;
        hits += h;
        if (if_filter && if_filter[l]) {
            double kcd = ceil(4.0 / sigma_line[l]); /* tube.f90:928 */
            if (!(kcd < 46341.0)) return fail("k_c**2 exceeds the default integer");
            rc = or_filter_spec(p->N, (int)kcd, omega);
            if (rc != 0) return rc;
        }
    }
    if (hits_out) *hits_out = hits;
    return 0;
}

int or_run(const or_params *p, int nline, const int32_t *npoint, const double *cx,
           const double *cy, const double *cz, const double *sigma_line,
           const double *Gamma_line, const int32_t *n_T_line, const int32_t *if_filter,
           double *omega, double *u, double *stats) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    or_geom g;
    int rc = or_preprocess(p, nline, npoint, cx, cy, cz, &g);
    if (rc == 0) rc = or_segments(p, cx, cy, cz, &g);
    if (rc != 0) { or_geom_free(&g); return rc; }
    const int64_t N3 = (int64_t)p->N * p->N * p->N;
    memset(omega, 0, sizeof(double) * 3 * N3);
    int64_t hits = 0;
    rc = or_deposit_filtered(p, sigma_line, Gamma_line, n_T_line, if_filter, &g, omega, &hits);
    or_geom_free(&g);
    if (rc != 0) return rc;
    rc = or_biot_savart(p->N, omega, u);
    if (rc != 0) return rc;
    rc = or_normalize(p->N, omega, u, stats);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    stats[4] = (double)hits;
    stats[5] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return rc;
}
