/*
 * FP64 CPU oracle of the woven-turbulence generator (test infrastructure only).
 *
 * This is a restatement, in plain C, of the reference Fortran path in tube.f90 and
 * subroutine.f90 (YYgroup/FastWeavTurb). It is the checker for the HIP library: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it. The product
 * (libwovengpu.so) never links or calls it.
 *
 * Every function cites the reference lines it follows. Deviations from the reference that
 * are forced by undefined behaviour there are listed next to the function concerned.
 */
#ifndef WOVEN_ORACLE_H
#define WOVEN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Global run parameters, in the order of parameters_input.dat (tube.f90:191-217). */
typedef struct {
    int32_t N;            /* grid size per axis (nx = ny = nz)            */
    int32_t n_d;          /* N_downsample                                   */
    int32_t line_s;       /* first tube deposited (1-based)                 */
    int32_t line_e;       /* last tube deposited (1-based, inclusive)       */
    double sigma_ratio;   /* sigma_max / sigma_min                          */
    double eta_t;         /* eta0, twist rate                               */
    int32_t ntf_times;    /* sub-segments per 2*pi of turning angle         */
    int32_t ntf_pre;      /* samples per section for the turning angle      */
    double Rtube_ratio;   /* cut-off radius / sigma                         */
    double dzeta_sup;     /* upper bound of the central-difference step     */
} or_params;

/* Geometry derived from the centerlines (tube.f90:392-751), flat over all tubes.
 * Sections: tube l owns n_sec_l + 1 sections (the last one is the periodic ghost section of
 * tube.f90:571-597). Sub-points: section s owns ntf[s] + 1 sub-points. Sub-segments: for the
 * real sections (not the ghost), sub-segment j of section s spans sub-points j and j+1. */
typedef struct {
    int32_t nline;
    int64_t npoint_all;
    int64_t nsec_all;      /* sum over tubes of (npoint + 1)   */
    int64_t nsp_all;       /* sum over sections of (ntf + 1)   */
    int64_t nseg_all;      /* sum over real sections of ntf    */
    /* per tube */
    int64_t *pt_base;      /* [nline+1] first point of each tube   */
    int64_t *sec_base;     /* [nline+1] first section of each tube */
    int64_t *seg_base;     /* [nline+1] first sub-segment of each tube */
    double *Len_Disc;      /* [nline] */
    double *length;        /* [nline] */
    /* per point */
    double *cze;           /* [npoint_all]   */
    double *ck;            /* [3*npoint_all] component-major within a point: ck[3*p+c] */
    double *ckk;           /* [3*npoint_all] */
    /* per section */
    double *coe;           /* [nsec_all*18]: coe[s*18 + c*6 + j], j = power */
    double *cze_l;         /* [nsec_all] cze_pl(t_i)   */
    double *cze_r;         /* [nsec_all] cze_pl(t_i+1) */
    int32_t *ntf;          /* [nsec_all] */
    double *dtf;           /* [nsec_all] */
    double *theta;         /* [nsec_all] turning angle estimate */
    int64_t *sp_base;      /* [nsec_all+1] first sub-point of each section */
    /* per sub-point */
    double *ci;            /* [3*nsp_all] positions after the seam fix */
    double *s;             /* [nsp_all] parameter zeta of the sub-point */
    double *tci;           /* [3*nsp_all] chord unit vectors (real sections only) */
    double *ndci;          /* [nsp_all] chord lengths (real sections only) */
    /* per sub-segment */
    int32_t *seg_sec;      /* [nseg_all] section of the sub-segment */
    int32_t *seg_j;        /* [nseg_all] 0-based index t_j-1 inside its section */
} or_geom;

/* Per sub-segment binning result (tube.f90:779-798). */
typedef struct {
    int32_t lo[3], hi[3];  /* extended (unwrapped) 1-based cell index ranges, inclusive */
    double R;              /* Rtube of the sub-segment */
    double ci0[3], ci1[3]; /* end points after the mark1 unwrap (tube.f90:779-785) */
} or_bin_rec;

const char *or_last_error(void);
/* Threads the oracle's parallel loops use (OMP_NUM_THREADS, else every core). */
int or_threads(void);

/* tube.f90:392-515 */
int or_preprocess(const or_params *p, int nline, const int32_t *npoint, const double *cx,
                  const double *cy, const double *cz, or_geom *g);
/* tube.f90:528-751 with subroutine.f90:440-560, 695-712 */
int or_segments(const or_params *p, const double *cx, const double *cy, const double *cz,
                or_geom *g);
void or_geom_free(or_geom *g);

/* tube.f90:760-798 for every sub-segment of tubes line_s..line_e (others get R = -1). */
int or_bin(const or_params *p, const double *sigma_line, const int32_t *n_T_line,
           const or_geom *g, or_bin_rec *out);

/* tube.f90:760-892: vorticity deposition, exact reference scatter loop. omega is
 * 3*N^3 doubles, component-major, each component Fortran order (x fastest). Adds to omega.
 * hits (optional) receives the number of (cell, sub-segment) pairs that deposited. */
int or_deposit(const or_params *p, const double *sigma_line, const double *Gamma_line,
               const int32_t *n_T_line, const or_geom *g, double *omega, int64_t *hits);

/* Same as or_deposit restricted to global sub-segment indices [seg_lo, seg_hi). */
int or_deposit_range(const or_params *p, const double *sigma_line, const double *Gamma_line,
                     const int32_t *n_T_line, const or_geom *g, int64_t seg_lo, int64_t seg_hi,
                     double *omega, int64_t *hits);

/* or_deposit restricted to the grid planes k = kfirst..klast (1-based, wrapped): those planes
 * receive exactly what or_deposit gives them, the others are left unchanged; hits counts the
 * deposits on those planes. The cost scales with the planes' share of the boxes, so a slab of a
 * large grid can be checked in seconds. */
int or_deposit_planes(const or_params *p, const double *sigma_line, const double *Gamma_line,
                      const int32_t *n_T_line, const or_geom *g, int32_t kfirst, int32_t klast,
                      double *omega, int64_t *hits);

/* Brute-force candidate test of one (cell, sub-segment) pair: returns 1 when the reference's
 * loop would deposit on extended cell (ip, jp, kp) from sub-segment seg (tube.f90:800-818). */
int or_pair_hits(const or_params *p, const or_geom *g, const or_bin_rec *b, int64_t seg,
                 int32_t ip, int32_t jp, int32_t kp);

/* subroutine.f90:983-1024 / 1037-1084: 3-D c2c transforms with the reference scaling
 * (forward scaled by 1/N per axis, backward unscaled). data: N^3 complex (re, im interleaved),
 * Fortran order. */
int or_fft3d(int N, double *data, int forward);
/* 1-D DFT of length n (any n), forward sign -1 when forward != 0, unscaled. */
int or_fft1d(int n, double *data, int forward);

/* filter_spec, subroutine.f90:1225-1277: each component of omega (3*N^3) is transformed, every
 * mode with k2 > kc^2 set to zero, and transformed back (real part kept). */
int or_filter_spec(int N, int kc, double *omega);

/* The reference's tube loop with filters (tube.f90:529-936): tubes line_s..line_e deposit in
 * order; after a tube with if_filter[l] != 0 the accumulated omega is filtered with
 * kc = ceiling(4/sigma_line[l]) (tube.f90:926-929). Adds to omega like or_deposit. */
int or_deposit_filtered(const or_params *p, const double *sigma_line, const double *Gamma_line,
                        const int32_t *n_T_line, const int32_t *if_filter, const or_geom *g,
                        double *omega, int64_t *hits);

/* tube.f90:985-1023: u = Re(IFFT(i k x w_hat / k^2)), DC zeroed. omega, u: 3*N^3. */
int or_biot_savart(int N, const double *omega, double *u);

/* tube.f90:1026-1043 with subroutine.f90:1550-1569. Scales omega and u in place by 1/up.
 * out[0] = Etot before scaling, out[1] = up, out[2] = Enstrophy after scaling,
 * out[3] = Etot after scaling (Etot/up^2). */
int or_normalize(int N, double *omega, double *u, double *out);

/* subroutine.f90:1604-1638 in FP64: E[i-1] for shells i = 1..nek. Returns nek. */
int or_spectrum(int N, const double *f, double *E, int nek_cap);
int or_nek(int N);

/* subroutine.f90:64-129: Sp[(p-1)*nr + ir] for p = 1..order, separations r_list[ir].
 * Returns nr; r_list receives the separations (FP32 rule of subroutine.f90:78, 98). */
int or_sp_separations(int N, int32_t *r_list, int cap);
int or_sp_calculate(int N, int order, const double *u, double *Sp, int32_t *r_list, int cap);

/* subroutine.f90:131-162 generalised to a list of separations and orders pmin..pmax:
 * S[(p-pmin)*nsep + m] = sum (ux(i+sep_m) - ux(i))^p / N^3. */
int or_sf_x(int N, const double *ux, int pmin, int pmax, int nsep, const int32_t *seps, double *S);

/* subroutine.f90:164-173 (sigma_func). */
double or_sigma(double zeta, double length, double smax, double smin, int n_T);

/* Whole pipeline in the reference order (tube.f90:392-1043): geometry, deposition (with the
 * per-tube filters when if_filter is not NULL), Biot-Savart, normalisation. omega and u are
 * 3*N^3 each (outputs, normalised). stats[0..3] as or_normalize; stats[4] = hits;
 * stats[5] = seconds (wall). */
int or_run(const or_params *p, int nline, const int32_t *npoint, const double *cx,
           const double *cy, const double *cz, const double *sigma_line,
           const double *Gamma_line, const int32_t *n_T_line, const int32_t *if_filter,
           double *omega, double *u, double *stats);

#ifdef __cplusplus
}
#endif
#endif
