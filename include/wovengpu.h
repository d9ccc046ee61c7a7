/*
 * wovengpu: C ABI of the MI355X (gfx950) woven-turbulence field generator.
 *
 * The reference (YYgroup/FastWeavTurb) is one Fortran program with no plugin API; its hot
 * path is inline code in tube.f90 plus external subroutines in subroutine.f90. Each entry
 * point below replaces one stretch of that path and cites it. A Fortran host calls these
 * through ISO_C_BINDING (fastweavturb_amd/fortran/wovengpu_mod.f90); Python calls them
 * through ctypes (fastweavturb_amd/api.py).
 *
 * Conventions
 *  - Return 0 on success, a negative WT_E* code on failure; wt_last_error() gives the text.
 *    Nothing aborts or throws across the ABI.
 *  - Scalars by value, arrays as plain pointers to contiguous host memory owned by the caller.
 *  - Fields are N^3 values per component in Fortran order (x fastest), like the reference's
 *    vorx(nx, ny, nz). The context owns all device memory and one HIP stream; every call is
 *    stream-ordered and synchronous on return.
 *  - Indices in results are the reference's: 1-based grid indices, extended (unwrapped)
 *    indices for bounding boxes (tube.f90:793-798).
 *  - One context = one GPU; single host thread per context.
 */
#ifndef WOVENGPU_H
#define WOVENGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WT_OK 0
#define WT_EINVAL (-1)     /* bad argument or inconsistent input                */
#define WT_ESTATE (-2)     /* call out of order (e.g. deposit before segments)  */
#define WT_EDEVICE (-3)    /* HIP runtime error                                 */
#define WT_ENOMEM (-4)     /* device allocation failed or memory cap exceeded   */
#define WT_ERANGE (-5)     /* value outside what the kernels support            */

typedef struct wt_ctx wt_ctx;

/* Scalar header of parameters_input.dat, same order and meaning (tube.f90:191-216). */
typedef struct {
    int32_t N;            /* grid points per axis, even                       */
    int32_t n_d;          /* N_downsample (field_wh.dat resolution)           */
    int32_t line_s;       /* first tube deposited, 1-based                    */
    int32_t line_e;       /* last tube deposited, 1-based, inclusive          */
    double sigma_ratio;   /* sigma_max / sigma_min                            */
    double eta_t;         /* twist rate eta0                                  */
    int32_t ntf_times;    /* sub-segments per 2*pi of turning angle           */
    int32_t ntf_pre;      /* samples per section for the turning angle        */
    double Rtube_ratio;   /* cut-off radius / sigma                           */
    double dzeta_sup;     /* upper bound of the dsigma/dzeta step             */
} wt_params;

/* Geometry counters after wt_build_segments (tube.f90:623-738 prints/writes these). */
typedef struct {
    int64_t npoint_all;   /* centerline points                                */
    int64_t nsec_all;     /* sections incl. one periodic ghost per tube       */
    int64_t nsub_all;     /* sub-points: sum over sections of (ntf + 1)       */
    int64_t nseg_all;     /* sub-segments deposited from: sum of ntf, real sections */
    int32_t ntf_max;      /* max over all sections                             */
    int32_t reserved;
    double dtf_min;       /* min over all sections (tube.f90:737)              */
    int64_t ntf_sum;      /* sum of ntf over all sections                      */
    int64_t theta_exact_sections; /* sections whose turning angle used the full
                                     reference sum (all of them if WT_THETA_EXACT=1) */
} wt_seg_stats;

/* Counters of the deposition (wt_deposit). */
typedef struct {
    int64_t columns;      /* grid columns scanned                                */
    int64_t candidates;   /* cells tested with the reference predicate          */
    int64_t hits;         /* (cell, sub-segment) pairs that deposited            */
    int32_t fixed_point_shift; /* FP32 mode: accumulator scale 2^shift (the smallest over the
                                  instances of a batch, each scaled from its own tubes); 0 in FP64 */
    int32_t overflow;     /* nonzero if a fixed-point contribution overflowed    */
    int64_t pairs;        /* (sub-segment, brick) work pairs of the binning      */
} wt_deposit_stats;

/* Field selectors for wt_get_field / wt_set_field. */
#define WT_FIELD_WX 0
#define WT_FIELD_WY 1
#define WT_FIELD_WZ 2
#define WT_FIELD_UX 3
#define WT_FIELD_UY 4
#define WT_FIELD_UZ 5

/* Create a context on HIP device `device` for an N^3 grid computing in FP64 (precision=64,
 * reference arithmetic) or FP32 (precision=32; geometry, binning and hit tests stay FP64).
 * Replaces: set_fft_plans + wavenumber + field allocation (tube.f90:279-306;
 * subroutine.f90:860-919). */
int wt_create(wt_ctx **ctx, int32_t n, int32_t precision, int32_t device);
void wt_destroy(wt_ctx *ctx);

/* Ensemble of ninst independent woven fields with grid sizes n[i] in one context (no
 * counterpart in the reference, which makes one field per run). Their tubes are concatenated,
 * so geometry, binning and deposition run as single launches over all instances, and the FFTs
 * run once per group of equal N. Each instance gets its own wt_set_params/wt_set_tubes after
 * wt_select_instance; wt_generate and the stage calls then process every instance. */
int wt_create_batch(wt_ctx **ctx, int32_t ninst, const int32_t *n, int32_t precision, int32_t device);
/* Instance addressed by wt_set_params, wt_set_tubes, wt_spectrum, wt_structure_functions,
 * wt_get_field, wt_set_field and wt_instance_scalars (0 after creation). */
int wt_select_instance(wt_ctx *ctx, int32_t inst);
/* Etot before normalisation, up and Enstrophy of the selected instance (after Biot-Savart). */
int wt_instance_scalars(wt_ctx *ctx, double *etot, double *up, double *enstrophy);
const char *wt_last_error(const wt_ctx *ctx);

/* Replaces the header read of parameters_input.dat (tube.f90:191-216). p->N must equal the
 * context's N. */
int wt_set_params(wt_ctx *ctx, const wt_params *p);

/* Replaces the per-tube rows of parameters_input.dat (tube.f90:218-221) and the centerline
 * read + broadcast (tube.f90:341-379): nline tubes, npoint[l] points each, coordinates as
 * three flat FP64 arrays of sum(npoint) entries. if_filter may be NULL (all zero); a nonzero
 * entry filters the accumulated vorticity after that tube's deposition (filter_spec,
 * tube.f90:926-936), in single-instance contexts. */
int wt_set_tubes(wt_ctx *ctx, int32_t nline, const int32_t *npoint, const double *sigma,
                 const double *gamma, const int32_t *n_T, const int32_t *if_filter,
                 const double *cx, const double *cy, const double *cz);

/* Replaces tube.f90:392-751: cze/Len_Disc/ck/ckk, quintic splines, turning angle -> ntf/dtf,
 * sub-points with the seam fix. FP64, bit-identical to the reference's operation order. */
int wt_build_segments(wt_ctx *ctx, wt_seg_stats *out);
/* Counters of the last wt_build_segments (also run inside wt_generate). */
int wt_get_seg_stats(wt_ctx *ctx, wt_seg_stats *out);

/* Replaces tube.f90:766-798 (unwrap, Rtube, integer bounding box) for every sub-segment of
 * tubes line_s..line_e, and lays out the column work of the deposition. Integer-exact. */
int wt_bin(wt_ctx *ctx);

/* Replaces tube.f90:800-892: vorticity of every (cell, sub-segment) pair that passes the
 * reference's tests, summed over sub-segments. Leaves omega resident on the device.
 * With if_filter set on tubes of line_s..line_e it also replaces tube.f90:926-936 and
 * subroutine.f90:1225-1277: the tubes up to each such tube are binned and deposited, then every
 * Fourier mode of the accumulated vorticity with kx^2 + ky^2 + kz^2 > k_c^2, k_c =
 * ceiling(4/sigma) of that tube, is set to zero; wt_bin is then not needed beforehand. */
int wt_deposit(wt_ctx *ctx, wt_deposit_stats *out);
/* Replaces filter_spec (subroutine.f90:1225-1277) on the resident vorticity of the selected
 * instance: every Fourier mode with kx^2 + ky^2 + kz^2 > k_c^2 is set to zero. */
int wt_filter_spec(wt_ctx *ctx, int32_t k_c);
/* Counters of the last deposition (also run inside wt_generate). */
int wt_get_deposit_stats(wt_ctx *ctx, wt_deposit_stats *out);

/* Replaces tube.f90:985-1043 (+ subroutine.f90:983-1084, 1550-1569): forward FFT of omega,
 * u_hat = i k x w_hat / k^2 with the DC mode zeroed, inverse FFT, then normalisation of omega
 * and u by up = sqrt(2 Etot / 3). Outputs (any may be NULL): Etot before normalisation, up,
 * and Enstrophy after normalisation. Also prepares the spectra of raw omega and of u. */
int wt_biot_savart(wt_ctx *ctx, double *etot, double *up, double *enstrophy);

/* Whole woven-field path: wt_build_segments + wt_bin + wt_deposit + wt_biot_savart. */
int wt_generate(wt_ctx *ctx, double *etot, double *up, double *enstrophy);

/* Replaces spectrum/get_spectrum (subroutine.f90:1604-1638), accumulated in FP64:
 * E[i-1] for integer shells i = 1..nek, nek = wt_nek(N). field 0 = omega before
 * normalisation (spectrum_vor_temp.dat, tube.f90:972-978), 1 = normalised u (spectrum.dat,
 * tube.f90:1158-1164). Returns nek (>0) or an error code. */
int wt_spectrum(wt_ctx *ctx, int32_t field, double *E, int32_t cap);
int32_t wt_nek(int32_t n);

/* Structure functions of the resident velocity.
 * mode 0 replaces Sp_calculate (subroutine.f90:64-129): S[(p-1)*nsep + m] for p = 1..pmax,
 *   longitudinal increments along x, y and z averaged, separations as given (use
 *   wt_sp_separations for the reference's list); pmin must be 1.
 * mode 1 generalises getStructureFunc_x (subroutine.f90:131-162): x-increments of u_x only,
 *   S[(p-pmin)*nsep + m] = sum (u_x(i+sep_m) - u_x(i))^p / N^3, p = pmin..pmax. */
int wt_structure_functions(wt_ctx *ctx, int32_t mode, int32_t pmin, int32_t pmax,
                           int32_t nsep, const int32_t *seps, double *S);
/* wt_structure_functions without any device allocation: no counterpart in the reference; it is
 * for callers that cannot allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc). seps_dev (nsep int32) and S_dev (nsep * (pmax - pmin + 1) doubles, same layout
 * as S above) are device pointers on the context's GPU. The per-block partial sums live in a fixed
 * 32 MiB device array per GPU (held until the process ends), so nsep * (pmax - pmin + 1) may not
 * exceed 2^22 / rows, rows = 2 x CUs in mode 1 and at most 4 x CUs in mode 0 (8192 and 4096 on a
 * 256-CU MI355X); larger requests return WT_ERANGE. Every call on a GPU shares that array: calls that
 * overlap on one GPU (several contexts or host threads) race. Synchronous on return; the results
 * equal wt_structure_functions' bitwise. */
int wt_structure_functions_noalloc(wt_ctx *ctx, int32_t mode, int32_t pmin, int32_t pmax, int32_t nsep,
                                   const int32_t *seps_dev, double *S_dev);
/* Separations of Sp_calculate (subroutine.f90:78, 97-98). Returns the count. */
int32_t wt_sp_separations(int32_t n, int32_t *r, int32_t cap);

/* Statistics of every instance of a context in one call (no counterpart in the reference, which
 * makes one field per run; these are the ensemble workload's outputs). E (may be NULL) receives
 * each instance's integer-shell spectrum of the normalised velocity, instances in order, nek(N_i)
 * values each (as wt_spectrum with field 1). S receives x-increment structure functions of u_x
 * (as mode 1 of wt_structure_functions) for every instance in one launch:
 * S[(i*(pmax-pmin+1) + p - pmin)*nsep + m] at separation seps[i*nsep + m] of instance i. */
int wt_batch_statistics(wt_ctx *ctx, int32_t pmin, int32_t pmax, int32_t nsep, const int32_t *seps, double *E,
                        double *S);

/* hipGraph replay (no counterpart in the reference). enable = 1: the spectral stage of
 * wt_biot_savart / wt_generate (every FFT pass, Biot-Savart, spectra, reductions and the
 * normalising inverse, for all instances) is captured after the first eager run and launched as
 * one graph afterwards, on any tubes. enable = 2: the whole wt_generate of the current tubes
 * (geometry, binning, deposition and the spectral stage, every instance) is captured after an
 * eager run and replayed as one graph while the tubes, parameters and tile shapes stay unchanged;
 * a change runs the next generation eagerly and captures again. Replays leave the stage times
 * [1..3] of wt_stage_times at -1 ([0] is the whole graph) and the counters of wt_get_seg_stats
 * and wt_get_deposit_stats at those of the eager run, which the same tubes reproduce. Contexts
 * with if_filter tubes are not captured. enable = 0: eager. */
int wt_set_graph(wt_ctx *ctx, int32_t enable);

/* Autotuner (no counterpart in the reference): times the candidate tile shapes of the FFT passes
 * (x lines per block, y and z columns per tile, threads per block) for every grid size of the
 * context on the device and keeps the fastest for later calls. When every instance has its
 * parameters and tubes (and none filters), it also times the binning plus deposition of those
 * tubes for the candidate deposition tiles of each grid size (brick shapes of 8^3..16^3 cells in
 * several aspect ratios, then columns per pass 4, 8, 16 for the fastest shape; the other grid
 * sizes keep theirs meanwhile) and keeps the fastest (wt_set_deposit_tiles). With cache_path,
 * shapes already recorded there for this GPU architecture and precision (deposition tiles also
 * for the grid size and the tubes' radius class) are reused without measuring, and the chosen
 * shapes are written back (JSON). report (may be NULL) receives one line per grid size and stage.
 * The half spectrum, and after a deposition tuning the resident fields, are overwritten; call
 * before wt_generate / wt_biot_savart. */
int wt_autotune(wt_ctx *ctx, const char *cache_path, char *report, int32_t report_cap);

/* Deposition tiles (no counterpart in the reference) of the instances with grid size N (N = 0:
 * every instance): tiles = {bx, by, bz, colcap}, the cells of a brick that one workgroup
 * accumulates in LDS (powers of two from 4 to 64, at most 4096 cells) and the columns of a row
 * segment walked per pass (1..64). Defaults {16, 16, 16, 8}. The field does not depend on them:
 * FP32 results are bitwise the same for every choice, FP64 results equal up to the order of the
 * brick's additions. A change re-bins at the next deposition. wt_get_deposit_tiles reads those of
 * grid size N (0: the selected instance's). */
int wt_set_deposit_tiles(wt_ctx *ctx, int32_t N, const int32_t *tiles);
int wt_get_deposit_tiles(wt_ctx *ctx, int32_t N, int32_t *tiles);

/* Replaces cal_pdf / export_pdf (subroutine.f90:1-62) as called at tube.f90:1166-1185: the PDF
 * of all three components of the resident omega (field 0) or u (field 1), taken as FP32 values
 * like the reference's real(4) copy, over nbins bins between their minimum and maximum
 * (the reference uses bin_num = 1000, tube.f90:129). val[l] are the bin centres and pdf[l] the
 * densities in the reference's FP32 arithmetic; counts (may be NULL) the number of values in each
 * bin. A value goes to the first bin whose interval [val - bs/2, val + bs/2) holds it, so values in
 * no bin (the maximum, usually) are not counted, as in the reference. Counts are exact; the
 * reference counts in FP32, which stops incrementing at 2^24 per bin and rank. */
int wt_pdf(wt_ctx *ctx, int32_t field, int32_t nbins, float *val, float *pdf, int64_t *counts);

/* Copy a resident field component to/from host memory, Fortran order, FP64 or FP32 host
 * buffer (output_v/output_puredata copy-out, subroutine.f90:1342-1399). wt_set_field replaces
 * the resident omega or u (used to drive wt_biot_savart / statistics on a given field). */
int wt_get_field(wt_ctx *ctx, int32_t which, void *host, int32_t host_is_fp64);
int wt_set_field(wt_ctx *ctx, int32_t which, const void *host, int32_t host_is_fp64);

/* Inspection of intermediate results, for parity checks against the reference's arrays:
 * per section (tube.f90 ntf_list, dtf_list; sections include one ghost per tube),
 * per sub-point (ci_list2 after the seam fix, s_list2), per sub-point used as a sub-segment
 * start (bounding box lo[3], hi[3] and Rtube; entries that start no sub-segment get R = -1). */
int wt_get_sections(wt_ctx *ctx, int32_t *ntf, double *dtf);
int wt_get_subpoints(wt_ctx *ctx, double *ci_xyz, double *s);
/* Length 2 pi Len_Disc of every tube of the selected instance, in tube order (tube.f90:501-502);
 * their sum is info.dat's "Total length of curve" (tube.f90:502, 1199). After segments are built. */
int wt_get_tube_lengths(wt_ctx *ctx, double *length);
int wt_get_bins(wt_ctx *ctx, int32_t *lohi, double *R);

/* Timings in milliseconds, measured with HIP events on the context stream. [0..3]: stages of
 * the last wt_generate (segments, bin, deposit, biot_savart). [4..]: the last launch of each
 * main kernel: turning angle, sub-points, bin count, deposit, fixed-point conversion, FFT x,
 * FFT y, FFT z + Biot-Savart, FFT y inverse, FFT x inverse, spectral reduction (-1 if it has
 * not run). Returns the number of entries available (15). */
int wt_stage_times(wt_ctx *ctx, double *ms, int32_t cap);

/* ---- fBb centerline generator ---------------------------------------------------------------
 * The producer of the input centerlines, FBB_with_parameters.m:73-120 (its wfbm draw replaced by
 * an exact fBm sampler), on the context's GPU. Tube l gets npoint[l] points of a closed
 * fractional Brownian bridge with Hurst index hurst[l]: fractional Gaussian noise by circulant
 * embedding (Davies & Harte) of m = wt_fbb_embedding(npoint[l]) points, cumulated, bridged by
 * f(t) = (t^2H + 1 - (1-t)^2H)/2 (FBB:92-102), scaled to a step rms of wiggle * 2 pi |w| / npoint
 * (FBB:104-112), shifted uniformly, drifted by 2 pi w t along the integer lattice vector
 * w = winding[3l..3l+2] and wrapped into [-pi, pi) (FBB:117-120). Outputs cx, cy, cz: host arrays
 * of sum(npoint) points. Random inputs: with gauss = NULL a Philox4x64-10 stream keyed (seed, k_l),
 * k_l = keys[l] (or l when keys is NULL) -- numpy's Philox bit generator with key seed + 2^64 k_l --,
 * normals by Box-Muller; otherwise gauss holds, tube after tube, 4m normals (real parts of the two
 * complex draws, then their imaginary parts) and 3 uniforms in [0, 1) for the shift. A tube's
 * random numbers depend only on (seed, k_l), not on its Hurst index or the other tubes.
 * npoint >= 3 and at most 2^19. */
int wt_fbb_centerlines(wt_ctx *ctx, int32_t ntube, const int32_t *npoint, const double *hurst,
                       const int32_t *winding, double wiggle, uint64_t seed, const uint64_t *keys,
                       const double *gauss, double *cx, double *cy, double *cz);
/* The same construction as FBB_with_parameters.m:73-120 writes it (reference mode): no winding
 * drift, the bridge scaled to the step rms dB[l] (the level's dB_1 r_s^(i-1), FBB:37, 70), a tube of
 * npoint[l] < 128 points sampled from a 128-point path at ceil(j 128 / npoint), j = 1..npoint
 * (FBB:78-84), the shift (u - 1/2) 2 pi added after the scaling and the wrap mod(x, 2 pi) - pi with
 * mod(x, y) = x - floor(x / y) y (FBB:117-120). Random inputs as for wt_fbb_centerlines, with m the
 * embedding of max(npoint[l], 128) points. */
int wt_fbb_reference_centerlines(wt_ctx *ctx, int32_t ntube, const int32_t *npoint, const double *hurst,
                                 const double *dB, uint64_t seed, const uint64_t *keys, const double *gauss,
                                 double *cx, double *cy, double *cz);
/* Embedding size m of a tube of npoint points: 2 x the smallest 2^a 3^b >= npoint - 1. */
int32_t wt_fbb_embedding(int32_t npoint);
/* The first count standard normals of tube `tube`'s random stream (inspection). */
int wt_fbb_normals(wt_ctx *ctx, uint64_t seed, int32_t tube, int64_t count, double *out);

/* ---- Inverse-design surrogate network ------------------------------------------------------
 * No counterpart in the reference, which has no learned component (its closest analogue is the
 * closed-form level scaling of FBB_with_parameters.m:25-41, 185-193). A multilayer perceptron
 * sizes[0] -> sizes[1] -> ... -> sizes[nlayer] with GELU after every hidden layer and a linear
 * output, trained and evaluated on one GPU by hand-written BF16 MFMA kernels with FP32
 * accumulation. Loss: masked mean squared error sum(mask (y - t)^2) / sum(mask). Optimiser:
 * Adam (0.9, 0.999, 1e-8) on FP32 master weights. Parameters are exchanged as one flat FP32
 * array: per layer the weights W[out][in] row-major, then the biases b[out]. */
typedef struct wt_mlp wt_mlp;
int wt_mlp_create(wt_mlp **m, int32_t nlayer, const int32_t *sizes, uint64_t seed, int32_t device);
void wt_mlp_destroy(wt_mlp *m);
const char *wt_mlp_last_error(const wt_mlp *m);
int64_t wt_mlp_num_params(const wt_mlp *m);
int wt_mlp_get_params(wt_mlp *m, float *flat);
/* Replaces the parameters and resets the optimiser state. */
int wt_mlp_set_params(wt_mlp *m, const float *flat);
/* Y[n][out] = network(X[n][in]). */
int wt_mlp_forward(wt_mlp *m, int32_t n, const float *X, float *Y);
/* Loss and its gradient (flat parameter order) on one batch; mask may be NULL (all ones). */
int wt_mlp_loss_grad(wt_mlp *m, int32_t n, const float *X, const float *T, const float *mask, double *loss,
                     float *grad);
/* One Adam step on one batch; *loss is the batch loss before the step. */
int wt_mlp_train_step(wt_mlp *m, int32_t n, const float *X, const float *T, const float *mask, float lr,
                      double *loss);
/* Keeps a training set resident on the device (mask may be NULL). */
int wt_mlp_set_data(wt_mlp *m, int32_t n, const float *X, const float *T, const float *mask);
/* Mini-batch Adam over the resident set: `epochs` seeded shuffles, learning rate decaying from lr
 * to lr_final on a cosine; loss_per_epoch[e] (may be NULL) = mean batch loss of epoch e. */
int wt_mlp_train(wt_mlp *m, int32_t epochs, int32_t batch, float lr, float lr_final, uint64_t seed,
                 double *loss_per_epoch);

#ifdef __cplusplus
}
#endif
#endif
