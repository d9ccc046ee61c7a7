// Internal declarations shared by the wovengpu host code and the HIP kernel files.
//
// A context holds one or more instances (independent woven fields, possibly of different grid
// sizes). Their tubes are concatenated into one tube table, so geometry, binning and
// deposition run as single launches over all instances; the FFT runs once per group of
// instances with equal N. A plain wt_create context is a batch of one.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wovengpu.h"

namespace wt {

constexpr double kPi = 3.141592653589793115997963468544185161590576171875;   // 4*atan(1d0)
constexpr double kPi2 = 6.28318530717958623199592693708837032318115234375;   // 8*atan(1d0)

// Per-instance run description (device copy in Geo::runs, indexed by Geo::tube_inst).
struct Run {
    int32_t N;
    int32_t line_s, line_e;     // 0-based inclusive tube range (relative to tube0) that deposits
    int32_t tube0;              // first global tube of the instance
    double sigma_ratio, eta, Rtube_ratio, dzeta_sup;
    int32_t ntf_times, ntf_pre;
    double dx;                  // 2*pi/N
    double start;               // -pi
    int64_t field_off;          // element offset of the instance's 3*N^3 field block
    int64_t brick0;             // first global brick of the instance
    int32_t nb[3];              // bricks per axis, ceil(N / bd[axis])
    int32_t bd[3], lbd[3];      // deposition brick: cells per axis (powers of two) and their log2
    int32_t colcap;             // deposition: columns per row segment and pass
    int32_t fx_shift;           // FP32 mode: the instance's accumulator scale 2^fx_shift, from its own
    double fx_scale, fx_inv;    // tubes only, so a batched instance rounds like a single run
};

// The deposition accumulates one brick of Run::bd[0] x bd[1] x bd[2] cells at a time in LDS:
// powers of two from kBrickMin to kBrickMax per axis and at most kBrickCells cells; the column
// stage walks at most Run::colcap columns of a row segment per pass (1..kColCapMax). Defaults
// apply until wt_set_deposit_tiles or wt_autotune choose others for a grid size.
constexpr int kBrickCells = 4096;
constexpr int kBrickMin = 4, kBrickMax = 64;
constexpr int kColCapMax = 64;
constexpr int kDefaultTiles[4] = {16, 16, 16, 8};   // bx, by, bz, colcap

// Device-side view of the tube table and geometry (all pointers are device memory).
struct Geo {
    int32_t nline;
    int64_t npoint_all, nsec_all, nsub_all;
    const Run *runs;            // [ninst]
    const int32_t *tube_inst;   // [nline] instance of each tube
    // per tube
    const int64_t *pt_base;     // [nline+1]
    const int64_t *sec_base;    // [nline+1]
    const double *sigma;        // [nline] sigma_min
    const double *gamma;        // [nline]
    const int32_t *n_T;         // [nline]
    double *Len_Disc;           // [nline]
    double *length;             // [nline]
    // per point
    const double *cx, *cy, *cz; // [P]
    int32_t *pt_tube;           // [P]
    double *chord;              // [P] scratch: chord to the previous point (closing chord at the first)
    double *cze;                // [P]
    double *ck;                 // [3P]
    double *ckk;                // [3P]
    // per section
    int32_t *sec_tube;          // [S]
    double *coe;                // [18S]: coe[s*18 + c*6 + j]
    double *czl, *czr;          // [S]
    int32_t *ntf;               // [S]
    double *dtf;                // [S]
    int64_t *sp_base;           // [S+1]
    // per sub-point
    int32_t *sub_sec;           // [NSP]
    int32_t *sub_tube;          // [NSP] tube of each sub-point
    double *ci;                 // [3*NSP] xyz interleaved
};

// Work layout of the deposition: for every brick, the sub-segments whose bounding box overlaps
// it (pairs[brick_off[b] .. brick_off[b+1]) hold sub-point indices, in no particular order, and
// pair_img which periodic image of the brick the box meets).
struct Bins {
    int64_t nbricks;
    int4 *brick_tab;            // [nbricks] (instance, bx, by, bz)
    int64_t *brick_off;         // [nbricks+1]
    uint32_t *brick_wide;       // [nbricks] leading pairs of each brick's list that come from wide boxes
    uint32_t *pairs;            // [npairs]
    uint8_t *pair_img;          // [npairs] periodic image of the brick, kx + 3ky + 9kz (k = 0..2)
    int64_t npairs;
};

// Work buffers of the geometry and binning stages, kept from the last eager run so that the
// stages can be replayed (recorded into a hipGraph) with no allocation.
struct SegWs {
    double *tot = nullptr;              // [nline] chord sums
    int32_t *err = nullptr;             // [1] non-finite turning angle
    int32_t *redo = nullptr;            // [nsec] sections the fast turning angle could not decide
    unsigned int *nredo = nullptr;      // [1]
    int64_t *scan = nullptr;            // scan workspace of the sub-point layout
};
// What the count pass of the binning learnt about one sub-point's sub-segment, so that the fill
// pass needs no geometry and the deposition takes the box as it is: its box in extended indices
// (tube.f90:793-798) and, for boxes of at most kRecMaskBits (brick, image) entries, which entries
// meet the end-plane slab (bit t: entry t in the order of the walk). A sub-point that starts no
// deposited sub-segment holds an empty box (lo > hi).
constexpr int kRecMaskBits = 32;
struct BinRec {
    int16_t lo[3], hi[3];
    uint32_t mask;
};
struct BinWs {
    BinRec *rec = nullptr;              // [nsub_all] count pass -> fill pass
    uint32_t *cnt = nullptr;            // [nbricks] pairs per brick (fill pass: slots after the wide ones)
    uint32_t *wslot = nullptr;          // [nbricks] fill pass: slots of the wide boxes' pairs
    int32_t *err = nullptr;             // [1]
    uint32_t *big = nullptr;            // wide boxes expanded by a wave each
    uint32_t *nbig = nullptr;           // [1]
    uint32_t big_cap = 0;
    unsigned big_grid = 0;
    int64_t *scan = nullptr;            // scan workspace of the brick offsets
};

struct DeviceScalars {          // small device-resident results, one per instance
    double etot_raw, ens_raw, up, inv_up, enstrophy, etot_norm;
    unsigned long long overflow, hits, candidates;
};

// Host copy of one instance's inputs and its places in the arenas.
struct Instance {
    int N = 0;
    wt_params params{};
    bool have_params = false, have_tubes = false;
    std::vector<int32_t> npoint, n_T, if_filter;
    std::vector<double> sigma, gamma, cx, cy, cz;
    int64_t field_off = 0;      // elements, in the omega/u arenas (3*N^3 per instance)
    int64_t spec_off = 0;       // complex elements, in the spectrum arena
    int nek = 0;
    int64_t E_off = 0;          // doubles, in the spectra arenas
    int32_t tube0 = 0, ntube = 0;
};

}  // namespace wt

struct wt_ctx {
    int device = 0;
    int precision = 32;
    int N = 0;                            // grid of the selected instance
    hipStream_t stream = nullptr;
    std::string err;
    bool have_segments = false, have_bins = false, have_omega = false, have_u = false;
    bool tubes_dirty = true;
    // instances
    std::vector<wt::Instance> inst;
    int cur = 0;                          // instance addressed by the per-instance calls
    std::vector<int> order;               // instances sorted by N (arena order)
    wt::Run *runs_dev = nullptr;
    std::vector<wt::Run> runs_host;
    wt::Run run{};                        // parameters shared by all instances (from instance 0)
    // host copy of the concatenated tube table
    std::vector<int64_t> pt_base, sec_base;
    // device buffers
    wt::Geo geo{};
    std::vector<void *> geo_allocs;       // freed on re-upload
    std::vector<void *> seg_allocs;       // sub-point level buffers
    wt::Bins bins{};
    std::vector<void *> bin_allocs;
    wt::SegWs segws;
    wt::BinWs binws;
    unsigned long long *dep_cnt = nullptr;   // deposition counters [0..3] and brick queue cursor [4]
    uint32_t *pairs_buf = nullptr;        // brick pair list, grown on demand
    size_t pairs_cap = 0;
    void *omega = nullptr;                // arena: sum of 3*N_i^3 values (float or double)
    void *u = nullptr;
    void *spec = nullptr;                 // arena: three half spectra per instance, rows padded to 128 B
    int64_t field_elems = 0, spec_elems = 0;
    double *partials = nullptr;           // per-block partial sums of the spectral pass
    size_t partials_cap = 0;
    std::vector<int> tw_N;                // grid sizes with a twiddle table
    std::vector<void *> tw_tab;           // twiddles exp(-2 pi i k / N) + digit-reversal table
    wt::DeviceScalars *scal = nullptr;    // [ninst]
    double *E_w = nullptr, *E_u = nullptr;   // spectra arenas
    int64_t E_elems = 0;
    bool have_spectra = false;
    // stats
    wt_seg_stats seg_stats{};
    wt_deposit_stats dep_stats{};
    int fixed_shift = 0;
    bool theta_exact_all = false;         // WT_THETA_EXACT=1: reference sum for every section
    unsigned int theta_redo = 0;          // sections re-evaluated exactly by the fast path
    double stage_ms[4] = {0, 0, 0, 0};
    hipEvent_t ev[5] = {};
    // per-kernel timing slots (see wt::kSlot*), events recorded on `stream`
    hipEvent_t kev[16][2] = {};
    bool kvalid[16] = {};
    // hipGraph of the spectral stage (all groups), captured once after the first eager run
    bool use_graph = false;
    bool capturing = false;
    hipGraph_t spec_graph = nullptr;
    hipGraphExec_t spec_exec = nullptr;
    // hipGraph of the whole generation (wt_set_graph(ctx, 2)), captured after an eager run for the
    // current tubes and replayed while they stay unchanged
    bool gen_graph_mode = false;
    hipGraph_t gen_graph = nullptr;
    hipGraphExec_t gen_exec = nullptr;
    // FFT tile shapes chosen by wt_autotune, 7 ints per grid size: N, x lines per block, x threads,
    // y columns per tile, y threads, z columns per tile, z threads (defaults for other N)
    std::vector<int32_t> fft_tiles;
    // deposition tiles per grid size, 5 ints each: N, bx, by, bz, colcap (kDefaultTiles for others)
    std::vector<int32_t> dep_tiles;
    // batched structure functions: work items and per-position separations
    void *sf_items = nullptr;
    int64_t sf_nitems = 0;
    int sf_items_lpg = 0;
    int ninst() const { return (int)inst.size(); }
};

namespace wt {
// kernel timing slots reported by wt_stage_times after the four stage totals
enum : int {
    kSlotTheta = 0, kSlotSubpoints, kSlotBinCount, kSlotDeposit, kSlotBinFill,
    kSlotFftX, kSlotFftY, kSlotFftZ, kSlotFftYInv, kSlotFftXInv, kSlotReduce, kNumSlots
};
inline void kt_begin(wt_ctx *c, int slot) {
    if (c->capturing) return;
    if (c->kev[slot][0] && hipEventRecord(c->kev[slot][0], c->stream) == hipSuccess) c->kvalid[slot] = false;
}
inline void kt_end(wt_ctx *c, int slot) {
    if (c->capturing) return;
    if (c->kev[slot][1] && hipEventRecord(c->kev[slot][1], c->stream) == hipSuccess) c->kvalid[slot] = true;
}
// segments.hip
hipError_t launch_segments(wt_ctx *c);       // points -> sections -> sub-points
hipError_t replay_segments(wt_ctx *c);       // the same kernels on the last layout (graph capture)
hipError_t launch_bins(wt_ctx *c);           // per sub-point boxes -> work items
hipError_t replay_bins(wt_ctx *c);           // the same kernels on the last layout (graph capture)
hipError_t launch_get_bins(wt_ctx *c, int32_t *d_lohi, double *d_R);
hipError_t launch_seg_stats(wt_ctx *c, wt_seg_stats *out);
// deposit.hip: accumulate = false writes omega, true adds to it
hipError_t launch_deposit(wt_ctx *c, bool accumulate = false);
hipError_t replay_deposit(wt_ctx *c);        // the deposit kernel alone, counters left on the device
// the deposition tiles of grid size N (bx, by, bz, colcap)
void dep_tiles_of(const wt_ctx *c, int N, int32_t out[4]);
bool dep_tiles_valid(const int32_t t[4]);
// fft.hip: mode 0 = Biot-Savart + normalisation of every instance; mode 1 = spectrum of the
// resident velocity of instance `only`
hipError_t launch_spectral(wt_ctx *c, int mode, int only);
hipError_t spectral_reserve(wt_ctx *c);
// filter_spec (subroutine.f90:1225-1277) on the vorticity of instance `inst`: every mode with
// kx^2 + ky^2 + kz^2 > kc2 set to zero
hipError_t launch_filter(wt_ctx *c, int inst, int64_t kc2);
// Times the candidate tile shapes of the x, y and z passes for every grid size of the context and
// keeps the fastest in c->fft_tiles; `cached` (7 ints per N, may be empty) is used instead of
// measuring for the sizes it holds. `report` receives one line per grid size.
hipError_t spectral_autotune(wt_ctx *c, const std::vector<int32_t> &cached, std::string &report);
bool fft_size_supported(int N);
int64_t spec_elems_of(int N, int precision);   // complex elements of one instance's three half spectra
void fft_reversal(int N, std::vector<int32_t> &rev, bool r16 = false);   // for make_plan(N, r16)
// stats.hip: structure functions of instance c->cur
hipError_t launch_sf(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps,
                     double *d_out);
// Allocation-free structure functions: the per-block partial sums in a file-scope device array of
// stats.hip whose address sf_scratch_init (called by wt_create with the device current) resolves
// once per device. launch_sf_noalloc returns hipErrorInvalidValue when rows * nsep * orders exceeds
// it; sf_noalloc_max_cols gives the largest nsep * orders for the selected instance.
hipError_t sf_scratch_init(int device);
int64_t sf_noalloc_max_cols(const wt_ctx *c, int mode);
hipError_t launch_sf_noalloc(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps,
                             double *d_out);
// PDF of the three components of omega (field 0) or u (field 1) of instance c->cur (cal_pdf)
hipError_t launch_pdf(wt_ctx *c, int field, int nbins, float *val, float *pdf, int64_t *counts);
// x-increment structure functions of every instance in one launch: d_seps[pos*nsep + m] per arena
// position, d_out[(pos*np + p)*nsep + m]
hipError_t launch_sf_batch(wt_ctx *c, int pmin, int pmax, int nsep, const int32_t *d_seps, double *d_out);
// fbb.hip: fBb centerlines of ntube tubes into host arrays (gauss: fixed random inputs or NULL);
// winding mode with winding[3 ntube] and wiggle, reference mode (winding NULL) with dB[ntube]
hipError_t fbb_centerlines(wt_ctx *c, int ntube, const int32_t *npoint, const double *hurst, const int32_t *winding,
                           double wiggle, const double *dB, uint64_t seed, const uint64_t *keys, const double *gauss,
                           double *cx, double *cy, double *cz);
// points of a tube's generated path (128 for a reference-mode tube of fewer points)
int32_t fbb_generated_points(int32_t npoint, bool ref);
hipError_t fbb_normals(wt_ctx *c, uint64_t seed, uint64_t tube, int64_t count, double *host);
int64_t fbb_embedding(int64_t npoint);
// scan.hip: exclusive prefix sums with a caller-held workspace of scan_scratch_elems(n) int64, and
// a zeroing kernel (no memset nodes in captured stages)
int64_t scan_scratch_elems(int64_t n);
hipError_t exclusive_scan_i64(const int64_t *d_in, int64_t *d_out, int64_t n, hipStream_t s, int64_t *ws);
hipError_t exclusive_scan_i32_to_i64(const int32_t *d_in, int add, int64_t *d_out, int64_t n,
                                     hipStream_t s, int64_t *ws);
hipError_t fill_zero(void *p, size_t bytes, hipStream_t st);
int64_t nek_of(int N);
}  // namespace wt

#define WT_HIP(expr)                                   \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return _e;               \
    } while (0)
