// Internal declarations shared by the wovengpu host code and the HIP kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wovengpu.h"

namespace wt {

constexpr double kPi = 3.141592653589793115997963468544185161590576171875;   // 4*atan(1d0)
constexpr double kPi2 = 6.28318530717958623199592693708837032318115234375;   // 8*atan(1d0)

// Device-side view of the tube table and geometry (all pointers are device memory).
struct Geo {
    int32_t nline;
    int64_t npoint_all, nsec_all, nsub_all;
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
    double *ci;                 // [3*NSP] xyz interleaved
};

// Run parameters copied to kernels by value.
struct Run {
    int32_t N;
    int32_t line_s, line_e;     // 0-based inclusive tube range that deposits
    double sigma_ratio, eta, Rtube_ratio, dzeta_sup;
    int32_t ntf_times, ntf_pre;
    double dx;                  // 2*pi/N
    double start;               // -pi
};

// Work layout of the deposition: sub-point q starts a sub-segment iff ncol[q] > 0.
struct Bins {
    int64_t *col_count;         // [NSP] columns of the sub-segment's box (0 = not a sub-segment)
    int64_t *item_off;          // [NSP+1] prefix sum of work items (column blocks)
    uint32_t *item_q;           // [items] sub-point of each work item
    int64_t items;
    int64_t columns;
};

struct DeviceScalars {          // small device-resident results
    double etot_raw, ens_raw, up, inv_up, enstrophy, etot_norm;
    unsigned long long overflow, hits, candidates;
};

}  // namespace wt

struct wt_ctx {
    int device = 0;
    int precision = 32;
    int N = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool have_params = false, have_tubes = false, have_segments = false, have_bins = false,
         have_omega = false, have_u = false;
    wt_params params{};
    wt::Run run{};
    // host copy of the tube table
    std::vector<int32_t> npoint;
    std::vector<int64_t> pt_base, sec_base;
    // device buffers
    wt::Geo geo{};
    std::vector<void *> geo_allocs;       // freed on re-upload
    std::vector<void *> seg_allocs;       // sub-point level buffers
    wt::Bins bins{};
    std::vector<void *> bin_allocs;
    void *omega = nullptr;                // 3*N^3 of float or double
    void *u = nullptr;                    // 3*N^3
    void *spec = nullptr;                 // 3 * N * N * (N/2+1) complex
    long long *acc = nullptr;             // 3*N^3 fixed-point accumulators (FP32 mode)
    double *partials = nullptr;           // per-block partial sums of the spectral pass
    size_t partials_cap = 0;
    void *twiddle = nullptr;              // N complex twiddles exp(-2 pi i k / N)
    wt::DeviceScalars *scal = nullptr;
    double *E_w = nullptr, *E_u = nullptr;   // device spectra [nek]
    int nek = 0;
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
};

namespace wt {
// kernel timing slots reported by wt_stage_times after the four stage totals
enum : int {
    kSlotTheta = 0, kSlotSubpoints, kSlotBinCount, kSlotDeposit, kSlotFixedToFloat,
    kSlotFftX, kSlotFftY, kSlotFftZ, kSlotFftYInv, kSlotFftXInv, kSlotReduce, kNumSlots
};
inline void kt_begin(wt_ctx *c, int slot) {
    if (c->kev[slot][0] && hipEventRecord(c->kev[slot][0], c->stream) == hipSuccess) c->kvalid[slot] = false;
}
inline void kt_end(wt_ctx *c, int slot) {
    if (c->kev[slot][1] && hipEventRecord(c->kev[slot][1], c->stream) == hipSuccess) c->kvalid[slot] = true;
}
// segments.hip
hipError_t launch_segments(wt_ctx *c);       // points -> sections -> sub-points (needs sp_base alloc inside)
hipError_t launch_bins(wt_ctx *c);           // per sub-point boxes -> work items
hipError_t launch_get_bins(wt_ctx *c, int32_t *d_lohi, double *d_R);
// deposit.hip
hipError_t launch_deposit(wt_ctx *c);
// fft.hip
hipError_t launch_biot_savart(wt_ctx *c);
// stats.hip
hipError_t launch_sf(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps,
                     double *d_out);
// scan.hip
hipError_t exclusive_scan_i64(const int64_t *d_in, int64_t *d_out, int64_t n, hipStream_t s,
                              void **tmp, size_t *tmp_bytes);
hipError_t exclusive_scan_i32_to_i64(const int32_t *d_in, int add, int64_t *d_out, int64_t n,
                                     hipStream_t s, void **tmp, size_t *tmp_bytes);
int64_t nek_of(int N);
}  // namespace wt

#define WT_HIP(expr)                                   \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return _e;               \
    } while (0)
