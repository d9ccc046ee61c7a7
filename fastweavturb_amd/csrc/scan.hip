// Exclusive prefix sums (int64) used to lay out sub-points and deposition work items.
// out has n+1 entries: out[0] = 0, out[i+1] = out[i] + in[i].
#include <algorithm>

#include "wt_internal.h"

namespace wt {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanThreads * kScanPer;

template <typename In>
__device__ __forceinline__ int64_t load_in(const In *in, int64_t i, int64_t n, int add) {
    return i < n ? (int64_t)in[i] + add : 0;
}

// Block-local exclusive scan of one tile; writes per-element prefixes and the tile total.
template <typename In>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const In *__restrict__ in, int add, int64_t n,
                                                             int64_t *__restrict__ out, int64_t *__restrict__ sums) {
    __shared__ int64_t s[kScanThreads];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int64_t v[kScanPer];
    int64_t t = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        v[k] = load_in(in, base + k, n, add);
        t += v[k];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int64_t x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    int64_t run = s[threadIdx.x] - t;  // exclusive prefix of this thread within the tile
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_add(int64_t *__restrict__ out, int64_t n, const int64_t *__restrict__ pref) {
    int64_t i = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
    int64_t off = pref[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanPer; k++, i += kScanThreads)
        if (i < n) out[i] += off;
}

// Zeroing kernel used instead of hipMemsetAsync, so that stages recorded into a hipGraph hold no
// memset nodes.
__global__ void k_zero_words(uint32_t *__restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

hipError_t fill_zero(void *p, size_t bytes, hipStream_t st) {
    if (bytes % 4) return hipErrorInvalidValue;
    const int64_t n = (int64_t)(bytes / 4);
    if (n == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t *)p, n);
    return hipGetLastError();
}

// Tile sums of every level of the recursive scan of n values, laid out level after level in one
// workspace of scan_ws_elems(n) int64.
int64_t scan_ws_elems(int64_t n) {
    int64_t tot = 0;
    int64_t m = n > 0 ? n : 1;
    for (;;) {
        const int64_t tiles = std::max<int64_t>(1, (m + kScanTile - 1) / kScanTile);
        tot += tiles + 1;
        if (tiles == 1) break;
        m = tiles;
    }
    return tot;
}

// Recursive scan over int64 values already in `buf` (n values) -> exclusive in place,
// total written to *total; the tile sums go to ws (scan_ws_elems(n) entries).
static hipError_t scan_level(int64_t *buf, int64_t n, int64_t *total, hipStream_t st, int64_t *ws) {
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles < 1) tiles = 1;
    int64_t *sums = ws;
    hipLaunchKernelGGL(k_scan_tiles<int64_t>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st,
                       (const int64_t *)buf, 0, n, buf, sums);
    WT_HIP(hipGetLastError());
    if (tiles == 1) {
        WT_HIP(hipMemcpyAsync(total, sums, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
        WT_HIP(fill_zero(sums, sizeof(int64_t), st));
    } else {
        WT_HIP(scan_level(sums, tiles, total, st, ws + tiles + 1));
    }
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, buf, n, sums);
    return hipGetLastError();
}

template <typename In>
static hipError_t scan_impl(const In *d_in, int add, int64_t *d_out, int64_t n, hipStream_t st, int64_t *ws) {
    if (n <= 0) return fill_zero(d_out, sizeof(int64_t), st);
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    int64_t *sums = ws;
    hipLaunchKernelGGL(k_scan_tiles<In>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, d_in, add, n,
                       d_out, sums);
    WT_HIP(hipGetLastError());
    // scan the tile sums in place; the grand total lands in d_out[n]
    WT_HIP(scan_level(sums, tiles, d_out + n, st, ws + tiles + 1));
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, d_out, n, sums);
    return hipGetLastError();
}

// The workspace is the caller's (scan_ws_elems(n) + scan_ws_elems(tiles of n) entries suffice:
// scan_scratch_elems), so a scan allocates nothing and can be recorded into a graph.
int64_t scan_scratch_elems(int64_t n) {
    const int64_t tiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
    return tiles + 1 + scan_ws_elems(tiles);
}

hipError_t exclusive_scan_i64(const int64_t *d_in, int64_t *d_out, int64_t n, hipStream_t s, int64_t *ws) {
    return scan_impl<int64_t>(d_in, 0, d_out, n, s, ws);
}

hipError_t exclusive_scan_i32_to_i64(const int32_t *d_in, int add, int64_t *d_out, int64_t n, hipStream_t s,
                                     int64_t *ws) {
    return scan_impl<int32_t>(d_in, add, d_out, n, s, ws);
}

}  // namespace wt
