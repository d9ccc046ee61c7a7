// Exclusive prefix sums (int64) used to lay out sub-points and deposition work items.
// out has n+1 entries: out[0] = 0, out[i+1] = out[i] + in[i].
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

// Recursive scan over int64 values already in `buf` (n values) -> exclusive in place,
// total written to *total.
static hipError_t scan_level(int64_t *buf, int64_t n, int64_t *total, hipStream_t st) {
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles < 1) tiles = 1;
    int64_t *sums = nullptr;
    WT_HIP(hipMallocAsync((void **)&sums, sizeof(int64_t) * (tiles + 1), st));
    hipLaunchKernelGGL(k_scan_tiles<int64_t>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st,
                       (const int64_t *)buf, 0, n, buf, sums);
    WT_HIP(hipGetLastError());
    if (tiles == 1) {
        WT_HIP(hipMemcpyAsync(total, sums, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
        WT_HIP(hipMemsetAsync(sums, 0, sizeof(int64_t), st));
    } else {
        WT_HIP(scan_level(sums, tiles, total, st));
    }
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, buf, n, sums);
    WT_HIP(hipGetLastError());
    WT_HIP(hipFreeAsync(sums, st));
    return hipSuccess;
}

template <typename In>
static hipError_t scan_impl(const In *d_in, int add, int64_t *d_out, int64_t n, hipStream_t st) {
    if (n <= 0) return hipMemsetAsync(d_out, 0, sizeof(int64_t), st);
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    int64_t *sums = nullptr;
    WT_HIP(hipMallocAsync((void **)&sums, sizeof(int64_t) * (tiles + 1), st));
    hipLaunchKernelGGL(k_scan_tiles<In>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, d_in, add, n,
                       d_out, sums);
    WT_HIP(hipGetLastError());
    // scan the tile sums in place; the grand total lands in d_out[n]
    WT_HIP(scan_level(sums, tiles, d_out + n, st));
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, d_out, n, sums);
    WT_HIP(hipGetLastError());
    WT_HIP(hipFreeAsync(sums, st));
    return hipSuccess;
}

hipError_t exclusive_scan_i64(const int64_t *d_in, int64_t *d_out, int64_t n, hipStream_t s, void **,
                              size_t *) {
    return scan_impl<int64_t>(d_in, 0, d_out, n, s);
}

hipError_t exclusive_scan_i32_to_i64(const int32_t *d_in, int add, int64_t *d_out, int64_t n, hipStream_t s,
                                     void **, size_t *) {
    return scan_impl<int32_t>(d_in, add, d_out, n, s);
}

}  // namespace wt
