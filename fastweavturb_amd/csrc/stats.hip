// Structure functions of the resident velocity (subroutine.f90:64-129, 131-162).
//
// mode 0 (Sp_calculate): S_p(r) = (1/3N^3) sum over all points of
//   (ux(i+r,j,k)-ux)^p + (uy(i,j+r,k)-uy)^p + (uz(i,j,k+r)-uz)^p, p = 1..pmax.
// mode 1 (x-only, getStructureFunc_x generalised): S_p(r) = (1/N^3) sum (ux(i+r)-ux(i))^p,
//   p = pmin..pmax, x-lines staged in LDS, one separation per thread.
// Increments are formed in the field's precision; powers and sums in FP64. Blocks write one
// partial row each and the rows are reduced in a fixed order, so results are reproducible.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "wt_internal.h"

namespace wt {

constexpr int kSfThreads = 256;
constexpr int kMaxOrder = 12;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds the block-wide sum of acc[0..np) into lds_acc[0..np) (called by all threads).
__device__ __forceinline__ void block_accumulate(const double *acc, int np, double *lds_acc, double *red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int p = 0; p < np; p++) {
        double v = wave_sum(acc[p]);
        if (lane == 0) red[wv * kMaxOrder + p] = v;
    }
    __syncthreads();
    if (threadIdx.x < np) {
        double s = 0.0;
        for (int w = 0; w < kSfThreads / 64; w++) s += red[w * kMaxOrder + threadIdx.x];
        lds_acc[threadIdx.x] += s;
    }
    __syncthreads();
}

// x-increments with one separation per thread: thread t takes separation m = t % nsep and point
// phase g = t / nsep (G = blockDim / nsep phases), accumulates the orders' power sums of its
// separation over the points x = g, g + G, ... of every staged line of every group the block
// takes, in registers, and the phases are summed once per block in a fixed order. Lanes of a wave
// read the same base point (an LDS broadcast) and consecutive shifted points.
template <typename T, int NP>
__device__ __forceinline__ void sf_accumulate_n(const T *line, int nl, int N, int s, int g, int G, int pmin,
                                                double *acc) {
    for (int li = 0; li < nl; li++) {
        const T *ln = line + li * N;
        for (int x = g; x < N; x += G) {
            int xs = x + s;
            if (xs >= N) xs -= N;
            const double d = (double)ln[xs] - (double)ln[x];
            double pw = 1.0;
            for (int p = 1; p < pmin; p++) pw *= d;
#pragma unroll
            for (int p = 0; p < NP; p++) {
                pw *= d;
                acc[p] += pw;
            }
        }
    }
}

// The order count as a compile-time constant, so that only the requested powers are formed.
template <typename T>
__device__ __forceinline__ void sf_accumulate(const T *line, int nl, int N, int s, int g, int G, int pmin, int np,
                                              double *acc) {
    switch (np) {
        case 1: sf_accumulate_n<T, 1>(line, nl, N, s, g, G, pmin, acc); break;
        case 2: sf_accumulate_n<T, 2>(line, nl, N, s, g, G, pmin, acc); break;
        case 3: sf_accumulate_n<T, 3>(line, nl, N, s, g, G, pmin, acc); break;
        case 4: sf_accumulate_n<T, 4>(line, nl, N, s, g, G, pmin, acc); break;
        case 5: sf_accumulate_n<T, 5>(line, nl, N, s, g, G, pmin, acc); break;
        case 6: sf_accumulate_n<T, 6>(line, nl, N, s, g, G, pmin, acc); break;
        case 7: sf_accumulate_n<T, 7>(line, nl, N, s, g, G, pmin, acc); break;
        case 8: sf_accumulate_n<T, 8>(line, nl, N, s, g, G, pmin, acc); break;
        case 9: sf_accumulate_n<T, 9>(line, nl, N, s, g, G, pmin, acc); break;
        case 10: sf_accumulate_n<T, 10>(line, nl, N, s, g, G, pmin, acc); break;
        case 11: sf_accumulate_n<T, 11>(line, nl, N, s, g, G, pmin, acc); break;
        default: sf_accumulate_n<T, 12>(line, nl, N, s, g, G, pmin, acc); break;
    }
}

// Sums the G phases' accumulators into out[m*np + p] (all threads call it).
__device__ __forceinline__ void sf_reduce_phases(const double *acc, double *red, int nsep, int np, int m, int g,
                                                 bool active, int G, double *out) {
    if (active)
#pragma unroll
        for (int p = 0; p < kMaxOrder; p++)
            if (p < np) red[((size_t)g * nsep + m) * np + p] = acc[p];
    __syncthreads();
    for (int i = threadIdx.x; i < nsep * np; i += blockDim.x) {
        double v = 0.0;
        for (int q = 0; q < G; q++) v += red[(size_t)q * nsep * np + i];
        out[i] = v;
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(1024) void k_sf_x_sep(const T *__restrict__ ux, int N, int64_t nlines, int lpg, int pmin,
                                                   int np, int nsep, const int32_t *__restrict__ seps, int row_stride,
                                                   int col0, double *__restrict__ partials) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *red = (double *)smem_raw;                                   // G * nsep * np
    const int G = (int)blockDim.x / nsep > 0 ? (int)blockDim.x / nsep : 1;
    T *line = (T *)(red + (size_t)G * nsep * np);                      // lpg * N
    const int m = threadIdx.x % nsep, g = threadIdx.x / nsep;
    const bool active = g < G;
    const int s = ((seps[m] % N) + N) % N;
    double acc[kMaxOrder];
#pragma unroll
    for (int p = 0; p < kMaxOrder; p++) acc[p] = 0.0;
    const int64_t ngroups = (nlines + lpg - 1) / lpg;
    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const int64_t l0 = grp * lpg;
        const int nl = (int)((nlines - l0) < lpg ? (nlines - l0) : lpg);
        for (int i = threadIdx.x; i < nl * N; i += blockDim.x) line[i] = ux[l0 * N + i];
        __syncthreads();
        if (active) sf_accumulate(line, nl, N, s, g, G, pmin, np, acc);
        __syncthreads();
    }
    sf_reduce_phases(acc, red, nsep, np, m, g, active, G, partials + (int64_t)blockIdx.x * row_stride + col0);
}

template <typename T>
__global__ __launch_bounds__(kSfThreads) void k_sf_long(const T *__restrict__ u, int N, int np, int nsep,
                                                        const int32_t *__restrict__ seps, double *__restrict__ partials) {
    extern __shared__ unsigned char smem_raw[];
    double *lds_acc = (double *)smem_raw;                       // nsep * np
    double *red = lds_acc + (size_t)nsep * np;
    for (int i = threadIdx.x; i < nsep * np; i += blockDim.x) lds_acc[i] = 0.0;
    __syncthreads();
    const int64_t N2 = (int64_t)N * N, N3 = N2 * N;
    const T *ux = u, *uy = u + N3, *uz = u + 2 * N3;
    const int64_t rows = N2;                       // (j, k) rows of N points
    const int rows_per_group = kSfThreads * 8 / N > 0 ? kSfThreads * 8 / N : 1;
    const int64_t ngroups = (rows + rows_per_group - 1) / rows_per_group;
    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        int64_t r0 = grp * rows_per_group;
        int nr = (int)((rows - r0) < rows_per_group ? (rows - r0) : rows_per_group);
        for (int m = 0; m < nsep; m++) {
            int s = ((seps[m] % N) + N) % N;
            double acc[kMaxOrder];
            for (int p = 0; p < np; p++) acc[p] = 0.0;
            for (int t = threadIdx.x; t < nr * N; t += blockDim.x) {
                int64_t row = r0 + t / N;
                int i = t % N;
                int j = (int)(row % N), k = (int)(row / N);
                int64_t id = row * N + i;
                int is = i + s; if (is >= N) is -= N;
                int js = j + s; if (js >= N) js -= N;
                int ks = k + s; if (ks >= N) ks -= N;
                double dx = (double)ux[row * N + is] - (double)ux[id];
                double dy = (double)uy[((int64_t)k * N + js) * N + i] - (double)uy[id];
                double dz = (double)uz[((int64_t)ks * N + j) * N + i] - (double)uz[id];
                double px = 1.0, py = 1.0, pz = 1.0;
                for (int p = 0; p < np; p++) {
                    px *= dx;
                    py *= dy;
                    pz *= dz;
                    acc[p] += (px + py) + pz;
                }
            }
            block_accumulate(acc, np, lds_acc + (size_t)m * np, red);
        }
    }
    double *row = partials + (int64_t)blockIdx.x * nsep * np;
    for (int i = threadIdx.x; i < nsep * np; i += blockDim.x) row[i] = lds_acc[i];
}

__global__ void k_reduce_rows(const double *__restrict__ partials, int64_t rows, int cols, double scale,
                              double *__restrict__ out) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= cols) return;
    double s = 0.0;
    for (int64_t r = 0; r < rows; r++) s += partials[r * cols + col];
    out[col] = s * scale;
}

// Same sums as k_reduce_rows (rows added in order) written in the output layout: partial column
// m * np + p goes to out[p * nsep + m].
__global__ void k_reduce_rows_sf(const double *__restrict__ partials, int64_t rows, int nsep, int np, double scale,
                                 double *__restrict__ out) {
    const int cols = nsep * np;
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= cols) return;
    double s = 0.0;
    for (int64_t r = 0; r < rows; r++) s += partials[r * cols + col];
    const int m = col / np, p = col - m * np;
    out[(int64_t)p * nsep + m] = s * scale;
}

__global__ void k_fill_zero(double *__restrict__ a, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0.0;
}

// Partial rows (blocks) of the structure-function launch of an N^3 field.
static int64_t sf_rows(const wt_ctx *c, int mode, int N) {
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cus = prop.multiProcessorCount;
    if (mode == 1) return (int64_t)cus * 2;
    const int rpg = kSfThreads * 8 / N > 0 ? kSfThreads * 8 / N : 1;
    const int64_t ngroups = ((int64_t)N * N + rpg - 1) / rpg;
    return ngroups < (int64_t)cus * 4 ? ngroups : (int64_t)cus * 4;
}

// Structure functions of instance c->cur into d_out[(p - pmin) * nsep + m] (device memory, the
// caller's), with `partials` (device, rows * nsep * np doubles, rows = sf_rows) as the per-block
// partial sums. Stream-ordered on the context stream; allocates nothing.
template <typename T>
static hipError_t sf_core(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps, double *partials,
                          double *d_out) {
    hipStream_t st = c->stream;
    const wt::Instance &I = c->inst[c->cur];
    const int N = I.N;
    const int np = pmax - pmin + 1;
    int dev_lds = 65536;
    (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    const T *u = (const T *)c->u + I.field_off;
    size_t acc_bytes = sizeof(double) * ((size_t)nsep * np + 4 * kMaxOrder);
    const int64_t blocks = sf_rows(c, mode, N);
    const int cols = nsep * np;
    if (mode == 1) {
        // separations in chunks of at most 1024, one thread per separation (and point phase); rows
        // of blocks beyond a chunk's group count stay zero
        hipLaunchKernelGGL(k_fill_zero, dim3(256), dim3(256), 0, st, partials, blocks * cols);
        WT_HIP(hipGetLastError());
        const int64_t nlines = (int64_t)N * N;
        for (int s0 = 0; s0 < nsep; s0 += 1024) {
            const int nc = nsep - s0 < 1024 ? nsep - s0 : 1024;
            const int threads = nc <= 256 ? 256 : (nc + 63) / 64 * 64;
            const int G = threads / nc;
            const size_t red = sizeof(double) * (size_t)G * nc * np;
            int lpg = 8192 / N > 0 ? 8192 / N : 1;
            while (lpg > 1 && red + (size_t)lpg * N * sizeof(T) > (size_t)dev_lds) lpg /= 2;
            const size_t sh = red + (size_t)lpg * N * sizeof(T);
            if (sh > (size_t)dev_lds) return hipErrorInvalidValue;
            const int64_t ngroups = (nlines + lpg - 1) / lpg;
            const int64_t b = ngroups < blocks ? ngroups : blocks;
            WT_HIP(hipFuncSetAttribute((const void *)k_sf_x_sep<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
            hipLaunchKernelGGL(k_sf_x_sep<T>, dim3((unsigned)b), dim3(threads), sh, st, u, N, nlines, lpg, pmin, np, nc,
                               d_seps + s0, nsep * np, s0 * np, partials);
            WT_HIP(hipGetLastError());
        }
    } else {
        if (acc_bytes > (size_t)dev_lds) return hipErrorInvalidValue;
        WT_HIP(hipFuncSetAttribute((const void *)k_sf_long<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)acc_bytes));
        hipLaunchKernelGGL(k_sf_long<T>, dim3((unsigned)blocks), dim3(kSfThreads), acc_bytes, st, u, N, np, nsep,
                           d_seps, partials);
    }
    WT_HIP(hipGetLastError());
    const double N3 = (double)N * N * N;
    const double scale = mode == 0 ? 1.0 / (3.0 * N3) : 1.0 / N3;
    hipLaunchKernelGGL(k_reduce_rows_sf, dim3((unsigned)((cols + 63) / 64)), dim3(64), 0, st, (const double *)partials,
                       blocks, nsep, np, scale, d_out);
    return hipGetLastError();
}

template <typename T>
static hipError_t sf_t(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps, double *d_out) {
    hipStream_t st = c->stream;
    const int64_t rows = sf_rows(c, mode, c->inst[c->cur].N);
    double *partials = nullptr;
    WT_HIP(hipMallocAsync((void **)&partials, sizeof(double) * rows * nsep * (pmax - pmin + 1), st));
    hipError_t e = sf_core<T>(c, mode, pmin, pmax, nsep, d_seps, partials, d_out);
    const hipError_t f = hipFreeAsync(partials, st);
    if (e == hipSuccess) e = f;
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

// ---- allocation-free variant -----------------------------------------------------------------
// The per-block partial sums of wt_structure_functions_noalloc live in this file-scope array, one
// copy per GPU, held until the process ends. Host code never passes `scratch` itself to a kernel
// (on the host it names a stand-in, not device memory): sf_scratch_init resolves each device's
// address with hipGetSymbolAddress once, and the kernels get that address.
constexpr int64_t kSfScratchDoubles = (int64_t)4 << 20;     // 32 MiB per GPU
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kSfScratchDoubles];
static double *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Called with `device` current (wt_create), once per device.
hipError_t sf_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    WT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (double *)p;
    return hipSuccess;
}

int64_t sf_noalloc_max_cols(const wt_ctx *c, int mode) {
    return kSfScratchDoubles / sf_rows(c, mode, c->inst[c->cur].N);
}

hipError_t launch_sf_noalloc(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps,
                             double *d_out) {
    if (pmax - pmin + 1 > kMaxOrder || pmin < 1) return hipErrorInvalidValue;
    double *partials = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c->device >= 0 && c->device < kMaxDevices) partials = g_scratch_ptr[c->device];
    }
    if (!partials) return hipErrorNotInitialized;
    const int64_t rows = sf_rows(c, mode, c->inst[c->cur].N);
    if (rows * nsep * (pmax - pmin + 1) > kSfScratchDoubles) return hipErrorInvalidValue;
    hipError_t e = c->precision == 64 ? sf_core<double>(c, mode, pmin, pmax, nsep, d_seps, partials, d_out)
                                      : sf_core<float>(c, mode, pmin, pmax, nsep, d_seps, partials, d_out);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

// ---- every instance of a batch in one launch ------------------------------------------------
// Work item = up to 8192 points of x-lines of one instance; items of an instance are contiguous,
// each writes one partial row (separation-major, order-minor), and the rows of an instance are
// summed in item order.
struct SfItem {
    int32_t pos, nl;      // arena position, lines in the item
    int64_t line0;        // first line within the instance
};
struct SfInst {
    int32_t N, pad;
    int64_t field_off;
};
constexpr int kSfItemPoints = 8192;

template <typename T>
__global__ __launch_bounds__(1024) void k_sf_x_batch(const T *__restrict__ u, const SfInst *__restrict__ inst,
                                                     const SfItem *__restrict__ items, int64_t nitems, int pmin,
                                                     int np, int nsep, const int32_t *__restrict__ seps,
                                                     double *__restrict__ partials) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *red = (double *)smem_raw;                                   // G * nsep * np
    const int G = (int)blockDim.x / nsep > 0 ? (int)blockDim.x / nsep : 1;
    T *line = (T *)(red + (size_t)G * nsep * np);                      // kSfItemPoints
    const int m = threadIdx.x % nsep, g = threadIdx.x / nsep;
    const bool active = g < G;
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const SfItem w = items[it];
        const SfInst I = inst[w.pos];
        const int N = I.N;
        const T *src = u + I.field_off + w.line0 * N;             // u_x: first component
        for (int i = threadIdx.x; i < w.nl * N; i += blockDim.x) line[i] = src[i];
        __syncthreads();
        double acc[kMaxOrder];
#pragma unroll
        for (int p = 0; p < kMaxOrder; p++) acc[p] = 0.0;
        if (active) {
            const int s = ((seps[(int64_t)w.pos * nsep + m] % N) + N) % N;
            sf_accumulate(line, w.nl, N, s, g, G, pmin, np, acc);
        }
        sf_reduce_phases(acc, red, nsep, np, m, g, active, G, partials + it * nsep * np);
    }
}

// out[(pos*np + p)*nsep + m] = sum over the position's items of column m*np + p, / N^3.
__global__ void k_sf_reduce_items(const double *__restrict__ partials, const int64_t *__restrict__ ibase,
                                  const SfInst *__restrict__ inst, int np, int nsep, double *__restrict__ out) {
    const int pos = blockIdx.y;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int cols = nsep * np;
    if (col >= cols) return;
    double s = 0.0;
    for (int64_t r = ibase[pos]; r < ibase[pos + 1]; r++) s += partials[r * cols + col];
    const double N = (double)inst[pos].N;
    const int m = col / np, p = col - m * np;
    out[((int64_t)pos * np + p) * nsep + m] = s / (N * N * N);
}

template <typename T>
static hipError_t sf_batch_t(wt_ctx *c, int pmin, int pmax, int nsep, const int32_t *d_seps, double *d_out) {
    hipStream_t st = c->stream;
    const int np = pmax - pmin + 1;
    const int npos = c->ninst();
    // work items, built once per context (the layout is fixed at creation)
    if (!c->sf_items) {
        std::vector<SfItem> items;
        std::vector<int64_t> ibase(npos + 1, 0);
        std::vector<SfInst> inst(npos);
        for (int pos = 0; pos < npos; pos++) {
            const wt::Instance &I = c->inst[c->order[pos]];
            inst[pos] = SfInst{I.N, 0, I.field_off};
            const int lpg = kSfItemPoints / I.N > 0 ? kSfItemPoints / I.N : 1;
            const int64_t nlines = (int64_t)I.N * I.N;
            ibase[pos] = (int64_t)items.size();
            for (int64_t l0 = 0; l0 < nlines; l0 += lpg)
                items.push_back(SfItem{pos, (int32_t)std::min<int64_t>(lpg, nlines - l0), l0});
        }
        ibase[npos] = (int64_t)items.size();
        const size_t bytes = sizeof(SfItem) * items.size() + sizeof(int64_t) * (npos + 1) + sizeof(SfInst) * npos;
        void *d = nullptr;
        WT_HIP(hipMalloc(&d, bytes));
        unsigned char *b = (unsigned char *)d;
        hipError_t e = hipMemcpy(b, items.data(), sizeof(SfItem) * items.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b + sizeof(SfItem) * items.size(), ibase.data(), sizeof(int64_t) * (npos + 1),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b + sizeof(SfItem) * items.size() + sizeof(int64_t) * (npos + 1), inst.data(),
                          sizeof(SfInst) * npos, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return e;
        }
        c->sf_items = d;
        c->sf_nitems = (int64_t)items.size();
    }
    const SfItem *d_items = (const SfItem *)c->sf_items;
    const int64_t *d_ibase = (const int64_t *)(d_items + c->sf_nitems);
    const SfInst *d_inst = (const SfInst *)(d_ibase + npos + 1);
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    int dev_lds = 65536;
    (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    if (nsep > 1024) return hipErrorInvalidValue;
    const int threads = nsep <= 256 ? 256 : (nsep + 63) / 64 * 64;
    const size_t sh = sizeof(double) * (size_t)(threads / nsep) * nsep * np + sizeof(T) * kSfItemPoints;
    if (sh > (size_t)dev_lds) return hipErrorInvalidValue;
    const int64_t blocks = std::min<int64_t>(c->sf_nitems, (int64_t)cus * 4);
    double *partials = nullptr;
    WT_HIP(hipMallocAsync((void **)&partials, sizeof(double) * c->sf_nitems * nsep * np, st));
    WT_HIP(hipFuncSetAttribute((const void *)k_sf_x_batch<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    hipLaunchKernelGGL(k_sf_x_batch<T>, dim3((unsigned)blocks), dim3(threads), sh, st, (const T *)c->u, d_inst,
                       d_items, c->sf_nitems, pmin, np, nsep, d_seps, partials);
    WT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sf_reduce_items, dim3((unsigned)((nsep * np + 63) / 64), (unsigned)npos), dim3(64), 0, st,
                       (const double *)partials, d_ibase, d_inst, np, nsep, d_out);
    WT_HIP(hipGetLastError());
    return hipFreeAsync(partials, st);
}

hipError_t launch_sf_batch(wt_ctx *c, int pmin, int pmax, int nsep, const int32_t *d_seps, double *d_out) {
    if (pmax - pmin + 1 > kMaxOrder || pmin < 1) return hipErrorInvalidValue;
    return c->precision == 64 ? sf_batch_t<double>(c, pmin, pmax, nsep, d_seps, d_out)
                              : sf_batch_t<float>(c, pmin, pmax, nsep, d_seps, d_out);
}

// ---- PDFs (cal_pdf, subroutine.f90:1-40) -----------------------------------------------------
// Values are taken as FP32 (the reference copies the fields to real(4), tube.f90:1169-1176).
template <typename T>
__global__ __launch_bounds__(256) void k_minmax(const T *__restrict__ x, int64_t n, float *__restrict__ out) {
    __shared__ float red[2][4];
    float mn = __builtin_huge_valf(), mx = -__builtin_huge_valf();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = (float)x[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, o, 64));
        mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = mn;
        red[1][threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            mn = fminf(mn, red[0][w]);
            mx = fmaxf(mx, red[1][w]);
        }
        out[2 * blockIdx.x] = red[0][0] < mn ? red[0][0] : mn;
        out[2 * blockIdx.x + 1] = red[1][0] > mx ? red[1][0] : mx;
    }
}

// Each value goes to the first bin l with lo[l] <= v < hi[l] (the reference's linear search).
// Both edge tables are non-decreasing in l (every step that forms them rounds monotonically), so
// that bin is the smallest L with hi[L] > v, provided lo[L] <= v, and no bin matches otherwise.
// L is found by walking from the estimate (v - mn) / bs, which is within a step or two of it unless
// bs is below the rounding unit of the values. Per-block counts in LDS, then one atomic per bin and
// block.
template <typename T>
__global__ __launch_bounds__(256) void k_pdf_hist(const T *__restrict__ x, int64_t n, int nb, float mn, float bs,
                                                  const float *__restrict__ edges, unsigned long long *__restrict__ counts) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *lo = (float *)smem_raw, *hi = lo + nb;
    unsigned int *h = (unsigned int *)(hi + nb);
    for (int l = threadIdx.x; l < nb; l += blockDim.x) {
        lo[l] = edges[l];
        hi[l] = edges[nb + l];
        h[l] = 0u;
    }
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = (float)x[i];
        const float t = bs > 0.0f ? (v - mn) / bs : 0.0f;
        int l = !(t >= 0.0f) ? 0 : (t >= (float)nb ? nb - 1 : (int)t);
        while (l > 0 && hi[l - 1] > v) l--;
        while (l < nb && !(hi[l] > v)) l++;
        if (l < nb && lo[l] <= v) atomicAdd(&h[l], 1u);
    }
    __syncthreads();
    for (int l = threadIdx.x; l < nb; l += blockDim.x)
        if (h[l]) atomicAdd(&counts[l], (unsigned long long)h[l]);
}

template <typename T>
static hipError_t pdf_t(wt_ctx *c, int field, int nb, float *val, float *pdf, int64_t *counts_out) {
    hipStream_t st = c->stream;
    const wt::Instance &I = c->inst[c->cur];
    const int N = I.N;
    const int64_t n = (int64_t)3 * N * N * N;
    const T *x = (const T *)(field == 0 ? c->omega : c->u) + I.field_off;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    const int blocks = cus * 4;
    float *d_mm = nullptr, *d_edges = nullptr;
    unsigned long long *d_cnt = nullptr;
    WT_HIP(hipMallocAsync((void **)&d_mm, sizeof(float) * 2 * blocks, st));
    hipError_t e = hipSuccess;
    std::vector<float> mm(2 * blocks);
    hipLaunchKernelGGL(k_minmax<T>, dim3(blocks), dim3(256), 0, st, x, n, d_mm);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(mm.data(), d_mm, sizeof(float) * 2 * blocks, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(d_mm, st);
    if (e != hipSuccess) return e;
    float mn = mm[0], mx = mm[1];
    for (int b = 1; b < blocks; b++) {
        mn = fminf(mn, mm[2 * b]);
        mx = fmaxf(mx, mm[2 * b + 1]);
    }
    // subroutine.f90:16-20: bin_size = (max - min) / real(bin_num);
    // val(l) = min + 0.5d0*bin_size + real(l-1)*bin_size (double sum of a real(4) product, stored
    // real(4)); the interval test uses val(l) -/+ 0.5*bin_size in real(4) (:27)
    const float bs = (mx - mn) / (float)nb;
    std::vector<float> edges(2 * (size_t)nb);
    for (int l = 0; l < nb; l++) {
        const float prod = (float)l * bs;
        val[l] = (float)(((double)mn + 0.5 * (double)bs) + (double)prod);
        edges[l] = val[l] - 0.5f * bs;
        edges[nb + l] = val[l] + 0.5f * bs;
    }
    WT_HIP(hipMallocAsync((void **)&d_edges, sizeof(float) * 2 * nb, st));
    e = hipMallocAsync((void **)&d_cnt, sizeof(unsigned long long) * nb, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_edges, edges.data(), sizeof(float) * 2 * nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long) * nb, st);
    const size_t sh = (sizeof(float) * 2 + sizeof(unsigned int)) * (size_t)nb;
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_pdf_hist<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_pdf_hist<T>, dim3(blocks), dim3(256), sh, st, x, n, nb, mn, bs, (const float *)d_edges,
                           d_cnt);
        e = hipGetLastError();
    }
    std::vector<unsigned long long> cnt(nb);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_cnt, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_edges) (void)hipFreeAsync(d_edges, st);
    if (d_cnt) (void)hipFreeAsync(d_cnt, st);
    if (e != hipSuccess) return e;
    // subroutine.f90:37-39: pdf_v(l) / (bin_size*nx*ny*nzp*nproc) in real(4), nzp = 3 N (the three
    // components stacked along z), nproc = 1
    const float denom = (((bs * (float)N) * (float)N) * (float)(3 * N)) * 1.0f;
    for (int l = 0; l < nb; l++) {
        pdf[l] = (float)cnt[l] / denom;
        if (counts_out) counts_out[l] = (int64_t)cnt[l];
    }
    return hipSuccess;
}

hipError_t launch_pdf(wt_ctx *c, int field, int nbins, float *val, float *pdf, int64_t *counts) {
    return c->precision == 64 ? pdf_t<double>(c, field, nbins, val, pdf, counts)
                              : pdf_t<float>(c, field, nbins, val, pdf, counts);
}

hipError_t launch_sf(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps, double *d_out) {
    if (pmax - pmin + 1 > kMaxOrder || pmin < 1) return hipErrorInvalidValue;
    return c->precision == 64 ? sf_t<double>(c, mode, pmin, pmax, nsep, d_seps, d_out)
                              : sf_t<float>(c, mode, pmin, pmax, nsep, d_seps, d_out);
}

}  // namespace wt
