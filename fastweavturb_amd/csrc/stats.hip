// Structure functions of the resident velocity (subroutine.f90:64-129, 131-162).
//
// mode 0 (Sp_calculate): S_p(r) = (1/3N^3) sum over all points of
//   (ux(i+r,j,k)-ux)^p + (uy(i,j+r,k)-uy)^p + (uz(i,j,k+r)-uz)^p, p = 1..pmax.
// mode 1 (x-only, getStructureFunc_x generalised): S_p(r) = (1/N^3) sum (ux(i+r)-ux(i))^p,
//   p = pmin..pmax, x-lines staged in LDS.
// Increments are formed in the field's precision; powers and sums in FP64. Persistent blocks
// accumulate per-(separation, order) sums in LDS and write one partial row each; the rows are
// reduced in a fixed order, so results are reproducible.
#include <algorithm>
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
__device__ __forceinline__ void block_accumulate(const double *acc, int np, double *lds_acc, double *scratch) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int p = 0; p < np; p++) {
        double v = wave_sum(acc[p]);
        if (lane == 0) scratch[wv * kMaxOrder + p] = v;
    }
    __syncthreads();
    if (threadIdx.x < np) {
        double s = 0.0;
        for (int w = 0; w < kSfThreads / 64; w++) s += scratch[w * kMaxOrder + threadIdx.x];
        lds_acc[threadIdx.x] += s;
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(kSfThreads) void k_sf_x(const T *__restrict__ ux, int N, int64_t nlines, int lines_per_group,
                                                     int pmin, int np, int nsep, const int32_t *__restrict__ seps,
                                                     double *__restrict__ partials) {
    extern __shared__ unsigned char smem_raw[];
    double *lds_acc = (double *)smem_raw;                       // nsep * np
    double *scratch = lds_acc + (size_t)nsep * np;               // 4 * kMaxOrder
    T *line = (T *)(scratch + 4 * kMaxOrder);                    // lines_per_group * N
    for (int i = threadIdx.x; i < nsep * np; i += blockDim.x) lds_acc[i] = 0.0;
    __syncthreads();
    const int64_t ngroups = (nlines + lines_per_group - 1) / lines_per_group;
    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        int64_t l0 = grp * lines_per_group;
        int nl = (int)((nlines - l0) < lines_per_group ? (nlines - l0) : lines_per_group);
        for (int i = threadIdx.x; i < nl * N; i += blockDim.x) line[i] = ux[l0 * N + i];
        __syncthreads();
        const int npts = nl * N;
        for (int m = 0; m < nsep; m++) {
            int s = ((seps[m] % N) + N) % N;
            double acc[kMaxOrder];
            for (int p = 0; p < np; p++) acc[p] = 0.0;
            for (int i = threadIdx.x; i < npts; i += blockDim.x) {
                int li = i / N, x = i - li * N;
                int xs = x + s;
                if (xs >= N) xs -= N;
                double d = (double)line[li * N + xs] - (double)line[i];
                double pw = 1.0;
                for (int p = 1; p < pmin; p++) pw *= d;
                for (int p = 0; p < np; p++) {
                    pw *= d;
                    acc[p] += pw;
                }
            }
            block_accumulate(acc, np, lds_acc + (size_t)m * np, scratch);
        }
    }
    double *row = partials + (int64_t)blockIdx.x * nsep * np;
    for (int i = threadIdx.x; i < nsep * np; i += blockDim.x) row[i] = lds_acc[i];
}

template <typename T>
__global__ __launch_bounds__(kSfThreads) void k_sf_long(const T *__restrict__ u, int N, int np, int nsep,
                                                        const int32_t *__restrict__ seps, double *__restrict__ partials) {
    extern __shared__ unsigned char smem_raw[];
    double *lds_acc = (double *)smem_raw;                       // nsep * np
    double *scratch = lds_acc + (size_t)nsep * np;
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
            block_accumulate(acc, np, lds_acc + (size_t)m * np, scratch);
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

template <typename T>
static hipError_t sf_t(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps, double *d_out) {
    hipStream_t st = c->stream;
    const wt::Instance &I = c->inst[c->cur];
    const int N = I.N;
    const int np = pmax - pmin + 1;
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cus = prop.multiProcessorCount;
    int dev_lds = 65536;
    (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    const T *u = (const T *)c->u + I.field_off;
    size_t acc_bytes = sizeof(double) * ((size_t)nsep * np + 4 * kMaxOrder);
    int64_t blocks;
    double *partials = nullptr;
    if (mode == 1) {
        int lpg = 8192 / N > 0 ? 8192 / N : 1;
        while (lpg > 1 && acc_bytes + (size_t)lpg * N * sizeof(T) > (size_t)dev_lds) lpg /= 2;
        size_t sh = acc_bytes + (size_t)lpg * N * sizeof(T);
        if (sh > (size_t)dev_lds) return hipErrorInvalidValue;
        int64_t nlines = (int64_t)N * N;
        int64_t ngroups = (nlines + lpg - 1) / lpg;
        blocks = ngroups < (int64_t)cus * 4 ? ngroups : (int64_t)cus * 4;
        WT_HIP(hipMallocAsync((void **)&partials, sizeof(double) * blocks * nsep * np, st));
        WT_HIP(hipFuncSetAttribute((const void *)k_sf_x<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        hipLaunchKernelGGL(k_sf_x<T>, dim3((unsigned)blocks), dim3(kSfThreads), sh, st, u, N, nlines, lpg, pmin, np,
                           nsep, d_seps, partials);
    } else {
        if (acc_bytes > (size_t)dev_lds) return hipErrorInvalidValue;
        int rpg = kSfThreads * 8 / N > 0 ? kSfThreads * 8 / N : 1;
        int64_t ngroups = ((int64_t)N * N + rpg - 1) / rpg;
        blocks = ngroups < (int64_t)cus * 4 ? ngroups : (int64_t)cus * 4;
        WT_HIP(hipMallocAsync((void **)&partials, sizeof(double) * blocks * nsep * np, st));
        WT_HIP(hipFuncSetAttribute((const void *)k_sf_long<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)acc_bytes));
        hipLaunchKernelGGL(k_sf_long<T>, dim3((unsigned)blocks), dim3(kSfThreads), acc_bytes, st, u, N, np, nsep,
                           d_seps, partials);
    }
    WT_HIP(hipGetLastError());
    double N3 = (double)N * N * N;
    double scale = mode == 0 ? 1.0 / (3.0 * N3) : 1.0 / N3;
    int cols = nsep * np;
    // output layout S[(p - pmin) * nsep + m]; partial columns are m * np + p
    double *tmp = nullptr;
    WT_HIP(hipMallocAsync((void **)&tmp, sizeof(double) * cols, st));
    hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)((cols + 63) / 64)), dim3(64), 0, st, (const double *)partials,
                       blocks, cols, scale, tmp);
    WT_HIP(hipGetLastError());
    std::vector<double> h(cols), o(cols);
    WT_HIP(hipMemcpyAsync(h.data(), tmp, sizeof(double) * cols, hipMemcpyDeviceToHost, st));
    WT_HIP(hipStreamSynchronize(st));
    for (int m = 0; m < nsep; m++)
        for (int p = 0; p < np; p++) o[(size_t)p * nsep + m] = h[(size_t)m * np + p];
    WT_HIP(hipMemcpyAsync(d_out, o.data(), sizeof(double) * cols, hipMemcpyHostToDevice, st));
    WT_HIP(hipFreeAsync(tmp, st));
    WT_HIP(hipFreeAsync(partials, st));
    return hipStreamSynchronize(st);
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
__global__ __launch_bounds__(kSfThreads) void k_sf_x_batch(const T *__restrict__ u, const SfInst *__restrict__ inst,
                                                           const SfItem *__restrict__ items, int64_t nitems, int pmin,
                                                           int np, int nsep, const int32_t *__restrict__ seps,
                                                           double *__restrict__ partials) {
    extern __shared__ unsigned char smem_raw[];
    double *lds_acc = (double *)smem_raw;                       // nsep * np
    double *scratch = lds_acc + (size_t)nsep * np;               // 4 * kMaxOrder
    T *line = (T *)(scratch + 4 * kMaxOrder);                    // kSfItemPoints
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const SfItem w = items[it];
        const SfInst I = inst[w.pos];
        const int N = I.N;
        for (int i = threadIdx.x; i < nsep * np; i += blockDim.x) lds_acc[i] = 0.0;
        const T *src = u + I.field_off + w.line0 * N;             // u_x: first component
        const int npts = w.nl * N;
        for (int i = threadIdx.x; i < npts; i += blockDim.x) line[i] = src[i];
        __syncthreads();
        for (int m = 0; m < nsep; m++) {
            const int s = ((seps[(int64_t)w.pos * nsep + m] % N) + N) % N;
            double acc[kMaxOrder];
            for (int p = 0; p < np; p++) acc[p] = 0.0;
            for (int i = threadIdx.x; i < npts; i += blockDim.x) {
                const int li = i / N, x = i - li * N;
                int xs = x + s;
                if (xs >= N) xs -= N;
                const double d = (double)line[li * N + xs] - (double)line[i];
                double pw = 1.0;
                for (int p = 1; p < pmin; p++) pw *= d;
                for (int p = 0; p < np; p++) {
                    pw *= d;
                    acc[p] += pw;
                }
            }
            block_accumulate(acc, np, lds_acc + (size_t)m * np, scratch);
        }
        double *row = partials + it * nsep * np;
        for (int i = threadIdx.x; i < nsep * np; i += blockDim.x) row[i] = lds_acc[i];
        __syncthreads();
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
    const size_t sh = sizeof(double) * ((size_t)nsep * np + 4 * kMaxOrder) + sizeof(T) * kSfItemPoints;
    if (sh > (size_t)dev_lds) return hipErrorInvalidValue;
    const int64_t blocks = std::min<int64_t>(c->sf_nitems, (int64_t)cus * 4);
    double *partials = nullptr;
    WT_HIP(hipMallocAsync((void **)&partials, sizeof(double) * c->sf_nitems * nsep * np, st));
    WT_HIP(hipFuncSetAttribute((const void *)k_sf_x_batch<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    hipLaunchKernelGGL(k_sf_x_batch<T>, dim3((unsigned)blocks), dim3(kSfThreads), sh, st, (const T *)c->u, d_inst,
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

hipError_t launch_sf(wt_ctx *c, int mode, int pmin, int pmax, int nsep, const int32_t *d_seps, double *d_out) {
    if (pmax - pmin + 1 > kMaxOrder || pmin < 1) return hipErrorInvalidValue;
    return c->precision == 64 ? sf_t<double>(c, mode, pmin, pmax, nsep, d_seps, d_out)
                              : sf_t<float>(c, mode, pmin, pmax, nsep, d_seps, d_out);
}

}  // namespace wt
