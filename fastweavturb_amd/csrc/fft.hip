// Spectral Biot-Savart inversion (tube.f90:985-1043, subroutine.f90:888-919, 983-1084,
// 1550-1569, 1604-1638) with a hand-written pencil FFT.
//
// Transform layout: the half spectrum (kx = 0..N/2) of each component, spec[c][z][y][kx]
// with the row stride padded to whole 128-byte lines (spec_row), complex<T>. Passes:
//   x : R2C of two real lines at once packed as one complex line, scaled by 1/N^3
//   y : C2C along y on tiles of W kx-columns staged in LDS
//   z : C2C forward along z, then u_hat = i k~ x w_hat / k^2 (k~ = k with Nyquist components
//       zeroed, k^2 unchanged; this equals the real part the reference keeps, see below),
//       spectra of w_hat and u_hat into shells, Parseval sums, C2C inverse along z
//   y : inverse;  x : C2R, scaled by 1/up, and omega scaled by 1/up in the same pass.
// Why the Nyquist rule: the reference transforms real omega with full complex FFTs
// (wavenumber index N/2 maps to -N/2) and keeps dreal() of the inverse (subroutine.f90:1083).
// The real part equals the Hermitian projection (X(k) + conj X(-k))/2 of the spectrum; at a
// Nyquist index the partner -k has the same wavenumber -N/2, and the projection works out to
// zeroing that wavenumber component in the numerator while keeping it in k^2.
//
// The 1-D FFTs run in LDS: iterative in-place decimation in time over radices 16/8/4/3/2, data
// loaded in mixed-radix digit-reversed order, twiddles from a table exp(-2 pi i k / N).
#include <algorithm>
#include <cstdio>
#include <string>

#include "fft_lds.h"

namespace wt {

// Global loads per thread issued before the first of them is waited for, in the passes' staging
// loops (each loop step otherwise waits for its own load before storing it to LDS).
constexpr int kLoadBatch = 8;

// Quotient and remainder of i by d: shifts when d is a power of two (shift = log2 d, else -1),
// otherwise a multiply-high by m = ceil(2^32 / d), exact for i * d < 2^32.
struct Divider {
    int d, shift;
    uint32_t m;
};
__device__ __forceinline__ Divider make_divider(int d) {
    Divider v;
    v.d = d;
    v.shift = (d & (d - 1)) == 0 ? __ffs(d) - 1 : -1;
    v.m = (uint32_t)((0x100000000ull + (uint64_t)d - 1) / (uint64_t)d);
    return v;
}
__device__ __forceinline__ int div_of(int i, const Divider &v) {
    return v.shift >= 0 ? i >> v.shift : (int)__umulhi((uint32_t)i, v.m);
}

// Logical tile of a block such that each XCD works through one contiguous range of tiles. The
// hardware hands workgroup b to XCD b % 8, and each XCD has its own L2; the column tiles of the y
// and z passes that share 128-byte lines (tiles narrower than a line) are neighbours in tile order,
// so this keeps a line's readers on one L2 instead of fetching it once per XCD.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
    const unsigned per = nb / kXcds, rem = nb % kXcds, x = b % kXcds, k = b / kXcds;
    return x * per + (x < rem ? x : rem) + k;
}

// ---- x pass, forward: two real lines -> two half spectra --------------------------------
template <typename T>
__global__ void k_fft_x_r2c(const T *__restrict__ f, C2<T> *__restrict__ spec, FftPlan pl_g, int pairs_per_block,
                            int64_t npairs, T scale) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = (C2<T> *)smem_raw;
    const int N = pl_g.N, H = N / 2 + 1, P = lds_pitch(N), Hs = spec_row<T>(N);
    const FftPlan pl = stage_plan<T>(pl_g, smem_raw + sizeof(C2<T>) * (size_t)pairs_per_block * P);
    __syncthreads();
    const int64_t pair0 = (int64_t)blockIdx.x * pairs_per_block;
    int np = (int)((npairs - pair0) < pairs_per_block ? (npairs - pair0) : pairs_per_block);
    const Divider dN = make_divider(N), dH = make_divider(H);
    const T *f0 = f + 2 * pair0 * N;   // the block's lines are contiguous
    for (int i0 = threadIdx.x; i0 < np * N; i0 += kLoadBatch * blockDim.x) {
        T a[kLoadBatch], b[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < np * N) {
                const int pp = div_of(i, dN), x = i - pp * N;
                a[u] = f0[(2 * pp) * N + x];
                b[u] = f0[(2 * pp + 1) * N + x];
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < np * N) {
                const int pp = div_of(i, dN), x = i - pp * N;
                buf[pp * P + pidx(pl.rev[x])] = C2<T>{a[u], b[u]};
            }
        }
    }
    __syncthreads();
    fft_lds<T, false>(buf, np, P, pl);
    for (int i = threadIdx.x; i < np * H; i += blockDim.x) {
        int pp = div_of(i, dH), k = i - pp * H;
        C2<T> z = buf[pp * P + pidx(k == N / 2 ? N / 2 : k)];
        C2<T> zc = buf[pp * P + pidx((N - k) % N)];
        // A = (Z(k) + conj Z(N-k)) / 2, B = (Z(k) - conj Z(N-k)) / (2i)
        C2<T> A = {(z.x + zc.x) * (T)0.5, (z.y - zc.y) * (T)0.5};
        C2<T> B = {(z.y + zc.y) * (T)0.5, (zc.x - z.x) * (T)0.5};
        int64_t line = 2 * (pair0 + pp);
        spec[line * Hs + k] = C2<T>{A.x * scale, A.y * scale};
        spec[(line + 1) * Hs + k] = C2<T>{B.x * scale, B.y * scale};
    }
}

// ---- x pass, inverse: two half spectra -> two real lines, times inv_up; omega too (sc null:
// unscaled, omega untouched) ----------------------------------------------------------------
template <typename T>
__global__ void k_fft_x_c2r(const C2<T> *__restrict__ spec, T *__restrict__ u, T *__restrict__ omega, FftPlan pl_g,
                            int pairs_per_block, int64_t npairs, const DeviceScalars *__restrict__ sc,
                            int64_t pairs_per_inst) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = (C2<T> *)smem_raw;
    const int N = pl_g.N, P = lds_pitch(N), Hs = spec_row<T>(N);
    const FftPlan pl = stage_plan<T>(pl_g, smem_raw + sizeof(C2<T>) * (size_t)pairs_per_block * P);
    __syncthreads();
    const int64_t pair0 = (int64_t)blockIdx.x * pairs_per_block;
    int np = (int)((npairs - pair0) < pairs_per_block ? (npairs - pair0) : pairs_per_block);
    const Divider dN = make_divider(N);
    const C2<T> *s0 = spec + 2 * pair0 * Hs;   // the block's lines are contiguous
    for (int i0 = threadIdx.x; i0 < np * N; i0 += kLoadBatch * blockDim.x) {
        C2<T> A[kLoadBatch], B[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < np * N) {
                const int pp = div_of(i, dN), k = i - pp * N;
                const int kk = k <= N / 2 ? k : N - k;
                A[u] = s0[(2 * pp) * Hs + kk];
                B[u] = s0[(2 * pp + 1) * Hs + kk];
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < np * N) {
                const int pp = div_of(i, dN), k = i - pp * N;
                C2<T> a = A[u], b = B[u];
                if (k == 0 || k == N / 2) { a.y = 0; b.y = 0; }
                if (k > N / 2) { a.y = -a.y; b.y = -b.y; }   // Hermitian extension
                // Z = A + i B
                buf[pp * P + pidx(k)] = C2<T>{a.x - b.y, a.y + b.x};
            }
        }
    }
    __syncthreads();
    fft_lds_dif<T, true>(buf, np, P, pl);
    // a block's pairs may straddle two instances of the group
    T *u0 = u + 2 * pair0 * N;
    T *w0 = omega ? omega + 2 * pair0 * N : nullptr;
    for (int i0 = threadIdx.x; i0 < np * N; i0 += kLoadBatch * blockDim.x) {
        T oa[kLoadBatch], ob[kLoadBatch];
        if (w0) {
#pragma unroll
            for (int u = 0; u < kLoadBatch; u++) {
                const int i = i0 + u * blockDim.x;
                if (i < np * N) {
                    const int pp = div_of(i, dN), x = i - pp * N;
                    oa[u] = w0[(2 * pp) * N + x];
                    ob[u] = w0[(2 * pp + 1) * N + x];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < np * N) {
                const int pp = div_of(i, dN), x = i - pp * N;
                const T s = sc ? (T)sc[(pair0 + pp) / pairs_per_inst].inv_up : (T)1;
                C2<T> z = buf[pp * P + pidx(pl.rev[x])];
                u0[(2 * pp) * N + x] = z.x * s;
                u0[(2 * pp + 1) * N + x] = z.y * s;
                if (w0) {
                    w0[(2 * pp) * N + x] = oa[u] * s;
                    w0[(2 * pp + 1) * N + x] = ob[u] * s;
                }
            }
        }
    }
}

// ---- y pass: tiles of W kx columns x N y values ------------------------------------------
template <typename T, bool INV>
__global__ void k_fft_y(C2<T> *__restrict__ spec, FftPlan pl_g, int W, int nchunk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = (C2<T> *)smem_raw;
    const int N = pl_g.N, H = N / 2 + 1, P = lds_pitch(N), Hs = spec_row<T>(N);
    const FftPlan pl = stage_plan<T>(pl_g, smem_raw + sizeof(C2<T>) * (size_t)W * P);
    __syncthreads();
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    int chunk = tile % nchunk;
    int64_t plane = tile / nchunk;  // c*N + z
    int kx0 = chunk * W;
    int w_valid = H - kx0 < W ? H - kx0 : W;
    C2<T> *base = spec + plane * (int64_t)N * Hs;
    const Divider dW = make_divider(W);
    for (int i0 = threadIdx.x; i0 < W * N; i0 += kLoadBatch * blockDim.x) {
        C2<T> v[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            const int y = div_of(i, dW), w = i - y * W;
            if (i < W * N && w < w_valid) v[u] = base[(int64_t)y * Hs + kx0 + w];
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            const int y = div_of(i, dW), w = i - y * W;
            if (i < W * N && w < w_valid) buf[w * P + pidx(INV ? y : pl.rev[y])] = v[u];
        }
    }
    __syncthreads();
    if (INV) fft_lds_dif<T, true>(buf, w_valid, P, pl);
    else fft_lds<T, false>(buf, w_valid, P, pl);
    for (int i = threadIdx.x; i < W * N; i += blockDim.x) {
        const int y = div_of(i, dW), w = i - y * W;
        if (w < w_valid) base[(int64_t)y * Hs + kx0 + w] = buf[w * P + pidx(INV ? pl.rev[y] : y)];
    }
}

__device__ __forceinline__ int wavenum(int i, int N) { return ((i + N / 2) % N) - N / 2; }

// Shell of integer k2: ik = int(sqrt(k2 + 1e-8) + 0.5) = the m with (m-1/2)^2 <= k2 < (m+1/2)^2
__device__ __forceinline__ int shell_of(int64_t k2) {
    int m = (int)sqrt((double)k2);
    while ((int64_t)(2 * m + 1) * (2 * m + 1) <= 4 * k2) m++;
    while (m > 0 && (int64_t)(2 * m - 1) * (2 * m - 1) > 4 * k2) m--;
    return m;
}

// ---- z pass: forward, Biot-Savart (mode 0), spectrum only (mode 1) or the sharp spectral
// filter (mode 2: modes with k2 > kc2 set to zero, filter_spec, subroutine.f90:1225-1277),
// inverse. Partials row per block (modes 0 and 1): [0, nek) shells of w_hat (mode 0) |
// [nek, 2nek) shells of u_hat | [2nek] sum |w_hat|^2 weighted | [2nek+1] sum |u_hat|^2 weighted.
// One component's tile is in LDS at a time: each is loaded, transformed and copied to registers
// (thread t holds items i = t + k blockDim, i = w N + z, of every component), so the pointwise
// step has all three components at hand; the results go back through LDS one component at a
// time for the inverse transform. kZItems bounds the items per thread (W N <= kZItems blockDim).
template <typename T>
constexpr int kZItems = sizeof(T) == 4 ? 16 : 4;

template <typename T>
__device__ __forceinline__ void z_load_tile(C2<T> *buf, const C2<T> *__restrict__ col, int64_t zstride, int W,
                                            int w_valid, int N, int P, const int32_t *rev, const Divider &dW) {
    for (int i0 = threadIdx.x; i0 < W * N; i0 += kLoadBatch * blockDim.x) {
        C2<T> v[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            const int z = div_of(i, dW), w = i - z * W;
            v[u] = C2<T>{0, 0};
            if (i < W * N && w < w_valid) v[u] = col[z * zstride + w];
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int i = i0 + u * blockDim.x;
            const int z = div_of(i, dW), w = i - z * W;
            if (i < W * N) buf[w * P + pidx(rev[z])] = v[u];
        }
    }
}

template <typename T>
__device__ __forceinline__ void z_store_tile(const C2<T> *buf, C2<T> *__restrict__ col, int64_t zstride, int W,
                                             int w_valid, int N, int P, const int32_t *rev, const Divider &dW) {
    for (int i = threadIdx.x; i < W * N; i += blockDim.x) {
        const int z = div_of(i, dW), w = i - z * W;
        if (w < w_valid) col[z * zstride + w] = buf[w * P + pidx(rev[z])];
    }
}

template <typename T>
__global__ void k_fft_z(C2<T> *__restrict__ spec, FftPlan pl_g, int W, int nchunk, int mode, int nek,
                        double *__restrict__ partials, int64_t kc2) {
    constexpr int KM = kZItems<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = (C2<T> *)smem_raw;                       // W * P
    const int N = pl_g.N, H = N / 2 + 1, P = lds_pitch(N), Hs = spec_row<T>(N);
    unsigned char *plan_area = smem_raw + sizeof(C2<T>) * (size_t)W * P;
    double *shell = (double *)(plan_area + plan_lds_bytes(N, sizeof(C2<T>)));   // 2 * nek + 2
    const FftPlan pl = stage_plan<T>(pl_g, plan_area);
    const int64_t N3h = (int64_t)N * N * Hs;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    int chunk = tile % nchunk;
    int y = (tile / nchunk) % N;
    spec += (int64_t)3 * N3h * (tile / ((int64_t)nchunk * N));  // instance of the group
    int kx0 = chunk * W;
    int w_valid = H - kx0 < W ? H - kx0 : W;
    for (int i = threadIdx.x; i < 2 * nek + 2; i += blockDim.x) shell[i] = 0.0;
    const Divider dW = make_divider(W), dN = make_divider(N);
    const int64_t zstride = (int64_t)N * Hs;
    C2<T> *col0 = spec + (int64_t)y * Hs + kx0;           // component c: col0 + c * N3h
    const int ky = wavenum(y, N);
    __syncthreads();
    if (mode == 2) {   // each component on its own: forward, mask, inverse
        for (int c = 0; c < 3; c++) {
            z_load_tile<T>(buf, col0 + c * N3h, zstride, W, w_valid, N, P, pl.rev, dW);
            __syncthreads();
            fft_lds<T, false>(buf, W, P, pl);
            for (int i = threadIdx.x; i < w_valid * N; i += blockDim.x) {
                const int w = div_of(i, dN), z = i - w * N;
                const int kx = wavenum(kx0 + w, N), kz = wavenum(z, N);
                if ((int64_t)kx * kx + (int64_t)ky * ky + (int64_t)kz * kz > kc2) buf[w * P + pidx(z)] = C2<T>{0, 0};
            }
            __syncthreads();
            fft_lds_dif<T, true>(buf, W, P, pl);
            z_store_tile<T>(buf, col0 + c * N3h, zstride, W, w_valid, N, P, pl.rev, dW);
            __syncthreads();
        }
        return;
    }
    // forward transforms, one component at a time, into registers
    C2<T> R[3][KM];
    for (int c = 0; c < 3; c++) {
        z_load_tile<T>(buf, col0 + c * N3h, zstride, W, w_valid, N, P, pl.rev, dW);
        __syncthreads();
        fft_lds<T, false>(buf, W, P, pl);
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int i = threadIdx.x + k * blockDim.x;
            if (i < W * N) {
                const int w = div_of(i, dN), z = i - w * N;
                R[c][k] = buf[w * P + pidx(z)];
            }
        }
        __syncthreads();
    }
    // pointwise: (w, kz) items
    double e_w = 0.0, e_u = 0.0;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int i = threadIdx.x + k * blockDim.x;
        if (i >= w_valid * N) continue;
        const int w = div_of(i, dN), z = i - w * N;
        int kxi = kx0 + w;
        int kx = wavenum(kxi, N), kz = wavenum(z, N);
        int64_t k2 = (int64_t)kx * kx + (int64_t)ky * ky + (int64_t)kz * kz;
        double wt = (kxi == 0 || kxi == N / 2) ? 1.0 : 2.0;   // half-spectrum multiplicity
        const C2<T> a = R[0][k], b = R[1][k], cc = R[2][k];
        double ew = (double)a.x * a.x + (double)a.y * a.y + (double)b.x * b.x + (double)b.y * b.y +
                    (double)cc.x * cc.x + (double)cc.y * cc.y;
        int sh = shell_of(k2);
        if (mode == 0) {
            e_w += wt * ew;
            if (sh >= 1 && sh <= nek) atomicAdd(shell + sh - 1, 0.5 * wt * ew);
            // u_hat = i (k~ x w_hat) / k2, Nyquist components of the numerator zeroed
            T qx = (kxi == N / 2) ? (T)0 : (T)kx;
            T qy = (y == N / 2) ? (T)0 : (T)ky;
            T qz = (z == N / 2) ? (T)0 : (T)kz;
            C2<T> ux, uy, uz;
            if (k2 == 0) {
                ux = uy = uz = C2<T>{0, 0};
            } else {
                T ik2 = (T)(1.0 / (double)k2);
                // cross = q x w
                C2<T> cx = {qy * cc.x - qz * b.x, qy * cc.y - qz * b.y};
                C2<T> cy = {qz * a.x - qx * cc.x, qz * a.y - qx * cc.y};
                C2<T> cz = {qx * b.x - qy * a.x, qx * b.y - qy * a.y};
                ux = C2<T>{-cx.y * ik2, cx.x * ik2};
                uy = C2<T>{-cy.y * ik2, cy.x * ik2};
                uz = C2<T>{-cz.y * ik2, cz.x * ik2};
            }
            double eu = (double)ux.x * ux.x + (double)ux.y * ux.y + (double)uy.x * uy.x + (double)uy.y * uy.y +
                        (double)uz.x * uz.x + (double)uz.y * uz.y;
            e_u += wt * eu;
            if (sh >= 1 && sh <= nek) atomicAdd(shell + nek + sh - 1, 0.5 * wt * eu);
            R[0][k] = ux;
            R[1][k] = uy;
            R[2][k] = uz;
        } else {
            e_u += wt * ew;
            if (sh >= 1 && sh <= nek) atomicAdd(shell + nek + sh - 1, 0.5 * wt * ew);
        }
    }
    // block sums of the two energies in a fixed order (they set up, hence every normalised value):
    // a shuffle tree per wave, then the waves in order
    __shared__ double esum[2][16];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        e_w += __shfl_xor(e_w, o, 64);
        e_u += __shfl_xor(e_u, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        esum[0][threadIdx.x >> 6] = e_w;
        esum[1][threadIdx.x >> 6] = e_u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int wv = 0; wv < (int)(blockDim.x >> 6); wv++) {
            a += esum[0][wv];
            b += esum[1][wv];
        }
        shell[2 * nek] = a;
        shell[2 * nek + 1] = b;
    }
    __syncthreads();
    double *row = partials + (int64_t)tile * (2 * nek + 2);
    for (int i = threadIdx.x; i < 2 * nek + 2; i += blockDim.x) row[i] = shell[i];
    if (mode != 0) return;
    // inverse transforms, one component at a time
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int i = threadIdx.x + k * blockDim.x;
            if (i < W * N) {
                const int w = div_of(i, dN), z = i - w * N;
                buf[w * P + pidx(z)] = R[c][k];
            }
        }
        __syncthreads();
        fft_lds_dif<T, true>(buf, W, P, pl);
        z_store_tile<T>(buf, col0 + c * N3h, zstride, W, w_valid, N, P, pl.rev, dW);
        __syncthreads();
    }
}

// Reduce the per-block partial rows (columns are independent), then derive up and the
// normalised spectra.
// Column sums of the per-block partial rows, in a fixed order: thread (x, y) of a 64 x 16 block
// sums rows y, y + 16, ... of column 64 blockIdx.x + x, then the 16 partial sums are added in
// order of y.
__global__ __launch_bounds__(1024) void k_reduce_columns(const double *__restrict__ partials, int64_t rows, int cols,
                                                         double *__restrict__ out) {
    __shared__ double part[16][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + tx;
    partials += (int64_t)blockIdx.y * rows * cols;   // rows of instance blockIdx.y
    double s = 0.0;
    if (col < cols)
        for (int64_t r = ty; r < rows; r += 16) s += partials[r * cols + col];
    part[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && col < cols) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 16; k++) t += part[k][tx];
        out[(int64_t)blockIdx.y * cols + col] = t;
    }
}

__global__ void k_finalize(const double *__restrict__ sums, int nek, int mode, DeviceScalars *__restrict__ sc,
                           double *__restrict__ E_w, double *__restrict__ E_u) {
    // sums: [0,nek) w shells, [nek,2nek) u shells, [2nek] sum|w|^2, [2nek+1] sum|u|^2; one
    // block per instance of the group
    sums += (int64_t)blockIdx.x * (2 * nek + 2);
    sc += blockIdx.x;
    E_w += (int64_t)blockIdx.x * nek;
    E_u += (int64_t)blockIdx.x * nek;
    if (mode == 0) {
        double etot = 0.5 * sums[2 * nek + 1];
        double up = sqrt(2 * etot / 3);
        sc->etot_raw = etot;
        sc->up = up;
        sc->inv_up = 1.0 / up;
        sc->ens_raw = 0.5 * sums[2 * nek];
        sc->enstrophy = 0.5 * sums[2 * nek] / (up * up);
        sc->etot_norm = etot / (up * up);
        for (int i = threadIdx.x; i < nek; i += blockDim.x) {
            E_w[i] = sums[i];
            E_u[i] = sums[nek + i] / (up * up);
        }
    } else {
        for (int i = threadIdx.x; i < nek; i += blockDim.x) E_u[i] = sums[nek + i];
    }
}

// ---- host side ----------------------------------------------------------------------------
int64_t spec_elems_of(int N, int precision) {
    const int Hs = precision == 64 ? spec_row<double>(N) : spec_row<float>(N);
    return (int64_t)3 * N * N * Hs;
}

bool fft_size_supported(int N) {
    if (N < 4 || N > 1024 || N % 2) return false;
    int n = N;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

// Digit-reversal permutation matching fft_lds's pass order (radix[0] first) for make_plan(N, r16).
void fft_reversal(int N, std::vector<int32_t> &rev, bool r16) {
    FftPlan pl;
    make_plan(N, pl, r16);
    rev.resize(N);
    for (int i = 0; i < N; i++) {
        int x = i, pos = 0, mult = N;
        for (int p = pl.npass - 1; p >= 0; p--) {
            int R = pl.radix[p];
            mult /= R;
            pos += (x % R) * mult;
            x /= R;
        }
        rev[i] = pos;
    }
}

static const void *twiddles_for(wt_ctx *c, int N) {
    for (size_t i = 0; i < c->tw_N.size(); i++)
        if (c->tw_N[i] == N) return c->tw_tab[i];
    return nullptr;
}

// Static LDS of k_fft_z (the fixed-order energy sums) left free when sizing its dynamic LDS.
constexpr size_t kZStatic = 512;

// Tile shapes of the three passes: x lines (pairs) per block and threads, y columns per tile and
// threads, z columns per tile and threads.
struct FftTiles {
    int ppb, tx, wy, ty, wz, tz;
};

static int lds_limit(wt_ctx *c) {
    int dev_lds = 65536;
    (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    return dev_lds;
}
static size_t lds_x(int N, int ppb, size_t cs) { return (size_t)ppb * lds_pitch(N) * cs + plan_lds_bytes(N, cs); }
static size_t lds_y(int N, int wy, size_t cs) { return (size_t)wy * lds_pitch(N) * cs + plan_lds_bytes(N, cs); }
static size_t lds_z(int N, int wz, size_t cs, int nek) {
    return (size_t)wz * lds_pitch(N) * cs + plan_lds_bytes(N, cs) + sizeof(double) * (2 * nek + 2);
}
// whether a z tile of wz columns fits the registers of tz threads (kZItems items per thread)
static bool z_fits(int N, int wz, int tz, size_t cs) {
    return (int64_t)wz * N <= (int64_t)(cs == sizeof(C2<float>) ? kZItems<float> : kZItems<double>) * tz;
}

// The shapes used without tuning: as many lines/columns as fit, up to 8192 points (x) or 16
// columns (y, z), 256 threads.
static FftTiles default_tiles(wt_ctx *c, int N, int nek) {
    const size_t cs = c->precision == 64 ? sizeof(C2<double>) : sizeof(C2<float>);
    const size_t lim = (size_t)lds_limit(c);
    FftTiles t{};
    t.ppb = (int)(8192 / N) > 1 ? (int)(8192 / N) : 1;
    while (t.ppb > 1 && lds_x(N, t.ppb, cs) > lim) t.ppb /= 2;
    t.wy = 1;
    while (t.wy < 16 && lds_y(N, 2 * t.wy, cs) <= lim) t.wy *= 2;
    t.wz = 1;
    while (t.wz < 16 && lds_z(N, 2 * t.wz, cs, nek) + kZStatic <= lim) t.wz *= 2;
    t.tx = t.ty = t.tz = 256;
    while (!z_fits(N, t.wz, t.tz, cs) && t.tz < 1024) t.tz *= 2;
    while (!z_fits(N, t.wz, t.tz, cs) && t.wz > 1) t.wz /= 2;
    return t;
}

static FftTiles tiles_for(wt_ctx *c, int N, int nek) {
    for (size_t i = 0; i + 7 <= c->fft_tiles.size(); i += 7)
        if (c->fft_tiles[i] == N) {
            const int32_t *v = &c->fft_tiles[i + 1];
            FftTiles t{v[0], v[1], v[2], v[3], v[4], v[5]};
            const size_t cs = c->precision == 64 ? sizeof(C2<double>) : sizeof(C2<float>);
            while (!z_fits(N, t.wz, t.tz, cs) && t.tz < 1024) t.tz *= 2;   // shapes cached for other kernels
            while (!z_fits(N, t.wz, t.tz, cs) && t.wz > 1) t.wz /= 2;
            return t;
        }
    return default_tiles(c, N, nek);
}

// Bytes of per-block partial rows the z pass of K instances of grid N needs.
static size_t spectral_partials_bytes(wt_ctx *c, int N, int K, int nek, int Wz) {
    const int H = N / 2 + 1;
    const int ncz = (H + Wz - 1) / Wz;
    const int64_t zblocks = (int64_t)N * ncz * K;
    const int cols = 2 * nek + 2;
    return sizeof(double) * ((size_t)zblocks * cols + (size_t)K * cols);
}

// Grows the partial-row buffer to what every group of the context needs, so that a captured
// spectral stage never reallocates.
hipError_t spectral_reserve(wt_ctx *c) {
    size_t need = 0;
    const int n = c->ninst();
    for (int pos = 0; pos < n;) {
        int end = pos;
        while (end < n && c->inst[c->order[end]].N == c->inst[c->order[pos]].N) end++;
        const wt::Instance &I0 = c->inst[c->order[pos]];
        need = std::max(need, spectral_partials_bytes(c, I0.N, end - pos, I0.nek, tiles_for(c, I0.N, I0.nek).wz));
        pos = end;
    }
    if (need <= c->partials_cap) return hipSuccess;
    WT_HIP(hipStreamSynchronize(c->stream));
    if (c->partials) (void)hipFree(c->partials);
    c->partials = nullptr;
    c->partials_cap = 0;
    WT_HIP(hipMalloc(&c->partials, need));
    c->partials_cap = need;
    return hipSuccess;
}

// Biot-Savart (mode 0) or the velocity spectrum (mode 1) of K instances of equal N whose fields,
// spectra, scalars and spectra arrays are contiguous from arena position `pos`.
template <typename T>
static hipError_t spectral_group(wt_ctx *c, int mode, int pos, int K, bool timed, int64_t kc2 = 0) {
    hipStream_t st = c->stream;
    const wt::Instance &I0 = c->inst[c->order[pos]];
    const int N = I0.N, H = N / 2 + 1;
    FftPlan pl;
    make_plan(N, pl, true);
    pl.tw = twiddles_for(c, N);
    if (!pl.tw) return hipErrorInvalidValue;
    pl.rev = (const int32_t *)((const char *)pl.tw + sizeof(C2<T>) * N);
    const size_t cs = sizeof(C2<T>);
    T *field = (T *)(mode == 1 ? c->u : c->omega) + I0.field_off;
    C2<T> *spec = (C2<T> *)c->spec + I0.spec_off;
    const int nek = I0.nek;
    const FftTiles tl = tiles_for(c, N, nek);
    // x forward
    const int ppb = tl.ppb;  // pairs per block
    const int64_t pairs_per_inst = (int64_t)3 * N * N / 2;
    const int64_t npairs = pairs_per_inst * K;
    unsigned gx = (unsigned)((npairs + ppb - 1) / ppb);
    T scale = (T)(1.0 / ((double)N * N * N));
    size_t shx = lds_x(N, ppb, cs);
    WT_HIP(hipFuncSetAttribute((const void *)k_fft_x_r2c<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shx));
    if (timed) kt_begin(c, kSlotFftX);
    hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3(gx), dim3(tl.tx), shx, st, (const T *)field, spec, pl, ppb, npairs, scale);
    if (timed) kt_end(c, kSlotFftX);
    WT_HIP(hipGetLastError());
    // y forward
    const int Wy = tl.wy;
    int ncy = (H + Wy - 1) / Wy;
    size_t shy = lds_y(N, Wy, cs);
    WT_HIP(hipFuncSetAttribute((const void *)k_fft_y<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shy));
    WT_HIP(hipFuncSetAttribute((const void *)k_fft_y<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shy));
    if (timed) kt_begin(c, kSlotFftY);
    hipLaunchKernelGGL((k_fft_y<T, false>), dim3((unsigned)(3 * K * N * ncy)), dim3(tl.ty), shy, st, spec, pl, Wy, ncy);
    if (timed) kt_end(c, kSlotFftY);
    WT_HIP(hipGetLastError());
    // z fused
    const int Wz = tl.wz;
    int ncz = (H + Wz - 1) / Wz;
    size_t shz = lds_z(N, Wz, cs, nek);
    const int64_t zrows = (int64_t)N * ncz;    // blocks (partial rows) per instance
    const int64_t zblocks = zrows * K;
    const int cols = 2 * nek + 2;
    size_t need = sizeof(double) * ((size_t)zblocks * cols + (size_t)K * cols);
    if (mode != 2 && need > c->partials_cap) {            // the filter writes no partial rows
        if (c->capturing) return hipErrorInvalidValue;   // spectral_reserve runs before a capture
        WT_HIP(spectral_reserve(c));
        if (need > c->partials_cap) return hipErrorInvalidValue;
    }
    double *colsum = c->partials ? c->partials + (size_t)zblocks * cols : nullptr;
    if (!z_fits(N, Wz, tl.tz, cs)) return hipErrorInvalidConfiguration;
    WT_HIP(hipFuncSetAttribute((const void *)k_fft_z<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shz));
    if (timed) kt_begin(c, kSlotFftZ);
    hipLaunchKernelGGL(k_fft_z<T>, dim3((unsigned)zblocks), dim3(tl.tz), shz, st, spec, pl, Wz, ncz, mode, nek,
                       c->partials, kc2);
    if (timed) kt_end(c, kSlotFftZ);
    WT_HIP(hipGetLastError());
    if (mode == 2) {   // filtered omega: y inverse, x inverse unscaled into omega
        hipLaunchKernelGGL((k_fft_y<T, true>), dim3((unsigned)(3 * K * N * ncy)), dim3(tl.ty), shy, st, spec, pl, Wy, ncy);
        WT_HIP(hipGetLastError());
        WT_HIP(hipFuncSetAttribute((const void *)k_fft_x_c2r<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shx));
        hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3(gx), dim3(tl.tx), shx, st, (const C2<T> *)spec, field, (T *)nullptr, pl,
                           ppb, npairs, (const DeviceScalars *)nullptr, pairs_per_inst);
        return hipGetLastError();
    }
    if (timed) kt_begin(c, kSlotReduce);
    hipLaunchKernelGGL(k_reduce_columns, dim3((unsigned)((cols + 63) / 64), (unsigned)K), dim3(1024), 0, st,
                       (const double *)c->partials, zrows, cols, colsum);
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)K), dim3(256), 0, st, (const double *)colsum, nek, mode,
                       c->scal + pos, c->E_w + I0.E_off, c->E_u + I0.E_off);
    if (timed) kt_end(c, kSlotReduce);
    WT_HIP(hipGetLastError());
    if (mode != 0) return hipSuccess;
    // y inverse
    if (timed) kt_begin(c, kSlotFftYInv);
    hipLaunchKernelGGL((k_fft_y<T, true>), dim3((unsigned)(3 * K * N * ncy)), dim3(tl.ty), shy, st, spec, pl, Wy, ncy);
    if (timed) kt_end(c, kSlotFftYInv);
    WT_HIP(hipGetLastError());
    // x inverse (C2R), scaled by 1/up, omega scaled too
    WT_HIP(hipFuncSetAttribute((const void *)k_fft_x_c2r<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shx));
    if (timed) kt_begin(c, kSlotFftXInv);
    hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3(gx), dim3(tl.tx), shx, st, (const C2<T> *)spec, (T *)c->u + I0.field_off,
                       (T *)c->omega + I0.field_off, pl, ppb, npairs, (const DeviceScalars *)(c->scal + pos),
                       pairs_per_inst);
    if (timed) kt_end(c, kSlotFftXInv);
    return hipGetLastError();
}

template <typename T>
static hipError_t spectral_t(wt_ctx *c, int mode, int only) {
    const int n = c->ninst();
    if (only >= 0) {  // a single instance: find its arena position
        for (int pos = 0; pos < n; pos++)
            if (c->order[pos] == only) return spectral_group<T>(c, mode, pos, 1, false);
        return hipErrorInvalidValue;
    }
    // groups of equal N are contiguous in arena order; time the largest group
    int best = 0, bestw = -1;
    for (int pos = 0; pos < n;) {
        int end = pos;
        while (end < n && c->inst[c->order[end]].N == c->inst[c->order[pos]].N) end++;
        int64_t w = (int64_t)(end - pos) * c->inst[c->order[pos]].N * c->inst[c->order[pos]].N;
        if (w > bestw) { bestw = (int)w; best = pos; }
        pos = end;
    }
    for (int pos = 0; pos < n;) {
        int end = pos;
        while (end < n && c->inst[c->order[end]].N == c->inst[c->order[pos]].N) end++;
        WT_HIP(spectral_group<T>(c, mode, pos, end - pos, pos == best));
        pos = end;
    }
    return hipSuccess;
}

// ---- tile autotuning -----------------------------------------------------------------------
// Times every candidate shape of the x (forward), y (forward) and z (Biot-Savart) passes of one
// group on the context's own arenas and keeps the fastest of each. The x pass only reads omega;
// y and z work in place on the half spectrum, whose contents the next x pass rewrites; z also
// writes the partial rows. No result buffer (omega, u, scalars, spectra) is touched. The
// inverse passes use the forward passes' shapes.
template <typename T>
static hipError_t tune_group(wt_ctx *c, int pos, int K, FftTiles &best, std::string &rep) {
    hipStream_t st = c->stream;
    const wt::Instance &I0 = c->inst[c->order[pos]];
    const int N = I0.N, H = N / 2 + 1, nek = I0.nek;
    FftPlan pl;
    make_plan(N, pl, true);
    pl.tw = twiddles_for(c, N);
    if (!pl.tw) return hipErrorInvalidValue;
    pl.rev = (const int32_t *)((const char *)pl.tw + sizeof(C2<T>) * N);
    const size_t cs = sizeof(C2<T>), lim = (size_t)lds_limit(c);
    T *field = (T *)c->omega + I0.field_off;
    C2<T> *spec = (C2<T> *)c->spec + I0.spec_off;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    WT_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        return err;
    }
    auto timed = [&](auto launch, float *out) -> hipError_t {
        float b = 1e30f;
        for (int r = 0; r < 3; r++) {
            WT_HIP(hipEventRecord(e0, st));
            launch();
            WT_HIP(hipGetLastError());
            WT_HIP(hipEventRecord(e1, st));
            WT_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            WT_HIP(hipEventElapsedTime(&ms, e0, e1));
            b = ms < b ? ms : b;
        }
        *out = b;
        return hipSuccess;
    };
    const int threads[3] = {256, 512, 1024};
    best = default_tiles(c, N, nek);
    float bx = 1e30f, by = 1e30f, bz = 1e30f, t = 0.f;
    const int64_t npairs = (int64_t)3 * N * N / 2 * K;
    const T scale = (T)(1.0 / ((double)N * N * N));
    for (int ppb = 1; ppb <= 64 && err == hipSuccess; ppb *= 2) {
        const size_t sh = lds_x(N, ppb, cs);
        if (sh > lim || (int64_t)ppb * N > 32768) continue;
        err = hipFuncSetAttribute((const void *)k_fft_x_r2c<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        const unsigned gx = (unsigned)((npairs + ppb - 1) / ppb);
        for (int tx : threads) {
            if (err != hipSuccess) break;
            err = timed([&] { hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3(gx), dim3(tx), sh, st, (const T *)field, spec, pl,
                                                 ppb, npairs, scale); }, &t);
            if (err == hipSuccess && t < bx) { bx = t; best.ppb = ppb; best.tx = tx; }
        }
    }
    for (int wy = 1; wy <= 16 && err == hipSuccess; wy *= 2) {
        const size_t sh = lds_y(N, wy, cs);
        if (sh > lim) continue;
        err = hipFuncSetAttribute((const void *)k_fft_y<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        const int ncy = (H + wy - 1) / wy;
        for (int ty : threads) {
            if (err != hipSuccess) break;
            err = timed([&] { hipLaunchKernelGGL((k_fft_y<T, false>), dim3((unsigned)(3 * K * N * ncy)), dim3(ty), sh, st,
                                                 spec, pl, wy, ncy); }, &t);
            if (err == hipSuccess && t < by) { by = t; best.wy = wy; best.ty = ty; }
        }
    }
    // partial rows for the narrowest candidate
    if (err == hipSuccess) {
        const size_t need = spectral_partials_bytes(c, N, K, nek, 1);
        if (need > c->partials_cap) {
            err = hipStreamSynchronize(st);
            if (c->partials) (void)hipFree(c->partials);
            c->partials = nullptr;
            c->partials_cap = 0;
            if (err == hipSuccess) err = hipMalloc(&c->partials, need);
            if (err == hipSuccess) c->partials_cap = need;
        }
    }
    for (int wz = 1; wz <= 16 && err == hipSuccess; wz *= 2) {
        const size_t sh = lds_z(N, wz, cs, nek);
        if (sh + kZStatic > lim) continue;
        bool any = false;
        for (int tz : threads) any = any || z_fits(N, wz, tz, cs);
        if (!any) continue;
        err = hipFuncSetAttribute((const void *)k_fft_z<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        const int ncz = (H + wz - 1) / wz;
        const int64_t zblocks = (int64_t)N * ncz * K;
        for (int tz : threads) {
            if (err != hipSuccess) break;
            if (!z_fits(N, wz, tz, cs)) continue;
            err = timed([&] { hipLaunchKernelGGL(k_fft_z<T>, dim3((unsigned)zblocks), dim3(tz), sh, st, spec, pl, wz, ncz,
                                                 0, nek, c->partials, (int64_t)0); }, &t);
            if (err == hipSuccess && t < bz) { bz = t; best.wz = wz; best.tz = tz; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return err;
    char line[256];
    snprintf(line, sizeof line, "N=%d x(%d lines, %d threads) %.3f ms; y(%d cols, %d threads) %.3f ms; z(%d cols, %d threads) %.3f ms\n",
             N, best.ppb, best.tx, bx, best.wy, best.ty, by, best.wz, best.tz, bz);
    rep += line;
    return hipSuccess;
}

hipError_t spectral_autotune(wt_ctx *c, const std::vector<int32_t> &cached, std::string &report) {
    const int n = c->ninst();
    for (int pos = 0; pos < n;) {
        int end = pos;
        while (end < n && c->inst[c->order[end]].N == c->inst[c->order[pos]].N) end++;
        const int N = c->inst[c->order[pos]].N;
        bool have = false;
        for (size_t i = 0; i + 7 <= cached.size(); i += 7)
            if (cached[i] == N) {
                for (size_t j = 0; j + 7 <= c->fft_tiles.size(); j += 7)
                    if (c->fft_tiles[j] == N) c->fft_tiles.erase(c->fft_tiles.begin() + j, c->fft_tiles.begin() + j + 7);
                c->fft_tiles.insert(c->fft_tiles.end(), cached.begin() + i, cached.begin() + i + 7);
                have = true;
                report += "N=" + std::to_string(N) + " from cache\n";
            }
        if (!have) {
            FftTiles b{};
            WT_HIP(c->precision == 64 ? tune_group<double>(c, pos, end - pos, b, report)
                                      : tune_group<float>(c, pos, end - pos, b, report));
            for (size_t j = 0; j + 7 <= c->fft_tiles.size(); j += 7)
                if (c->fft_tiles[j] == N) c->fft_tiles.erase(c->fft_tiles.begin() + j, c->fft_tiles.begin() + j + 7);
            const int32_t v[7] = {N, b.ppb, b.tx, b.wy, b.ty, b.wz, b.tz};
            c->fft_tiles.insert(c->fft_tiles.end(), v, v + 7);
        }
        pos = end;
    }
    return spectral_reserve(c);
}

template <typename T>
static hipError_t filter_t(wt_ctx *c, int inst, int64_t kc2) {
    for (int pos = 0; pos < c->ninst(); pos++)
        if (c->order[pos] == inst) return spectral_group<T>(c, 2, pos, 1, false, kc2);
    return hipErrorInvalidValue;
}

hipError_t launch_filter(wt_ctx *c, int inst, int64_t kc2) {
    return c->precision == 64 ? filter_t<double>(c, inst, kc2) : filter_t<float>(c, inst, kc2);
}

hipError_t launch_spectral(wt_ctx *c, int mode, int only) {
    return c->precision == 64 ? spectral_t<double>(c, mode, only) : spectral_t<float>(c, mode, only);
}

}  // namespace wt
