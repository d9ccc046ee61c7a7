// In-LDS 1-D FFTs shared by the spectral stage (fft.hip) and the fBb centerline generator
// (fbb.hip): iterative in-place mixed-radix (16/8/4/3/2) transforms of lines staged in LDS, with
// twiddles exp(-2 pi i k / N) and a digit-reversal table from a plan (make_plan, fft_reversal).
#pragma once

#include "wt_internal.h"

namespace wt {

template <typename T> struct C2 { T x, y; };

template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ C2<T> mul_mi(C2<T> a) { return INV ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x}; }

struct FftPlan {
    int N;
    int npass;
    int radix[12];
    const int32_t *rev;   // digit-reversal permutation, N entries
    const void *tw;       // N complex twiddles exp(-2 pi i k / N) of type C2<T>
};

// The twiddle and digit-reversal tables of a plan, copied into LDS at `area` (16-byte aligned,
// plan_lds_bytes(N) long); the returned plan points at the copies. The caller synchronises.
__host__ __device__ constexpr size_t plan_lds_bytes(int N, size_t cs) {
    return ((size_t)N * cs + (size_t)N * sizeof(int32_t) + 15) / 16 * 16;
}
// Row stride of the half spectrum in HBM: N/2 + 1 columns rounded up to whole 128-byte lines, so
// the column tiles of the y and z passes start on line boundaries.
template <typename T>
__host__ __device__ constexpr int spec_row(int N) {
    return (N / 2 + 1 + (int)(128 / sizeof(C2<T>)) - 1) / (int)(128 / sizeof(C2<T>)) * (int)(128 / sizeof(C2<T>));
}

template <typename T>
__device__ __forceinline__ FftPlan stage_plan(const FftPlan &pl, unsigned char *area) {
    C2<T> *tw = (C2<T> *)area;
    int32_t *rv = (int32_t *)(tw + pl.N);
    const C2<T> *gtw = (const C2<T> *)pl.tw;
    for (int i = threadIdx.x; i < pl.N; i += blockDim.x) {
        tw[i] = gtw[i];
        rv[i] = pl.rev[i];
    }
    FftPlan q = pl;
    q.tw = tw;
    q.rev = rv;
    return q;
}


// LDS layout of the line buffers: element j of a line sits at pidx(j) = j + j/16 (one padding
// element per 16), and lines are lds_pitch(N) = N + N/16 + 1 elements apart (1 mod 32). Butterfly
// strides (powers of the radices) and the transposed tiles of the y and z passes (lanes one line
// apart) then spread over the LDS banks instead of piling onto a few.
__host__ __device__ constexpr int lds_pitch(int N) { return N + N / 16 + 1; }
__device__ __forceinline__ int pidx(int j) { return j + (j >> 4); }

template <typename T, bool INV>
__device__ __forceinline__ void dft4(C2<T> *v) {
    C2<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C2<T> s13 = cadd(v[1], v[3]), d13 = mul_mi<T, INV>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_small(C2<T> *v) {
    if (R == 2) {
        C2<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 3) {
        const T c1 = (T)-0.5, s1 = (T)(INV ? 0.86602540378443864676 : -0.86602540378443864676);
        C2<T> a = v[0], b = v[1], c = v[2];
        C2<T> s = cadd(b, c), d = csub(b, c);
        v[0] = cadd(a, s);
        C2<T> m = {a.x + c1 * s.x, a.y + c1 * s.y};
        C2<T> r = {-s1 * d.y, s1 * d.x};  // i*s1*d
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    } else if (R == 4) {
        dft4<T, INV>(v);
    } else if (R == 8) {
        C2<T> e[4] = {v[0], v[2], v[4], v[6]};
        C2<T> o[4] = {v[1], v[3], v[5], v[7]};
        dft4<T, INV>(e);
        dft4<T, INV>(o);
        const T h = (T)0.70710678118654752440;
        // w8^k for k = 0..3 (forward: exp(-i pi k / 4))
        C2<T> w1 = {h, INV ? h : -h};
        C2<T> o1 = cmul(o[1], w1);
        C2<T> o2 = mul_mi<T, INV>(o[2]);
        C2<T> w3 = {-h, INV ? h : -h};
        C2<T> o3 = cmul(o[3], w3);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    } else if (R == 16) {
        // 16 = 4 x 4: DFTs of the four decimated sequences v[n2], v[n2 + 4], ..., their outputs
        // k1 times w16^(n2 k1), then DFTs over n2; X[k1 + 4 k2] lands in v[k1 + 4 k2]
        const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
        const T sg = INV ? (T)1 : (T)-1;   // sign of the imaginary part of w16^m
        C2<T> a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
#pragma unroll
            for (int n1 = 0; n1 < 4; n1++) a[n2][n1] = v[4 * n1 + n2];
            dft4<T, INV>(a[n2]);
        }
        // w16^m for m = n2 k1 in {1, 2, 3, 4, 6, 9}
        a[1][1] = cmul(a[1][1], C2<T>{c1, sg * s1});
        a[1][2] = cmul(a[1][2], C2<T>{h, sg * h});
        a[1][3] = cmul(a[1][3], C2<T>{s1, sg * c1});
        a[2][1] = cmul(a[2][1], C2<T>{h, sg * h});
        a[2][2] = mul_mi<T, INV>(a[2][2]);
        a[2][3] = cmul(a[2][3], C2<T>{-h, sg * h});
        a[3][1] = cmul(a[3][1], C2<T>{s1, sg * c1});
        a[3][2] = cmul(a[3][2], C2<T>{-h, sg * h});
        a[3][3] = cmul(a[3][3], C2<T>{-c1, -sg * s1});
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            C2<T> b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            dft4<T, INV>(b);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = b[k2];
        }
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void fft_pass(C2<T> *buf, int nlines, int ldb, int N, int L,
                                         const C2<T> *__restrict__ tw) {
    const int nbf = N / R;
    const int total = nlines * nbf;
    const int step = N / (L * R);
    const bool pow2 = (N & (N - 1)) == 0;   // then nbf and L are powers of two: shifts, not divisions
    const int snbf = __ffs(nbf) - 1, sL = __ffs(L) - 1;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int line, b, g, k;
        if (pow2) {
            line = t >> snbf;
            b = t & (nbf - 1);
            g = b >> sL;
            k = b & (L - 1);
        } else {
            line = t / nbf;
            b = t - line * nbf;
            g = b / L;
            k = b - g * L;
        }
        C2<T> *p = buf + line * ldb;
        const int j0 = g * L * R + k;
        C2<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = p[pidx(j0 + r * L)];
        if (L > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                C2<T> w = tw[k * r * step];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<T, R, INV>(v);
#pragma unroll
        for (int r = 0; r < R; r++) p[pidx(j0 + r * L)] = v[r];
    }
}

// In-place FFT of nlines lines stored at buf + line*ldb, input already digit-reversed.
template <typename T, bool INV>
__device__ void fft_lds(C2<T> *buf, int nlines, int ldb, const FftPlan &pl) {
    const C2<T> *tw = (const C2<T> *)pl.tw;
    int L = 1;
    for (int p = 0; p < pl.npass; p++) {
        int R = pl.radix[p];
        switch (R) {
            case 16: fft_pass<T, 16, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            case 8: fft_pass<T, 8, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            case 4: fft_pass<T, 4, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            case 3: fft_pass<T, 3, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            default: fft_pass<T, 2, INV>(buf, nlines, ldb, pl.N, L, tw); break;
        }
        __syncthreads();
        L *= R;
    }
}

// Decimation in frequency: natural-order input, digit-reversed output (read back through
// rev). The same butterflies as fft_lds in reverse pass order, twiddles after the small DFT.
template <typename T, int R, bool INV>
__device__ __forceinline__ void fft_pass_dif(C2<T> *buf, int nlines, int ldb, int N, int L,
                                             const C2<T> *__restrict__ tw) {
    const int nbf = N / R;
    const int total = nlines * nbf;
    const int step = N / (L * R);
    const bool pow2 = (N & (N - 1)) == 0;   // then nbf and L are powers of two: shifts, not divisions
    const int snbf = __ffs(nbf) - 1, sL = __ffs(L) - 1;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int line, b, g, k;
        if (pow2) {
            line = t >> snbf;
            b = t & (nbf - 1);
            g = b >> sL;
            k = b & (L - 1);
        } else {
            line = t / nbf;
            b = t - line * nbf;
            g = b / L;
            k = b - g * L;
        }
        C2<T> *p = buf + line * ldb;
        const int j0 = g * L * R + k;
        C2<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = p[pidx(j0 + r * L)];
        dft_small<T, R, INV>(v);
        if (L > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                C2<T> w = tw[k * r * step];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) p[pidx(j0 + r * L)] = v[r];
    }
}

template <typename T, bool INV>
__device__ void fft_lds_dif(C2<T> *buf, int nlines, int ldb, const FftPlan &pl) {
    const C2<T> *tw = (const C2<T> *)pl.tw;
    int L = pl.N;
    for (int p = pl.npass - 1; p >= 0; p--) {
        int R = pl.radix[p];
        L /= R;
        switch (R) {
            case 16: fft_pass_dif<T, 16, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            case 8: fft_pass_dif<T, 8, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            case 4: fft_pass_dif<T, 4, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            case 3: fft_pass_dif<T, 3, INV>(buf, nlines, ldb, pl.N, L, tw); break;
            default: fft_pass_dif<T, 2, INV>(buf, nlines, ldb, pl.N, L, tw); break;
        }
        __syncthreads();
    }
}

// Radices of an N-point transform, largest first (N = 2^a 3^b). With r16, radix-16 passes are
// used where they make fewer passes than radix 8 (1024 = 16 16 4 instead of 8 8 8 2; 512 stays
// 8 8 8): every pass reads and writes the whole line in LDS once.
inline void make_plan(int N, FftPlan &pl, bool r16 = false) {
    auto fill = [&](bool use16) {
        pl.N = N;
        pl.npass = 0;
        int n = N;
        if (use16)
            while (n % 16 == 0) { pl.radix[pl.npass++] = 16; n /= 16; }
        while (n % 8 == 0) { pl.radix[pl.npass++] = 8; n /= 8; }
        while (n % 4 == 0) { pl.radix[pl.npass++] = 4; n /= 4; }
        while (n % 2 == 0) { pl.radix[pl.npass++] = 2; n /= 2; }
        while (n % 3 == 0) { pl.radix[pl.npass++] = 3; n /= 3; }
    };
    fill(false);
    if (r16) {
        const int n8 = pl.npass;
        fill(true);
        if (pl.npass >= n8) fill(false);
    }
}

}  // namespace wt
