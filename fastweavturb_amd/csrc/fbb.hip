// Fractional-Brownian-bridge centerlines on the GPU: the construction of FBB_with_parameters.m
// (lines 73-120) that produces the woven field's input centerlines, as fastweavturb_amd/fbb.py
// restates it with an exact fBm sampler:
//   - fractional Gaussian noise by circulant embedding (Davies & Harte 1987): with m = 2 n',
//     n' = the smallest 2^a 3^b >= n - 1, the eigenvalues lambda = FFT of the embedding's first
//     row gamma(0..n'), gamma(n'-1..1), gamma(k) = (|k+1|^2H - 2|k|^2H + |k-1|^2H) / 2; then two
//     complex draws w_d = FFT(sqrt(lambda / m) z_d) whose real and imaginary parts are independent
//     fGn samples (x from Re w_0, y from Re w_1, z from Im w_0);
//   - the bridge B_i - f_i B_{n-1} of the cumulative sums B, f_i = (t^2H + 1 - (1-t)^2H) / 2,
//     t = (i+1)/n (FBB:92-102);
//   - step-rms normalisation of the 3-D increments to dB (FBB:104-112);
//   - a uniform shift x0 in [-pi, pi)^3 and the wrap into [-pi, pi) (FBB:117-120).
// Two modes. Winding mode (the benchmark workloads): dB = wiggle * 2 pi |w| / n and a drift
// 2 pi w i/n along an integer lattice vector w, so that a closed centerline has a mean direction;
// the wrap is numpy's mod(x + pi, 2 pi) - pi. Reference mode: the construction as the reference
// writes it, with the level's dB and no drift; a tube of n < 128 points takes its points from a
// 128-point path at ceil(j 128 / n), j = 1..n (FBB:78-84); the wrap is mod(x, 2 pi) - pi with
// mod(x, y) = x - floor(x / y) y.
// Random numbers: Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed (seed, tube key), blocks
// at counters 1, 2, ... -- the word stream of numpy's Philox bit generator with that key; a
// uniform is (word >> 11) 2^-53, normals come from consecutive uniform pairs by Box-Muller. A
// tube's stream holds its 4m normals (Re z_0, Re z_1, Im z_0, Im z_1) and then 3 uniforms (x0).
// FFTs: the in-LDS transforms of fft_lds.h, in one step for m <= 1024, otherwise in four steps
// (m = m1 m2: length-m2 columns, twiddles exp(-2 pi i j1 k2 / m), length-m1 rows).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "fft_lds.h"

namespace wt {

using Z = C2<double>;

// ---- Philox4x64-10 ----------------------------------------------------------------------------
__device__ __forceinline__ void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
    const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
        const uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
        const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

// Uniform j of a tube's stream (words of the blocks at counters 1, 2, ...).
__device__ __forceinline__ void philox_uniform_pair(uint64_t seed, uint64_t tube, uint64_t j, double *u0,
                                                    double *u1) {
    uint64_t c[4] = {j / 4 + 1, 0, 0, 0};
    philox4x64_10(c, seed, tube);
    const int w = (int)(j % 4);
    const uint64_t a = w == 0 ? c[0] : (w == 1 ? c[1] : (w == 2 ? c[2] : c[3]));
    const uint64_t b = w == 0 ? c[1] : (w == 1 ? c[2] : (w == 2 ? c[3] : c[0]));
    *u0 = (double)(a >> 11) * (1.0 / 9007199254740992.0);
    *u1 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
}

// Normal q of a tube's stream: Box-Muller on uniforms 2p, 2p+1 (p = q/2; even q takes the cosine).
__device__ __forceinline__ double philox_normal(uint64_t seed, uint64_t tube, uint64_t q) {
    double u0, u1;
    philox_uniform_pair(seed, tube, q & ~1ull, &u0, &u1);
    const double r = sqrt(-2.0 * log(1.0 - u0));
    double s, c;
    sincospi(2.0 * u1, &s, &c);
    return (q & 1) ? r * s : r * c;
}

__device__ __forceinline__ double philox_uniform(uint64_t seed, uint64_t tube, uint64_t j) {
    double u0, u1;
    philox_uniform_pair(seed, tube, j, &u0, &u1);
    return u0;
}

// One tube of a group of equal embedding size m.
struct FbbTube {
    int64_t pt_off;      // first point in the output arrays
    int64_t gauss_off;   // first of its 4m + 3 inputs in the staged inputs (-1: Philox)
    uint64_t key;        // second Philox key word (the tube's index unless the caller gives keys)
    int32_t n;           // points
    int32_t ngen;        // points of the generated path: n, or 128 for a reference-mode tube of n < 128
    double H;
    double w[3];         // lattice winding (0 in reference mode)
    double dB;           // step rms of the scaled bridge
    int32_t ref;         // 1: reference mode
    int32_t pad;
};

// ---- kernels ------------------------------------------------------------------------------------
// Embedding rows: line b = gamma(min(k, m - k)), imaginary part 0.
__global__ void k_fbb_rows(Z *__restrict__ rows, const FbbTube *__restrict__ tubes, int ntube, int m) {
    const int64_t total = (int64_t)ntube * m;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(i / m), k = (int)(i - (int64_t)b * m);
        const double kk = (double)(k <= m / 2 ? k : m - k), e = 2.0 * tubes[b].H;
        const double g = 0.5 * ((pow(fabs(kk + 1.0), e) - 2.0 * pow(fabs(kk), e)) + pow(fabs(kk - 1.0), e));
        rows[i] = Z{g, 0.0};
    }
}

// Draw lines 2b, 2b+1: sqrt(max(Re lambda_k, 0) / m) (Re z_d,k + i Im z_d,k).
__global__ void k_fbb_draw(Z *__restrict__ draws, const Z *__restrict__ lam, const FbbTube *__restrict__ tubes,
                           int ntube, int m, const double *__restrict__ gauss, uint64_t seed) {
    const int64_t total = (int64_t)ntube * 2 * m;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t line = i / m;
        const int k = (int)(i - line * m);
        const int b = (int)(line >> 1), d = (int)(line & 1);
        const double s = sqrt(fmax(lam[(int64_t)b * m + k].x, 0.0) / (double)m);
        const FbbTube t = tubes[b];
        double re, im;
        if (t.gauss_off >= 0) {
            re = gauss[t.gauss_off + (int64_t)d * m + k];
            im = gauss[t.gauss_off + 2 * (int64_t)m + (int64_t)d * m + k];
        } else {
            re = philox_normal(seed, t.key, (uint64_t)d * m + k);
            im = philox_normal(seed, t.key, 2 * (uint64_t)m + (uint64_t)d * m + k);
        }
        draws[i] = Z{s * re, s * im};
    }
}

// Lines of length N <= 1024 (plan pl) transformed in LDS, lpb lines per block, in place.
__global__ void k_fbb_fft_lines(Z *__restrict__ data, int64_t nlines, FftPlan pl_g, int lpb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Z *buf = (Z *)smem_raw;
    const int N = pl_g.N, P = lds_pitch(N);
    const FftPlan pl = stage_plan<double>(pl_g, smem_raw + sizeof(Z) * (size_t)lpb * P);
    __syncthreads();
    const int64_t l0 = (int64_t)blockIdx.x * lpb;
    const int nl = (int)(nlines - l0 < lpb ? nlines - l0 : lpb);
    for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
        const int l = i / N, j = i - l * N;
        buf[l * P + pidx(pl.rev[j])] = data[(l0 + l) * N + j];
    }
    __syncthreads();
    fft_lds<double, false>(buf, nl, P, pl);
    for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
        const int l = i / N, k = i - l * N;
        data[(l0 + l) * N + k] = buf[l * P + pidx(k)];
    }
}

// Four-step, first half: for columns j1 = c0 .. c0 + W - 1 of every line, the length-m2 DFT over
// j2 of x[j1 + m1 j2], times exp(-2 pi i j1 k2 / m), written back to x[j1 + m1 k2].
__global__ void k_fbb_fft_cols(Z *__restrict__ data, int m1, int m2, FftPlan pl2_g, const Z *__restrict__ twm,
                               int W, int ncolblk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Z *buf = (Z *)smem_raw;
    const int P = lds_pitch(m2), m = m1 * m2;
    const FftPlan pl = stage_plan<double>(pl2_g, smem_raw + sizeof(Z) * (size_t)W * P);
    __syncthreads();
    const int64_t line = blockIdx.x / ncolblk;
    const int c0 = (int)(blockIdx.x - line * ncolblk) * W;
    const int nc = min(W, m1 - c0);
    Z *x = data + line * m;
    for (int i = threadIdx.x; i < nc * m2; i += blockDim.x) {
        const int j2 = i / nc, c = i - j2 * nc;
        buf[c * P + pidx(pl.rev[j2])] = x[(c0 + c) + (int64_t)m1 * j2];
    }
    __syncthreads();
    fft_lds<double, false>(buf, nc, P, pl);
    for (int i = threadIdx.x; i < nc * m2; i += blockDim.x) {
        const int k2 = i / nc, c = i - k2 * nc;
        const int j1 = c0 + c;
        const Z v = buf[c * P + pidx(k2)];
        const Z w = twm[(int)(((int64_t)j1 * k2) % m)];
        x[j1 + (int64_t)m1 * k2] = cmul(v, w);
    }
}

// Four-step, second half: rows k2 = r0 .. r0 + W - 1, the length-m1 DFT over j1 of
// x[j1 + m1 k2], written to out[k2 + m2 k1].
__global__ void k_fbb_fft_rows(const Z *__restrict__ data, Z *__restrict__ out, int m1, int m2, FftPlan pl1_g, int W,
                               int nrowblk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Z *buf = (Z *)smem_raw;
    const int P = lds_pitch(m1), m = m1 * m2;
    const FftPlan pl = stage_plan<double>(pl1_g, smem_raw + sizeof(Z) * (size_t)W * P);
    __syncthreads();
    const int64_t line = blockIdx.x / nrowblk;
    const int r0 = (int)(blockIdx.x - line * nrowblk) * W;
    const int nr = min(W, m2 - r0);
    const Z *x = data + line * m;
    Z *y = out + line * m;
    for (int i = threadIdx.x; i < nr * m1; i += blockDim.x) {
        const int r = i / m1, j1 = i - r * m1;
        buf[r * P + pidx(pl.rev[j1])] = x[j1 + (int64_t)m1 * (r0 + r)];
    }
    __syncthreads();
    fft_lds<double, false>(buf, nr, P, pl);
    for (int i = threadIdx.x; i < nr * m1; i += blockDim.x) {
        const int k1 = i / nr, r = i - k1 * nr;
        y[(r0 + r) + (int64_t)m2 * k1] = buf[r * P + pidx(k1)];
    }
}

constexpr int kPathThreads = 1024;

__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
    return s;
}

// numpy's float modulo (result with the sign of the divisor) of x + pi by 2 pi, minus pi.
__device__ __forceinline__ double wrap_pi(double x) {
    const double b = 6.28318530717958623199592693708837032318115234375;
    double r = fmod(x + 3.141592653589793115997963468544185161590576171875, b);
    if (r != 0.0) {
        if (r < 0.0) r += b;
    } else {
        r = 0.0;
    }
    return r - 3.141592653589793115997963468544185161590576171875;
}

// mod(x, 2 pi) - pi as the reference writes it (FBB:120), mod(x, y) = x - floor(x / y) y.
__device__ __forceinline__ double wrap_ref(double x) {
#pragma clang fp contract(off)
    const double b = 6.28318530717958623199592693708837032318115234375;
    const double q = floor(x / b);
    return (x - q * b) - 3.141592653589793115997963468544185161590576171875;
}

// A reference-mode tube of n < 128 points, by one thread: the 128-point path P_k (P_0 = 0, the
// cumulative fGn samples), its points P at ceil(j 128 / n) - 1 for j = 1..n, the bridge, the scale
// to dB and the shifted, wrapped positions.
__device__ void fbb_path_subsampled(const FbbTube &t, const Z *w0, const Z *w1, const double x0[3],
                                    double *const out[3]) {
    const int n = t.n, ng = t.ngen - 1;
    const double r = (double)t.ngen / (double)n;
    double s[3] = {0.0, 0.0, 0.0};
    int i = 0;
    for (int k = 0; k <= ng && i < n; k++) {
        while (i < n && (int)ceil((double)(i + 1) * r) - 1 == k) {
            for (int c = 0; c < 3; c++) out[c][i] = s[c];
            i++;
        }
        if (k < ng) {
            s[0] += w0[k].x;
            s[1] += w1[k].x;
            s[2] += w0[k].y;
        }
    }
    const double e = 2.0 * t.H, dn = (double)n;
    const double last[3] = {out[0][n - 1], out[1][n - 1], out[2][n - 1]};
    double ss = 0.0, prev[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < n; j++) {
        const double tt = (double)(j + 1) / dn;
        const double f = 0.5 * ((pow(tt, e) + 1.0) - pow(1.0 - tt, e));
        for (int c = 0; c < 3; c++) {
            const double v = out[c][j] - f * last[c];
            if (j > 0) ss += (v - prev[c]) * (v - prev[c]);
            out[c][j] = v;
            prev[c] = v;
        }
    }
    const double rms = ss / (double)(n - 1);
    const double scale = rms > 0.0 ? t.dB / sqrt(rms) : 0.0;
    for (int j = 0; j < n; j++)
        for (int c = 0; c < 3; c++) out[c][j] = wrap_ref(out[c][j] * scale + x0[c]);
}

// One block per tube: fGn samples of the two draws, cumulative sums, the bridge, the step-rms
// scale and the final wrapped positions, written to x, y, z at the tube's points. Thread t owns
// the points 1 + t*per .. (t+1)*per (thread 0 also point 0) and only ever reads back its own
// writes; values it needs from other chunks come through LDS.
__global__ __launch_bounds__(kPathThreads) void k_fbb_path(const Z *__restrict__ w, const FbbTube *__restrict__ tubes,
                                                          int m, const double *__restrict__ gauss,
                                                          uint64_t seed, double *__restrict__ X,
                                                          double *__restrict__ Y, double *__restrict__ Zc) {
    __shared__ double part[3][kPathThreads];
    __shared__ double total[3];
    __shared__ double red[kPathThreads / 64];
    const FbbTube t = tubes[blockIdx.x];
    const int n = t.n, ng = n - 1;
    const Z *w0 = w + (int64_t)(2 * blockIdx.x) * m, *w1 = w0 + m;
    double *const out[3] = {X + t.pt_off, Y + t.pt_off, Zc + t.pt_off};
    double x0[3];
    for (int c = 0; c < 3; c++) {
        const double u = t.gauss_off >= 0 ? gauss[t.gauss_off + 4 * (int64_t)m + c]
                                           : philox_uniform(seed, t.key, 4 * (uint64_t)m + c);
        x0[c] = (u - 0.5) * 2.0 * 3.141592653589793115997963468544185161590576171875;
    }
    if (t.ngen != t.n) {   // block-uniform: a short reference-mode tube
        if (threadIdx.x == 0) fbb_path_subsampled(t, w0, w1, x0, out);
        return;
    }
    const int per = (ng + kPathThreads - 1) / kPathThreads;
    const int g0 = min(ng, (int)threadIdx.x * per), g1 = min(ng, g0 + per);
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = g0; i < g1; i++) {
        s[0] += w0[i].x;
        s[1] += w1[i].x;
        s[2] += w0[i].y;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) part[c][threadIdx.x] = s[c];
    __syncthreads();
    if (threadIdx.x < 3) {   // exclusive prefix of the chunk sums, one coordinate per thread
        double acc = 0.0;
        for (int i = 0; i < kPathThreads; i++) {
            const double v = part[threadIdx.x][i];
            part[threadIdx.x][i] = acc;
            acc += v;
        }
        total[threadIdx.x] = acc;      // B_{n-1}
    }
    __syncthreads();
    const double e = 2.0 * t.H, dn = (double)n;
    const double last[3] = {total[0], total[1], total[2]};
    auto fbr = [&](int i) {            // bridge weight f of point i
        const double tt = (double)(i + 1) / dn;
        return 0.5 * ((pow(tt, e) + 1.0) - pow(1.0 - tt, e));
    };
    // B_0 = 0, B_{i+1} = B_i + g_i; bridge B_i - f_i B_{n-1}; squared steps of the bridge
    double b[3] = {part[0][threadIdx.x], part[1][threadIdx.x], part[2][threadIdx.x]};
    double ss = 0.0;
    {
        const double f = fbr(g0);
        double prev[3];
        for (int c = 0; c < 3; c++) prev[c] = b[c] - f * last[c];
        if (threadIdx.x == 0)
            for (int c = 0; c < 3; c++) out[c][0] = prev[c];
        for (int i = g0; i < g1; i++) {
            b[0] += w0[i].x;
            b[1] += w1[i].x;
            b[2] += w0[i].y;
            const double f1 = fbr(i + 1);
            for (int c = 0; c < 3; c++) {
                const double v = b[c] - f1 * last[c];
                const double d = v - prev[c];
                ss += d * d;
                out[c][i + 1] = v;
                prev[c] = v;
            }
        }
    }
    const double rms = block_sum(ss, red) / (double)(n - 1);
    const double scale = rms > 0.0 ? t.dB / sqrt(rms) : 0.0;
    const int p0 = threadIdx.x == 0 ? 0 : g0 + 1, p1 = g1;      // this thread's points
    for (int i = p0; i <= p1 && i < n; i++) {
        const double tt = (double)i / dn;
        for (int c = 0; c < 3; c++) {
            if (t.ref) {
                out[c][i] = wrap_ref(out[c][i] * scale + x0[c]);
            } else {
                const double drift = 2.0 * 3.141592653589793115997963468544185161590576171875 * t.w[c] * tt;
                out[c][i] = wrap_pi((x0[c] + drift) + out[c][i] * scale);
            }
        }
    }
}

// Uniform words of a tube's stream, for checking the generator against numpy's Philox.
__global__ void k_fbb_normals(double *out, int64_t count, uint64_t seed, uint64_t tube) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = philox_normal(seed, tube, (uint64_t)q);
}

// ---- host side ----------------------------------------------------------------------------------
int64_t fbb_smooth(int64_t n) {
    int64_t best = 1;
    while (best < n) best <<= 1;
    for (int64_t p3 = 1; p3 <= best; p3 *= 3) {
        int64_t v = p3;
        while (v < n) v <<= 1;
        best = std::min(best, v);
    }
    return best;
}

int64_t fbb_embedding(int64_t npoint) { return 2 * fbb_smooth(std::max<int64_t>(npoint - 1, 1)); }

namespace {

// FP64 plan tables (twiddles exp(-2 pi i k / N), digit reversal) on the device.
struct PlanTab {
    FftPlan pl{};
    void *mem = nullptr;
};

hipError_t plan_tab(int N, PlanTab &t, hipStream_t st) {
    make_plan(N, t.pl);
    std::vector<int32_t> rev;
    fft_reversal(N, rev);
    std::vector<unsigned char> host(sizeof(Z) * (size_t)N + sizeof(int32_t) * (size_t)N);
    Z *tw = (Z *)host.data();
    for (int k = 0; k < N; k++) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)N;
        tw[k] = Z{(double)cosl(a), (double)sinl(a)};
    }
    std::memcpy(host.data() + sizeof(Z) * (size_t)N, rev.data(), sizeof(int32_t) * (size_t)N);
    WT_HIP(hipMallocAsync(&t.mem, host.size(), st));
    WT_HIP(hipMemcpyAsync(t.mem, host.data(), host.size(), hipMemcpyHostToDevice, st));
    WT_HIP(hipStreamSynchronize(st));
    t.pl.tw = t.mem;
    t.pl.rev = (const int32_t *)((unsigned char *)t.mem + sizeof(Z) * (size_t)N);
    return hipSuccess;
}

bool smooth_23(int n) {
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

// m = m1 m2 with both 2^a 3^b and at most 1024, m1 as close to sqrt(m) as possible (m > 1024).
void split(int m, int &m1, int &m2) {
    m1 = 0;
    for (int d = 1; d <= 1024; d++)
        if (m % d == 0 && smooth_23(d) && m / d <= 1024 && smooth_23(m / d)) {
            if (m1 == 0 || std::abs((double)d * d - m) < std::abs((double)m1 * m1 - m)) m1 = d;
        }
    m2 = m1 ? m / m1 : 0;
}

size_t lines_lds(int N, int lines) { return sizeof(Z) * (size_t)lines * lds_pitch(N) + plan_lds_bytes(N, sizeof(Z)); }

// In-place forward DFT of nlines lines of length m at d (scratch: a second buffer of the same size).
hipError_t fft_batch(Z *d, Z *scratch, int64_t nlines, int m, hipStream_t st, std::map<int, PlanTab> &plans,
                     std::map<int, void *> &twm) {
    auto plan = [&](int N) -> const FftPlan * {
        auto it = plans.find(N);
        if (it == plans.end()) {
            PlanTab t;
            if (plan_tab(N, t, st) != hipSuccess) return nullptr;
            it = plans.emplace(N, t).first;
        }
        return &it->second.pl;
    };
    if (m <= 1024) {
        const FftPlan *pl = plan(m);
        if (!pl) return hipErrorOutOfMemory;
        const int lpb = std::max(1, std::min(8, 4096 / m));
        const size_t sh = lines_lds(m, lpb);
        WT_HIP(hipFuncSetAttribute((const void *)k_fbb_fft_lines, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        const int64_t blocks = (nlines + lpb - 1) / lpb;
        hipLaunchKernelGGL(k_fbb_fft_lines, dim3((unsigned)blocks), dim3(256), sh, st, d, nlines, *pl, lpb);
        return hipGetLastError();
    }
    int m1, m2;
    split(m, m1, m2);
    if (!m1) return hipErrorInvalidValue;
    const FftPlan *p1 = plan(m1), *p2 = plan(m2);
    if (!p1 || !p2) return hipErrorOutOfMemory;
    auto it = twm.find(m);
    if (it == twm.end()) {
        std::vector<Z> h(m);
        for (int k = 0; k < m; k++) {
            const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)m;
            h[k] = Z{(double)cosl(a), (double)sinl(a)};
        }
        void *p = nullptr;
        WT_HIP(hipMallocAsync(&p, sizeof(Z) * (size_t)m, st));
        WT_HIP(hipMemcpyAsync(p, h.data(), sizeof(Z) * (size_t)m, hipMemcpyHostToDevice, st));
        WT_HIP(hipStreamSynchronize(st));
        it = twm.emplace(m, p).first;
    }
    const int W = 4;
    const size_t sc = lines_lds(m2, W), sr = lines_lds(m1, W);
    WT_HIP(hipFuncSetAttribute((const void *)k_fbb_fft_cols, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sc));
    WT_HIP(hipFuncSetAttribute((const void *)k_fbb_fft_rows, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sr));
    const int ncol = (m1 + W - 1) / W, nrow = (m2 + W - 1) / W;
    hipLaunchKernelGGL(k_fbb_fft_cols, dim3((unsigned)(nlines * ncol)), dim3(256), sc, st, d, m1, m2, *p2,
                       (const Z *)it->second, W, ncol);
    WT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fbb_fft_rows, dim3((unsigned)(nlines * nrow)), dim3(256), sr, st, (const Z *)d, scratch, m1,
                       m2, *p1, W, nrow);
    WT_HIP(hipGetLastError());
    return hipMemcpyAsync(d, scratch, sizeof(Z) * (size_t)nlines * m, hipMemcpyDeviceToDevice, st);
}

}  // namespace

int32_t fbb_generated_points(int32_t npoint, bool ref) { return ref && npoint < 128 ? 128 : npoint; }

hipError_t fbb_centerlines(wt_ctx *c, int ntube, const int32_t *npoint, const double *hurst, const int32_t *winding,
                           double wiggle, const double *dB, uint64_t seed, const uint64_t *keys, const double *gauss,
                           double *cx, double *cy, double *cz) {
    hipStream_t st = c->stream;
    const bool ref = winding == nullptr;
    std::vector<int64_t> pt(ntube + 1, 0), goff(ntube + 1, 0);
    for (int l = 0; l < ntube; l++) {
        pt[l + 1] = pt[l] + npoint[l];
        goff[l + 1] = goff[l] + 4 * fbb_embedding(fbb_generated_points(npoint[l], ref)) + 3;
    }
    const int64_t P = pt[ntube];
    // tubes grouped by embedding size
    std::map<int, std::vector<int>> groups;
    for (int l = 0; l < ntube; l++) groups[(int)fbb_embedding(fbb_generated_points(npoint[l], ref))].push_back(l);
    double *dxyz = nullptr;
    WT_HIP(hipMallocAsync((void **)&dxyz, sizeof(double) * 3 * (size_t)P, st));
    std::map<int, PlanTab> plans;
    std::map<int, void *> twm;
    std::vector<void *> frees;
    hipError_t e = hipSuccess;
    const int64_t kChunkElems = (int64_t)1 << 25;   // complex elements of one chunk's draw lines
    for (auto &g : groups) {
        const int m = g.first;
        const std::vector<int> &ls = g.second;
        const int per = (int)std::max<int64_t>(1, kChunkElems / (2 * (int64_t)m));
        for (size_t a = 0; a < ls.size() && e == hipSuccess; a += per) {
            const int B = (int)std::min<size_t>(per, ls.size() - a);
            std::vector<FbbTube> tt(B);
            int64_t gsum = 0;
            for (int b = 0; b < B; b++) {
                const int l = ls[a + b];
                tt[b].pt_off = pt[l];
                tt[b].n = npoint[l];
                tt[b].ngen = fbb_generated_points(npoint[l], ref);
                tt[b].key = keys ? keys[l] : (uint64_t)l;
                tt[b].H = hurst[l];
                tt[b].ref = ref ? 1 : 0;
                for (int k = 0; k < 3; k++) tt[b].w[k] = ref ? 0.0 : (double)winding[3 * l + k];
                if (ref) {
                    tt[b].dB = dB[l];
                } else {
                    const double wn = std::sqrt((tt[b].w[0] * tt[b].w[0] + tt[b].w[1] * tt[b].w[1]) + tt[b].w[2] * tt[b].w[2]);
                    tt[b].dB = wiggle * (2.0 * 3.141592653589793115997963468544185161590576171875 * wn / (double)npoint[l]);
                }
                tt[b].gauss_off = gauss ? gsum : -1;
                gsum += 4 * (int64_t)m + 3;
            }
            FbbTube *dt = nullptr;
            Z *lines = nullptr, *scr = nullptr;
            double *dg = nullptr;
            e = hipMallocAsync((void **)&dt, sizeof(FbbTube) * B, st);
            if (e == hipSuccess) e = hipMallocAsync((void **)&lines, sizeof(Z) * (size_t)3 * B * m, st);
            if (e == hipSuccess) e = hipMallocAsync((void **)&scr, sizeof(Z) * (size_t)2 * B * m, st);
            if (e == hipSuccess && gauss) e = hipMallocAsync((void **)&dg, sizeof(double) * (size_t)gsum, st);
            if (e == hipSuccess) e = hipMemcpyAsync(dt, tt.data(), sizeof(FbbTube) * B, hipMemcpyHostToDevice, st);
            if (e == hipSuccess && gauss) {
                // the chunk's inputs, tube by tube
                std::vector<double> hg((size_t)gsum);
                for (int b = 0; b < B; b++) {
                    const int l = ls[a + b];
                    std::memcpy(hg.data() + tt[b].gauss_off, gauss + goff[l], sizeof(double) * (size_t)(4 * (int64_t)m + 3));
                }
                e = hipMemcpyAsync(dg, hg.data(), sizeof(double) * (size_t)gsum, hipMemcpyHostToDevice, st);
                if (e == hipSuccess) e = hipStreamSynchronize(st);
            }
            Z *lam = lines, *draws = lines + (size_t)B * m;
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_fbb_rows, dim3(1024), dim3(256), 0, st, lam, (const FbbTube *)dt, B, m);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = fft_batch(lam, scr, B, m, st, plans, twm);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_fbb_draw, dim3(2048), dim3(256), 0, st, draws, (const Z *)lam, (const FbbTube *)dt,
                                   B, m, (const double *)dg, seed);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = fft_batch(draws, scr, 2 * (int64_t)B, m, st, plans, twm);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_fbb_path, dim3(B), dim3(kPathThreads), 0, st, (const Z *)draws, (const FbbTube *)dt, m,
                                   (const double *)dg, seed, dxyz, dxyz + P, dxyz + 2 * P);
                e = hipGetLastError();
            }
            hipError_t f = hipStreamSynchronize(st);
            if (e == hipSuccess) e = f;
            for (void *p : {(void *)dt, (void *)lines, (void *)scr, (void *)dg})
                if (p) (void)hipFreeAsync(p, st);
        }
        if (e != hipSuccess) break;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(cx, dxyz, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(cy, dxyz + P, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(cz, dxyz + 2 * P, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(dxyz, st);
    for (auto &p : plans) (void)hipFreeAsync(p.second.mem, st);
    for (auto &p : twm) (void)hipFreeAsync(p.second, st);
    hipError_t f = hipStreamSynchronize(st);
    return e != hipSuccess ? e : f;
}

hipError_t fbb_normals(wt_ctx *c, uint64_t seed, uint64_t tube, int64_t count, double *host) {
    double *d = nullptr;
    WT_HIP(hipMallocAsync((void **)&d, sizeof(double) * (size_t)count, c->stream));
    hipLaunchKernelGGL(k_fbb_normals, dim3(256), dim3(256), 0, c->stream, d, count, seed, tube);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost, c->stream);
    (void)hipFreeAsync(d, c->stream);
    hipError_t f = hipStreamSynchronize(c->stream);
    return e != hipSuccess ? e : f;
}

}  // namespace wt
