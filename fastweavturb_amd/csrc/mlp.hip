// Inverse-design surrogate network: a multilayer perceptron (GELU hidden layers, linear output)
// trained and evaluated on the GPU with hand-written BF16 MFMA kernels and FP32 accumulation.
// The reference has no learned component; its closest analogue is the closed-form level
// scalings of the offline centerline generator (FBB_with_parameters.m:25-41, 185-193).
//
// Layout (all device-resident, row-major):
//   W_l  [O_l x I_l]  FP32 master weights, Adam moments m/v, and a BF16 copy used by the GEMMs
//   b_l  [O_l]        FP32
//   activations: X (BF16, B x I_0); per layer the pre-activation Z_l (FP32) and the BF16
//   post-activation A_l that feeds the next layer; output Y (FP32).
// Every dimension is padded to a multiple of 64 (batch rows, layer widths); padded weights,
// inputs and masks are zero, so padded outputs and gradients stay exactly zero.
//
// One GEMM kernel serves all products: C(m,n) = sum_k A(m,k) B(k,n) with arbitrary strides,
// 64x64 output tiles per workgroup (4 waves of 32x32), K staged through LDS 32 at a time,
// v_mfma_f32_16x16x32_bf16 on the tiles. Forward: Z = X W^T + b (+ GELU into BF16).
// Backward: dA_in = dZ W, dW = dZ^T A_in, db = column sums of dZ. Loss: masked mean squared
// error sum m (y - t)^2 / sum m. Optimiser: Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the
// FP32 masters, refreshing the BF16 copies. Every reduction has a fixed order, so training is
// bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/wovengpu.h"

namespace wtm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;       // output tile rows and columns; padding granularity
constexpr int kTileK = 32;      // K per MFMA step
constexpr int kLdsK = kTileK + 8;
constexpr int kMaxLayers = 8;

__host__ __device__ inline unsigned short f2bf(float f) {  // round to nearest even
    unsigned int u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { kEpiStore = 0, kEpiBias = 1, kEpiBiasGelu = 2 };

struct GemmArgs {
    const unsigned short *A;  // A(m,k) = A[m*sam + k*sak], BF16
    int64_t sam, sak;
    const unsigned short *B;  // B(k,n) = B[k*sbk + n*sbn], BF16
    int64_t sbk, sbn;
    int M, N, K;              // multiples of 64, 64, 32
    float *C;                 // C(m,n) = C[m*ldc + n], FP32 (pre-activation for GELU); may be null
    int64_t ldc;
    unsigned short *Cb;       // BF16 copy of the epilogue output (GELU applied); may be null
    int64_t ldcb;
    const float *bias;        // [N] for kEpiBias / kEpiBiasGelu
};

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) unsigned short As[kTile][kLdsK];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[kTile][kLdsK];  // Bs[n][k]
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bool a_kfast = g.sak == 1, b_kfast = g.sbk == 1;
    for (int k0 = 0; k0 < g.K; k0 += kTileK) {
        for (int e = tid; e < kTile * kTileK; e += 256) {
            int r, kk;
            if (a_kfast) { r = e / kTileK; kk = e % kTileK; } else { kk = e / kTile; r = e % kTile; }
            As[r][kk] = g.A[(int64_t)(m0 + r) * g.sam + (int64_t)(k0 + kk) * g.sak];
        }
        for (int e = tid; e < kTile * kTileK; e += 256) {
            int n, kk;
            if (b_kfast) { n = e / kTileK; kk = e % kTileK; } else { kk = e / kTile; n = e % kTile; }
            Bs[n][kk] = g.B[(int64_t)(k0 + kk) * g.sbk + (int64_t)(n0 + n) * g.sbn];
        }
        __syncthreads();
        // operand lanes: row (A) / column (B) = lane % 16, k = 8*(lane/16) .. +7 for both, so
        // the products pair the same k whatever order the unit sums them in
        const int r16 = lane & 15, kq = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; i++) {
            a[i] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&As[wm + i * 16 + r16][kq]);
            b[i] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&Bs[wn + i * 16 + r16][kq]);
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // result lanes: column = lane % 16, rows 4*(lane/16) .. +3
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + j * 16 + (lane & 15);
                float v = acc[i][j][r];
                if (EPI != kEpiStore) v += g.bias[n];
                if (g.C) g.C[(int64_t)m * g.ldc + n] = v;
                if (g.Cb) g.Cb[(int64_t)m * g.ldcb + n] = f2bf(EPI == kEpiBiasGelu ? gelu(v) : v);
            }
}

// Rows of a resident FP32 set (n x cols, unpadded) into a padded batch: BF16 inputs (rows x
// ldx), FP32 targets and masks (rows x ldt). Row i takes set row perm[base + i] (identity if
// perm is null); rows at or past `count` are zero.
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ T, const float *__restrict__ M,
                         const int32_t *__restrict__ perm, int64_t base, int count, int rows, int nin, int ldx,
                         int nout, int ldt, unsigned short *__restrict__ Xb, float *__restrict__ Tb,
                         float *__restrict__ Mb) {
    const int64_t total = (int64_t)rows * (ldx > ldt ? ldx : ldt);
    const int width = ldx > ldt ? ldx : ldt;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(e / width), c = (int)(e % width);
        bool live = i < count;
        int64_t src = live ? (perm ? perm[base + i] : base + i) : 0;
        if (c < ldx) Xb[(int64_t)i * ldx + c] = f2bf(live && c < nin ? X[src * nin + c] : 0.f);
        if (Tb && c < ldt) {
            Tb[(int64_t)i * ldt + c] = live && c < nout ? T[src * nout + c] : 0.f;
            Mb[(int64_t)i * ldt + c] = live && c < nout ? (M ? M[src * nout + c] : 1.f) : 0.f;
        }
    }
}

// Masked mean squared error of the output batch and its gradient dY = 2 m (y - t) / sum m
// (FP32 and BF16). One block; loss goes to *loss (FP64).
__global__ __launch_bounds__(1024) void k_loss_grad(const float *__restrict__ Y, const float *__restrict__ T,
                                                    const float *__restrict__ Mk, int64_t n, float *__restrict__ dZ,
                                                    unsigned short *__restrict__ dZb, double *__restrict__ loss) {
    __shared__ double red[2][16];
    double s_m = 0.0, s_l = 0.0;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        double m = Mk[e], d = (double)Y[e] - (double)T[e];
        s_m += m;
        s_l += m * d * d;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        s_m += __shfl_xor(s_m, o, 64);
        s_l += __shfl_xor(s_l, o, 64);
    }
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red[0][wv] = s_m; red[1][wv] = s_l; }
    __syncthreads();
    double msum = 0.0, lsum = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) { msum += red[0][i]; lsum += red[1][i]; }
    const float inv = msum > 0.0 ? (float)(2.0 / msum) : 0.f;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        float g = Mk[e] * (Y[e] - T[e]) * inv;
        dZ[e] = g;
        dZb[e] = f2bf(g);
    }
    if (threadIdx.x == 0 && loss) *loss = msum > 0.0 ? lsum / msum : 0.0;
}

// dZ = dA * gelu'(Z), FP32 and BF16.
__global__ void k_gelu_back(const float *__restrict__ dA, const float *__restrict__ Z, int64_t n,
                            float *__restrict__ dZ, unsigned short *__restrict__ dZb) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        float g = dA[e] * gelu_grad(Z[e]);
        dZ[e] = g;
        dZb[e] = f2bf(g);
    }
}

// db[c] = sum over rows of dZ[r][c], rows in order.
__global__ void k_colsum(const float *__restrict__ dZ, int rows, int cols, float *__restrict__ db) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += dZ[(int64_t)r * cols + c];
    db[c] = s;
}

// Adam on n parameters (FP32 masters), refreshing the BF16 copy where one is given.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int64_t n, float lr_t, float b1, float b2, float eps,
                       unsigned short *__restrict__ wb) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        float gi = g[e];
        float mi = b1 * m[e] + (1.f - b1) * gi;
        float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        float wi = w[e] - lr_t * mi / (sqrtf(vi) + eps);
        w[e] = wi;
        if (wb) wb[e] = f2bf(wi);
    }
}

__global__ void k_to_bf16(const float *__restrict__ x, int64_t n, unsigned short *__restrict__ y) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        y[e] = f2bf(x[e]);
}

inline int pad64(int n) { return (n + kTile - 1) / kTile * kTile; }

}  // namespace wtm

#define WT_HIP(expr)                                   \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return _e;               \
    } while (0)

struct wt_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int nl = 0;
    int in[wtm::kMaxLayers] = {}, out[wtm::kMaxLayers] = {};   // true sizes
    int ip[wtm::kMaxLayers] = {}, op[wtm::kMaxLayers] = {};    // padded sizes
    // parameters (padded), Adam state and gradients
    float *W[wtm::kMaxLayers] = {}, *b[wtm::kMaxLayers] = {};
    float *mW[wtm::kMaxLayers] = {}, *vW[wtm::kMaxLayers] = {}, *mb[wtm::kMaxLayers] = {}, *vb[wtm::kMaxLayers] = {};
    float *gW[wtm::kMaxLayers] = {}, *gb[wtm::kMaxLayers] = {};
    unsigned short *Wb[wtm::kMaxLayers] = {};
    int64_t adam_t = 0;
    // batch buffers for up to bcap rows
    int bcap = 0;
    unsigned short *Xb = nullptr;                    // bcap x ip[0]
    float *Z[wtm::kMaxLayers] = {};                  // bcap x op[l]
    unsigned short *Ab[wtm::kMaxLayers] = {};        // bcap x op[l] (hidden layers)
    float *Tb = nullptr, *Mb = nullptr;              // bcap x op[nl-1]
    float *dA = nullptr, *dZ = nullptr;              // bcap x maxw
    unsigned short *dZb = nullptr;
    double *loss = nullptr;                          // device scalar(s)
    int64_t loss_cap = 0;
    // resident training set
    float *setX = nullptr, *setT = nullptr, *setM = nullptr;
    int32_t *perm = nullptr;
    int64_t nset = 0;
    std::vector<void *> batch_allocs;
    int maxw() const {
        int m = ip[0];
        for (int l = 0; l < nl; l++) m = std::max(m, op[l]);
        return m;
    }
};

static thread_local char g_mlp_create_err[256] = "";

static int mlp_err(wt_mlp *m, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (m) m->err = buf;
    else snprintf(g_mlp_create_err, sizeof g_mlp_create_err, "%s", buf);
    return code;
}

static int mlp_hip(wt_mlp *m, hipError_t e, const char *where) {
    return mlp_err(m, e == hipErrorOutOfMemory ? WT_ENOMEM : WT_EDEVICE, "%s: HIP error %d (%s)", where, (int)e,
                   hipGetErrorString(e));
}

#define MLP_HIP(m, expr, where)                        \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return mlp_hip(m, _e, where); \
    } while (0)

#define MLP_CTX(m)                                     \
    do {                                               \
        if (!(m)) return WT_EINVAL;                    \
        (m)->err.clear();                              \
        hipError_t _s = hipSetDevice((m)->device);     \
        if (_s != hipSuccess) return mlp_hip(m, _s, "hipSetDevice"); \
    } while (0)

template <typename T>
static hipError_t dalloc(T **p, size_t n) {
    return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

static unsigned grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 4096));
}

static hipError_t ensure_batch(wt_mlp *m, int rows) {
    if (rows <= m->bcap) return hipSuccess;
    for (void *p : m->batch_allocs) (void)hipFree(p);
    m->batch_allocs.clear();
    m->bcap = 0;
    auto A = [&](auto **p, size_t n) -> hipError_t {
        hipError_t e = dalloc(p, n);
        if (e == hipSuccess) m->batch_allocs.push_back((void *)*p);
        return e;
    };
    const size_t R = (size_t)rows;
    WT_HIP(A(&m->Xb, R * m->ip[0]));
    for (int l = 0; l < m->nl; l++) {
        WT_HIP(A(&m->Z[l], R * m->op[l]));
        WT_HIP(A(&m->Ab[l], R * m->op[l]));
    }
    WT_HIP(A(&m->Tb, R * m->op[m->nl - 1]));
    WT_HIP(A(&m->Mb, R * m->op[m->nl - 1]));
    WT_HIP(A(&m->dA, R * m->maxw()));
    WT_HIP(A(&m->dZ, R * m->maxw()));
    WT_HIP(A(&m->dZb, R * m->maxw()));
    m->bcap = rows;
    return hipSuccess;
}

static hipError_t gemm(hipStream_t st, int epi, const wtm::GemmArgs &g) {
    if (g.M % wtm::kTile || g.N % wtm::kTile || g.K % wtm::kTileK) return hipErrorInvalidValue;
    dim3 grid((unsigned)(g.N / wtm::kTile), (unsigned)(g.M / wtm::kTile));
    if (epi == wtm::kEpiStore) hipLaunchKernelGGL(wtm::k_gemm<wtm::kEpiStore>, grid, dim3(256), 0, st, g);
    else if (epi == wtm::kEpiBias) hipLaunchKernelGGL(wtm::k_gemm<wtm::kEpiBias>, grid, dim3(256), 0, st, g);
    else hipLaunchKernelGGL(wtm::k_gemm<wtm::kEpiBiasGelu>, grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

// Forward of `rows` batch rows already in m->Xb; output layer pre-activations land in Z[nl-1].
static hipError_t forward(wt_mlp *m, int rows) {
    const unsigned short *x = m->Xb;
    for (int l = 0; l < m->nl; l++) {
        const bool last = l == m->nl - 1;
        wtm::GemmArgs g{};
        g.A = x; g.sam = m->ip[l]; g.sak = 1;
        g.B = m->Wb[l]; g.sbk = 1; g.sbn = m->ip[l];      // B(k=i, n=o) = W[o][i]
        g.M = rows; g.N = m->op[l]; g.K = m->ip[l];
        g.C = m->Z[l]; g.ldc = m->op[l];
        g.Cb = last ? nullptr : m->Ab[l]; g.ldcb = m->op[l];
        g.bias = m->b[l];
        WT_HIP(gemm(m->stream, last ? wtm::kEpiBias : wtm::kEpiBiasGelu, g));
        x = m->Ab[l];
    }
    return hipSuccess;
}

// Loss and gradients of the batch in Xb/Tb/Mb (rows padded); gradients into gW/gb.
static hipError_t backward(wt_mlp *m, int rows, double *d_loss) {
    const int L = m->nl - 1;
    hipStream_t st = m->stream;
    hipLaunchKernelGGL(wtm::k_loss_grad, dim3(1), dim3(1024), 0, st, (const float *)m->Z[L], (const float *)m->Tb,
                       (const float *)m->Mb, (int64_t)rows * m->op[L], m->dZ, m->dZb, d_loss);
    WT_HIP(hipGetLastError());
    for (int l = L; l >= 0; l--) {
        const unsigned short *ain = l == 0 ? m->Xb : m->Ab[l - 1];
        // dW = dZ^T A_in: A(m=o, k=r) = dZ[r][o], B(k=r, n=i) = A_in[r][i]
        wtm::GemmArgs gw{};
        gw.A = m->dZb; gw.sam = 1; gw.sak = m->op[l];
        gw.B = ain; gw.sbk = m->ip[l]; gw.sbn = 1;
        gw.M = m->op[l]; gw.N = m->ip[l]; gw.K = rows;
        gw.C = m->gW[l]; gw.ldc = m->ip[l];
        WT_HIP(gemm(st, wtm::kEpiStore, gw));
        hipLaunchKernelGGL(wtm::k_colsum, dim3((unsigned)((m->op[l] + 63) / 64)), dim3(64), 0, st,
                           (const float *)m->dZ, rows, m->op[l], m->gb[l]);
        WT_HIP(hipGetLastError());
        if (l == 0) break;
        // dA_in = dZ W: A(m=r, k=o) = dZ[r][o], B(k=o, n=i) = W[o][i]
        wtm::GemmArgs gd{};
        gd.A = m->dZb; gd.sam = m->op[l]; gd.sak = 1;
        gd.B = m->Wb[l]; gd.sbk = m->ip[l]; gd.sbn = 1;
        gd.M = rows; gd.N = m->ip[l]; gd.K = m->op[l];
        gd.C = m->dA; gd.ldc = m->ip[l];
        WT_HIP(gemm(st, wtm::kEpiStore, gd));
        const int64_t n = (int64_t)rows * m->ip[l];
        hipLaunchKernelGGL(wtm::k_gelu_back, dim3(grid_for(n)), dim3(256), 0, st, (const float *)m->dA,
                           (const float *)m->Z[l - 1], n, m->dZ, m->dZb);
        WT_HIP(hipGetLastError());
    }
    return hipSuccess;
}

static hipError_t adam(wt_mlp *m, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m->adam_t++;
    const double bc1 = 1.0 - std::pow((double)b1, (double)m->adam_t);
    const double bc2 = 1.0 - std::pow((double)b2, (double)m->adam_t);
    const float lr_t = (float)(lr * std::sqrt(bc2) / bc1);
    for (int l = 0; l < m->nl; l++) {
        const int64_t nw = (int64_t)m->op[l] * m->ip[l];
        hipLaunchKernelGGL(wtm::k_adam, dim3(grid_for(nw)), dim3(256), 0, m->stream, m->W[l], m->mW[l], m->vW[l],
                           (const float *)m->gW[l], nw, lr_t, b1, b2, eps, m->Wb[l]);
        hipLaunchKernelGGL(wtm::k_adam, dim3(1), dim3(256), 0, m->stream, m->b[l], m->mb[l], m->vb[l],
                           (const float *)m->gb[l], (int64_t)m->op[l], lr_t, b1, b2, eps, (unsigned short *)nullptr);
    }
    return hipGetLastError();
}

// Uploads an unpadded host batch and gathers it into the padded batch buffers.
static hipError_t load_host_batch(wt_mlp *m, int n, const float *X, const float *T, const float *Mk, int rows) {
    hipStream_t st = m->stream;
    const int nin = m->in[0], nout = m->out[m->nl - 1];
    float *dX = nullptr, *dT = nullptr, *dM = nullptr;
    WT_HIP(hipMallocAsync((void **)&dX, sizeof(float) * (size_t)n * nin + 4, st));
    hipError_t e = hipMemcpyAsync(dX, X, sizeof(float) * (size_t)n * nin, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && T) {
        e = hipMallocAsync((void **)&dT, sizeof(float) * (size_t)n * nout + 4, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dT, T, sizeof(float) * (size_t)n * nout, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && Mk) {
            e = hipMallocAsync((void **)&dM, sizeof(float) * (size_t)n * nout + 4, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dM, Mk, sizeof(float) * (size_t)n * nout, hipMemcpyHostToDevice, st);
        }
    }
    if (e == hipSuccess) {
        const int width = std::max(m->ip[0], m->op[m->nl - 1]);
        hipLaunchKernelGGL(wtm::k_gather, dim3(grid_for((int64_t)rows * width)), dim3(256), 0, st, (const float *)dX,
                           (const float *)dT, (const float *)dM, (const int32_t *)nullptr, (int64_t)0, n, rows, nin,
                           m->ip[0], nout, m->op[m->nl - 1], m->Xb, T ? m->Tb : (float *)nullptr, m->Mb);
        e = hipGetLastError();
    }
    if (dX) (void)hipFreeAsync(dX, st);
    if (dT) (void)hipFreeAsync(dT, st);
    if (dM) (void)hipFreeAsync(dM, st);
    return e;
}

extern "C" {

const char *wt_mlp_last_error(const wt_mlp *m) { return m ? m->err.c_str() : g_mlp_create_err; }

void wt_mlp_destroy(wt_mlp *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *p : m->batch_allocs) (void)hipFree(p);
    for (int l = 0; l < m->nl; l++) {
        void *ps[] = {m->W[l], m->b[l], m->mW[l], m->vW[l], m->mb[l], m->vb[l], m->gW[l], m->gb[l], m->Wb[l]};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
    void *ps[] = {m->loss, m->setX, m->setT, m->setM, m->perm};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int wt_mlp_create(wt_mlp **out, int32_t nlayer, const int32_t *sizes, uint64_t seed, int32_t device) {
    if (!out) return mlp_err(nullptr, WT_EINVAL, "wt_mlp_create: NULL handle pointer");
    *out = nullptr;
    if (nlayer < 1 || nlayer > wtm::kMaxLayers || !sizes)
        return mlp_err(nullptr, WT_EINVAL, "need 1..%d layers and a sizes array", wtm::kMaxLayers);
    for (int l = 0; l <= nlayer; l++)
        if (sizes[l] < 1 || sizes[l] > 8192) return mlp_err(nullptr, WT_EINVAL, "layer width %d out of range", sizes[l]);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mlp_err(nullptr, WT_EDEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return mlp_err(nullptr, WT_EINVAL, "device %d out of range", device);
    wt_mlp *m = new (std::nothrow) wt_mlp();
    if (!m) return mlp_err(nullptr, WT_ENOMEM, "host allocation failed");
    m->device = device;
    m->nl = nlayer;
    for (int l = 0; l < nlayer; l++) {
        m->in[l] = sizes[l];
        m->out[l] = sizes[l + 1];
        m->ip[l] = wtm::pad64(sizes[l]);
        m->op[l] = wtm::pad64(sizes[l + 1]);
    }
    auto fail = [&](hipError_t e, const char *w) {
        int code = e == hipErrorOutOfMemory ? WT_ENOMEM : WT_EDEVICE;
        mlp_err(nullptr, code, "%s: %s", w, hipGetErrorString(e));
        wt_mlp_destroy(m);
        return code;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return fail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    // He-uniform initial weights for the GELU layers, Glorot-uniform for the output layer,
    // zero biases, from a seeded generator on the host
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int l = 0; l < nlayer; l++) {
        const size_t nw = (size_t)m->op[l] * m->ip[l], nb = (size_t)m->op[l];
        float **bufs[] = {&m->W[l], &m->mW[l], &m->vW[l], &m->gW[l]};
        for (float **p : bufs)
            if ((e = dalloc(p, nw)) != hipSuccess || (e = hipMemset(*p, 0, nw * 4)) != hipSuccess) return fail(e, "weights");
        float **bb[] = {&m->b[l], &m->mb[l], &m->vb[l], &m->gb[l]};
        for (float **p : bb)
            if ((e = dalloc(p, nb)) != hipSuccess || (e = hipMemset(*p, 0, nb * 4)) != hipSuccess) return fail(e, "biases");
        if ((e = dalloc(&m->Wb[l], nw)) != hipSuccess) return fail(e, "weights");
        std::vector<float> w(nw, 0.f);
        const bool last = l == nlayer - 1;
        const double a = last ? std::sqrt(6.0 / (m->in[l] + m->out[l])) : std::sqrt(6.0 / m->in[l]);
        for (int o = 0; o < m->out[l]; o++)
            for (int i = 0; i < m->in[l]; i++) w[(size_t)o * m->ip[l] + i] = (float)(a * U(rng));
        std::vector<unsigned short> wb(nw);
        for (size_t i = 0; i < nw; i++) wb[i] = wtm::f2bf(w[i]);
        if ((e = hipMemcpy(m->W[l], w.data(), nw * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "weights");
        if ((e = hipMemcpy(m->Wb[l], wb.data(), nw * 2, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "weights");
    }
    if ((e = dalloc(&m->loss, 1)) != hipSuccess) return fail(e, "loss");
    m->loss_cap = 1;
    *out = m;
    return WT_OK;
}

int64_t wt_mlp_num_params(const wt_mlp *m) {
    if (!m) return WT_EINVAL;
    int64_t n = 0;
    for (int l = 0; l < m->nl; l++) n += (int64_t)m->out[l] * m->in[l] + m->out[l];
    return n;
}

// Copies between the flat unpadded order (per layer: W row-major [out][in], then b) and the
// padded device arrays. dir 0: device -> host, 1: host -> device.
static int copy_params(wt_mlp *m, float *flat, int dir, float *const *Wsrc, float *const *bsrc) {
    size_t off = 0;
    for (int l = 0; l < m->nl; l++) {
        const size_t nw = (size_t)m->op[l] * m->ip[l];
        std::vector<float> w(nw, 0.f), b(m->op[l], 0.f);
        if (dir == 0) {
            MLP_HIP(m, hipMemcpy(w.data(), Wsrc[l], nw * 4, hipMemcpyDeviceToHost), "copy");
            MLP_HIP(m, hipMemcpy(b.data(), bsrc[l], (size_t)m->op[l] * 4, hipMemcpyDeviceToHost), "copy");
        }
        for (int o = 0; o < m->out[l]; o++)
            for (int i = 0; i < m->in[l]; i++) {
                float &d = w[(size_t)o * m->ip[l] + i];
                if (dir == 0) flat[off++] = d;
                else d = flat[off++];
            }
        for (int o = 0; o < m->out[l]; o++) {
            if (dir == 0) flat[off++] = b[o];
            else b[o] = flat[off++];
        }
        if (dir == 1) {
            std::vector<unsigned short> wb(nw);
            for (size_t i = 0; i < nw; i++) wb[i] = wtm::f2bf(w[i]);
            MLP_HIP(m, hipMemcpy(Wsrc[l], w.data(), nw * 4, hipMemcpyHostToDevice), "copy");
            MLP_HIP(m, hipMemcpy(m->Wb[l], wb.data(), nw * 2, hipMemcpyHostToDevice), "copy");
            MLP_HIP(m, hipMemcpy(bsrc[l], b.data(), (size_t)m->op[l] * 4, hipMemcpyHostToDevice), "copy");
        }
    }
    return WT_OK;
}

int wt_mlp_get_params(wt_mlp *m, float *flat) {
    MLP_CTX(m);
    if (!flat) return mlp_err(m, WT_EINVAL, "NULL buffer");
    MLP_HIP(m, hipStreamSynchronize(m->stream), "sync");
    return copy_params(m, flat, 0, m->W, m->b);
}

int wt_mlp_set_params(wt_mlp *m, const float *flat) {
    MLP_CTX(m);
    if (!flat) return mlp_err(m, WT_EINVAL, "NULL buffer");
    MLP_HIP(m, hipStreamSynchronize(m->stream), "sync");
    int rc = copy_params(m, const_cast<float *>(flat), 1, m->W, m->b);
    if (rc) return rc;
    // a new starting point: reset the optimiser
    for (int l = 0; l < m->nl; l++) {
        const size_t nw = (size_t)m->op[l] * m->ip[l];
        MLP_HIP(m, hipMemset(m->mW[l], 0, nw * 4), "reset");
        MLP_HIP(m, hipMemset(m->vW[l], 0, nw * 4), "reset");
        MLP_HIP(m, hipMemset(m->mb[l], 0, (size_t)m->op[l] * 4), "reset");
        MLP_HIP(m, hipMemset(m->vb[l], 0, (size_t)m->op[l] * 4), "reset");
    }
    m->adam_t = 0;
    return WT_OK;
}

int wt_mlp_forward(wt_mlp *m, int32_t n, const float *X, float *Y) {
    MLP_CTX(m);
    if (n < 1 || !X || !Y) return mlp_err(m, WT_EINVAL, "need n >= 1 and non-NULL arrays");
    const int L = m->nl - 1, nout = m->out[L];
    const int chunk = 4096;
    for (int r0 = 0; r0 < n; r0 += chunk) {
        const int cnt = std::min(chunk, n - r0), rows = wtm::pad64(cnt);
        MLP_HIP(m, ensure_batch(m, rows), "batch buffers");
        MLP_HIP(m, load_host_batch(m, cnt, X + (size_t)r0 * m->in[0], nullptr, nullptr, rows), "upload");
        MLP_HIP(m, forward(m, rows), "forward");
        std::vector<float> z((size_t)rows * m->op[L]);
        MLP_HIP(m, hipMemcpyAsync(z.data(), m->Z[L], z.size() * 4, hipMemcpyDeviceToHost, m->stream), "copy");
        MLP_HIP(m, hipStreamSynchronize(m->stream), "forward");
        for (int i = 0; i < cnt; i++)
            for (int j = 0; j < nout; j++) Y[(size_t)(r0 + i) * nout + j] = z[(size_t)i * m->op[L] + j];
    }
    return WT_OK;
}

int wt_mlp_loss_grad(wt_mlp *m, int32_t n, const float *X, const float *T, const float *mask, double *loss,
                     float *grad) {
    MLP_CTX(m);
    if (n < 1 || !X || !T) return mlp_err(m, WT_EINVAL, "need n >= 1 and non-NULL inputs and targets");
    const int rows = wtm::pad64(n);
    MLP_HIP(m, ensure_batch(m, rows), "batch buffers");
    MLP_HIP(m, load_host_batch(m, n, X, T, mask, rows), "upload");
    MLP_HIP(m, forward(m, rows), "forward");
    MLP_HIP(m, backward(m, rows, m->loss), "backward");
    MLP_HIP(m, hipStreamSynchronize(m->stream), "loss_grad");
    if (loss) MLP_HIP(m, hipMemcpy(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost), "copy");
    if (grad) return copy_params(m, grad, 0, m->gW, m->gb);
    return WT_OK;
}

int wt_mlp_train_step(wt_mlp *m, int32_t n, const float *X, const float *T, const float *mask, float lr,
                      double *loss) {
    MLP_CTX(m);
    if (n < 1 || !X || !T) return mlp_err(m, WT_EINVAL, "need n >= 1 and non-NULL inputs and targets");
    if (!(lr > 0.f)) return mlp_err(m, WT_EINVAL, "learning rate must be positive");
    const int rows = wtm::pad64(n);
    MLP_HIP(m, ensure_batch(m, rows), "batch buffers");
    MLP_HIP(m, load_host_batch(m, n, X, T, mask, rows), "upload");
    MLP_HIP(m, forward(m, rows), "forward");
    MLP_HIP(m, backward(m, rows, m->loss), "backward");
    MLP_HIP(m, adam(m, lr), "adam");
    MLP_HIP(m, hipStreamSynchronize(m->stream), "train_step");
    if (loss) MLP_HIP(m, hipMemcpy(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost), "copy");
    return WT_OK;
}

int wt_mlp_set_data(wt_mlp *m, int32_t n, const float *X, const float *T, const float *mask) {
    MLP_CTX(m);
    if (n < 1 || !X || !T) return mlp_err(m, WT_EINVAL, "need n >= 1 and non-NULL inputs and targets");
    void *old[] = {m->setX, m->setT, m->setM, m->perm};
    for (void *p : old)
        if (p) (void)hipFree(p);
    m->setX = m->setT = m->setM = nullptr;
    m->perm = nullptr;
    m->nset = 0;
    const size_t nin = m->in[0], nout = m->out[m->nl - 1];
    MLP_HIP(m, dalloc(&m->setX, (size_t)n * nin), "training set");
    MLP_HIP(m, dalloc(&m->setT, (size_t)n * nout), "training set");
    MLP_HIP(m, dalloc(&m->setM, (size_t)n * nout), "training set");
    MLP_HIP(m, dalloc(&m->perm, (size_t)n), "training set");
    MLP_HIP(m, hipMemcpy(m->setX, X, (size_t)n * nin * 4, hipMemcpyHostToDevice), "copy");
    MLP_HIP(m, hipMemcpy(m->setT, T, (size_t)n * nout * 4, hipMemcpyHostToDevice), "copy");
    if (mask) {
        MLP_HIP(m, hipMemcpy(m->setM, mask, (size_t)n * nout * 4, hipMemcpyHostToDevice), "copy");
    } else {
        std::vector<float> ones((size_t)n * nout, 1.f);
        MLP_HIP(m, hipMemcpy(m->setM, ones.data(), ones.size() * 4, hipMemcpyHostToDevice), "copy");
    }
    m->nset = n;
    return WT_OK;
}

int wt_mlp_train(wt_mlp *m, int32_t epochs, int32_t batch, float lr, float lr_final, uint64_t seed,
                 double *loss_per_epoch) {
    MLP_CTX(m);
    if (m->nset < 1) return mlp_err(m, WT_ESTATE, "no training set (wt_mlp_set_data)");
    if (epochs < 1 || batch < 1 || !(lr > 0.f) || !(lr_final > 0.f))
        return mlp_err(m, WT_EINVAL, "need epochs >= 1, batch >= 1 and positive learning rates");
    const int rows = wtm::pad64(std::min<int64_t>(batch, m->nset));
    const int64_t steps_per_epoch = (m->nset + batch - 1) / batch;
    const int64_t total = steps_per_epoch * epochs;
    MLP_HIP(m, ensure_batch(m, rows), "batch buffers");
    if (m->loss_cap < steps_per_epoch) {
        (void)hipFree(m->loss);
        m->loss = nullptr;
        m->loss_cap = 0;
        MLP_HIP(m, dalloc(&m->loss, (size_t)steps_per_epoch), "loss buffer");
        m->loss_cap = steps_per_epoch;
    }
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm(m->nset);
    std::vector<double> losses(steps_per_epoch);
    const int nin = m->in[0], nout = m->out[m->nl - 1];
    const int width = std::max(m->ip[0], m->op[m->nl - 1]);
    int64_t step = 0;
    for (int ep = 0; ep < epochs; ep++) {
        for (int64_t i = 0; i < m->nset; i++) perm[i] = (int32_t)i;
        std::shuffle(perm.begin(), perm.end(), rng);
        MLP_HIP(m, hipMemcpyAsync(m->perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, m->stream), "perm");
        for (int64_t s = 0; s < steps_per_epoch; s++, step++) {
            const int64_t base = s * batch;
            const int cnt = (int)std::min<int64_t>(batch, m->nset - base);
            const int r = wtm::pad64(cnt);
            hipLaunchKernelGGL(wtm::k_gather, dim3(grid_for((int64_t)r * width)), dim3(256), 0, m->stream,
                               (const float *)m->setX, (const float *)m->setT, (const float *)m->setM,
                               (const int32_t *)m->perm, base, cnt, r, nin, m->ip[0], nout, m->op[m->nl - 1], m->Xb,
                               m->Tb, m->Mb);
            MLP_HIP(m, hipGetLastError(), "gather");
            MLP_HIP(m, forward(m, r), "forward");
            MLP_HIP(m, backward(m, r, m->loss + s), "backward");
            // cosine decay from lr to lr_final over all steps
            const double f = total > 1 ? (double)step / (double)(total - 1) : 1.0;
            const float lr_s = (float)(lr_final + 0.5 * (lr - lr_final) * (1.0 + std::cos(3.14159265358979 * f)));
            MLP_HIP(m, adam(m, lr_s), "adam");
        }
        MLP_HIP(m, hipMemcpyAsync(losses.data(), m->loss, sizeof(double) * steps_per_epoch, hipMemcpyDeviceToHost,
                                  m->stream), "loss");
        MLP_HIP(m, hipStreamSynchronize(m->stream), "train");
        if (loss_per_epoch) {
            double s = 0.0;
            for (double v : losses) s += v;
            loss_per_epoch[ep] = s / (double)steps_per_epoch;
        }
    }
    return WT_OK;
}

}  // extern "C"
