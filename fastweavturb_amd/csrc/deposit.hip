// Vorticity deposition (tube.f90:800-892) on the GPU.
//
// The reference walks every cell of each sub-segment's integer bounding box and keeps the
// cells that pass three tests (t0 >= 0, t1 < 0, rho_c < Rtube). Here the grid is cut into
// bricks of Run::bd cells, chosen per grid size (bricks.hip pairs every sub-segment with the bricks
// its box overlaps). A persistent workgroup owns one brick at a time and accumulates it in LDS, then
// writes the brick's cells once: the field sees no atomics and needs no clearing pass.
//
// Inside a brick each lane takes one (sub-segment, brick) pair and walks the box columns that
// fall in the brick, along the chord's dominant axis; per row of columns only the b-interval
// where the slab between the two end planes can reach the brick's cells is visited. For each
// column the two end planes give in closed form the short index range that can pass; only
// those candidates get the reference's exact FP64 test (the square root of the distance test
// is taken only when the squared distance is within 1e-14 of R^2), so the set of depositing
// (cell, sub-segment) pairs is exactly the reference's. Candidates are compacted per wavefront
// through an LDS queue and tested 64 at a time with every lane busy; the testing lane reads the
// candidate's sub-segment from the owning lane's registers by cross-lane shuffles, so no
// sub-segment is re-derived per candidate, and evaluates the contribution in place when the
// candidate is a hit (about 86% of them are).
//
// Precision modes:
//   double: the reference formulas in the reference order; LDS accumulation in FP64.
//   float : geometry, tests, zeta, the spline point and the offset to it in FP64; for constant
//           cores without twist the Gaussian core in FP32, otherwise the whole reference
//           formula in FP64; LDS accumulation in 64-bit fixed point (exact and
//           order-independent, so results are bitwise reproducible).
// This file is compiled with -ffp-contract=off; the FP32 evaluation re-enables contraction.
#include <type_traits>

#include "wt_device.h"

namespace wt {

constexpr int kDepWaves = 8;
constexpr int kDepThreads = kDepWaves * 64;

struct DepArgs {
    Geo g;
    Run r;                        // instance 0 (only used for the Run argument of subseg_load)
    Bins b;
    const BinRec *rec;            // the binning's boxes, per sub-point
    void *omega;                  // double* (FP64 mode) or float* (FP32 mode)
    double fx_limit;              // |value * scale| bound before overflow (FP32 mode; the scale
                                  // is the brick's instance's Run::fx_scale)
    unsigned long long *counters; // [0] columns [1] candidates [2] hits [3] overflow
    unsigned int *next;           // brick queue cursor
    int accumulate;               // 0: the brick's cells are written; 1: added to omega
};

// subseg_load for a sub-segment that the binning has already accepted (so it starts a deposited
// sub-segment of a tube in line_s..line_e of the brick's instance `ri`), given its section s and
// tube l (sub_sec[q], sub_tube[q]), with a shorter chain of dependent loads: the four sub-points
// q..q+3 (the last sub-segment of a section takes its second chord from the next section's first
// two sub-points, which are q+2 and q+3 there) and the section's and tube's values are read at
// once; the box is the one the binning computed (rc). Same values as subseg_load.
__device__ __forceinline__ void subseg_load_pair(const Geo &g, const Run &ri, int64_t q, int32_t s, int32_t l,
                                                 const BinRec &rc, SubSeg &o) {
    double p[12];
#pragma unroll
    for (int k = 0; k < 12; k++) p[k] = g.ci[3 * q + k];
    const int32_t j = (int32_t)(q - g.sp_base[s]);
    const int32_t nt = g.ntf[s];
    const double czl = g.czl[s], dtf = g.dtf[s];
    const double smin = g.sigma[l], len = g.length[l];
    const int nT = g.n_T[l];
    o.N = ri.N;
    o.dx = ri.dx;
    o.start = ri.start;
    o.field_off = ri.field_off;
    o.sec = s;
    o.tube = l;
    o.j = j;
    auto chord_of = [](const double *b, const double *a, double *t, double *nd) {
        double d[3];
        cminus(b, a, d);
        const double n = norm3(d);
        *nd = n;
#pragma unroll
        for (int c = 0; c < 3; c++) t[c] = d[c] / n;
    };
    double nd1;
    chord_of(p + 3, p, o.tci0, &o.ndci0);
    if (j + 1 < nt) chord_of(p + 6, p + 3, o.tci1, &nd1);
    else chord_of(p + 9, p + 6, o.tci1, &nd1);
    o.s0 = czl + dtf * (double)j;
    o.s1 = czl + dtf * (double)(j + 1);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o.ci0[c] = p[c];
        o.ci1[c] = p[3 + c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {  // mark1 (tube.f90:779-785)
        if (o.ci1[c] - o.ci0[c] > kPi) o.ci0[c] = o.ci0[c] + kPi2;
        else if (o.ci1[c] - o.ci0[c] < -kPi) o.ci1[c] = o.ci1[c] + kPi2;
    }
    const double smax = ri.sigma_ratio * smin;
    double sig = sigma_func(o.s0, len, smax, smin, nT);
    const double R = ri.Rtube_ratio * sig;
    sig = sigma_func(o.s1, len, smax, smin, nT);
    const double R2 = ri.Rtube_ratio * sig;
    o.R = (R2 > R) ? R2 : R;
#pragma unroll
    for (int c = 0; c < 3; c++) {  // tube.f90:793-798
        o.lo[c] = rc.lo[c];
        o.hi[c] = rc.hi[c];
    }
}

// Number of lanes below this one whose bit is set in the wave mask m, plus base.
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m, uint32_t base = 0u) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}


// ---- hit evaluation, FP64: tube.f90:819-892 in the reference order ----------------------
__device__ void hit_math_f64(const Geo &g, const SubSeg &sg, int32_t ip, int32_t jp, int32_t kp, double *v) {
    const Run &r = g.runs[g.tube_inst[sg.tube]];
    const int l = sg.tube;
    const double LD = g.Len_Disc[l], len = g.length[l], Gamma = g.gamma[l];
    const double smin = g.sigma[l], smax = r.sigma_ratio * smin;
    const int nT = g.n_T[l];
    const double *coe = g.coe + 18 * (int64_t)sg.sec;
    const double sfirst = g.czl[sg.sec];
    const double dzeta = r.dzeta_sup < g.dtf[sg.sec] ? r.dzeta_sup : g.dtf[sg.sec];
    double X[3] = {mesh(ip, sg), mesh(jp, sg), mesh(kp, sg)};
    double a0 = X[0] - sg.ci0[0], a1 = X[1] - sg.ci0[1], a2 = X[2] - sg.ci0[2];
    double t0 = (a0 * sg.tci0[0] + a1 * sg.tci0[1]) + a2 * sg.tci0[2];
    double t1 = ((X[0] - sg.ci1[0]) * sg.tci0[0] + (X[1] - sg.ci1[1]) * sg.tci0[1]) +
                (X[2] - sg.ci1[2]) * sg.tci0[2];
    double zeta = (sg.s1 * t0 - sg.s0 * t1) / sg.ndci0;
    double zsec = zeta - sfirst;
    double cz3[3], dc[3], ddc[3], cr[3], xc[3], T[3];
    spline_pos(zsec * LD, coe, cz3);
    spline_d1(zsec * LD, coe, dc);
    double ndc = norm3(dc);
#pragma unroll
    for (int c = 0; c < 3; c++) T[c] = dc[c] / ndc;
    spline_d2(zsec * LD, coe, ddc);
    cr[0] = dc[1] * ddc[2] - dc[2] * ddc[1];
    cr[1] = dc[2] * ddc[0] - dc[0] * ddc[2];
    cr[2] = dc[0] * ddc[1] - dc[1] * ddc[0];
    double ncr = norm3(cr);
    cminus(X, cz3, xc);
    double nxc = norm3(xc);
    double rho = nxc;
    double sigma = sigma_func(zeta, len, smax, smin, nT);
    double tp = exp(-(p2(rho - 0.0) / 2.0 / p2(sigma))) / (kPi2 * p2(sigma));
    double ftd = -exp(-(p2(sg.R) / 2.0 / p2(sigma))) / (kPi2 * p2(sigma));
    tp = tp + ftd;
    double sl = sigma_func(zeta - dzeta, len, smax, smin, nT);
    double sr = sigma_func(zeta + dzeta, len, smax, smin, nT);
    double dsds = (sr - sl) / (2.0 * dzeta);
    dsds = dsds / ndc / LD;
    const double eta = r.eta;
    double tg = tp * Gamma;
    if (ncr < 1.0e-8) {
        double er[3], eth[3];
#pragma unroll
        for (int c = 0; c < 3; c++) er[c] = xc[c] / (nxc + 1.0e-15);
        eth[0] = er[1] * T[2] - er[2] * T[1];
        eth[1] = er[2] * T[0] - er[0] * T[2];
        eth[2] = er[0] * T[1] - er[1] * T[0];
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = tg * ((T[c] + eta * rho * eth[c]) + (rho * dsds / sigma) * er[c]);
    } else {
        double B[3], Nv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) B[c] = cr[c] / ncr;
        Nv[0] = B[1] * T[2] - B[2] * T[1];
        Nv[1] = B[2] * T[0] - B[0] * T[2];
        Nv[2] = B[0] * T[1] - B[1] * T[0];
        double kap = ncr / p3(ndc);
        double costh = ((xc[0] * Nv[0] + xc[1] * Nv[1]) + xc[2] * Nv[2]) / (nxc + 1.0e-15);
        double sinth = ((xc[0] * B[0] + xc[1] * B[1]) + xc[2] * B[2]) / (nxc + 1.0e-15);
        double D = 1.0 - kap * rho * costh;
        double t5 = -(sinth * rho / D);
        double t6 = costh * rho / D;
        double t7 = costh * rho * dsds / sigma / D;
        double t8 = sinth * rho * dsds / sigma / D;
#pragma unroll
        for (int c = 0; c < 3; c++)
            v[c] = tg * ((T[c] + eta * (t5 * Nv[c] + t6 * B[c])) + (t7 * Nv[c] + t8 * B[c]));
    }
}

// ---- hit evaluation, FP32 mode ----------------------------------------------------------
// Constant core radius and no twist (sigma_ratio = 1, eta = 0; every benchmark workload): the
// reference's contribution reduces exactly to tp*Gamma*T (t5..t8 and the straight-branch extras
// vanish). The host selects this path only when every instance has sigma_ratio = 1 and eta = 0.
// The spline point, its tangent and the offset xc are FP64 values: the spline parameter is the
// reference's zeta rewritten as an affine function of t0 (equal up to rounding), point and
// tangent evaluated together by Horner's scheme with FMA, the spline
// wrapped to [-pi, pi), then cminus's single +-2pi correction, which is not a full minimum
// image: cells of an extended box more than 3pi from the wrapped spline point keep a far offset
// and receive the negative cut-off term, as in the reference. The Gaussian core runs in FP32 with
// the hardware reciprocal and reciprocal square root. Otherwise the curvature terms carry
// 1/(1 - kappa rho cos), which amplifies rounding near its zero, so the whole reference formula
// runs in FP64 (hit_math_f64).

// Per-hit values that live in global memory (the section's spline): loaded before the candidate
// test so that their latency overlaps it.
struct HitConst {
    double coe[18];
};

__device__ __forceinline__ void hit_const_load(const Geo &g, const SubSeg &sg, HitConst &hc) {
    const double *coe = g.coe + 18 * (int64_t)sg.sec;
#pragma unroll
    for (int k = 0; k < 18; k++) hc.coe[k] = coe[k];
}

// The Gaussian core of a pair's tube in single precision (constant core radius): the contribution
// is tp Gamma |c'|^-1 c' with tp Gamma = exp2(e rho^2) g - gc, e = -log2(e) / (2 sigma^2),
// g = Gamma / (2 pi sigma^2) and gc = g exp(-R^2 / (2 sigma^2)) (the cut-off term).
struct GaussK {
    float e, g, gc;
};

__device__ __forceinline__ GaussK gauss_consts(double sigma, double gamma, double R) {
    GaussK k;
    const float s = (float)sigma, is2 = __builtin_amdgcn_rcpf(s * s), Rf = (float)R;
    k.e = -0.5f * 1.4426950408889634f * is2;
    k.g = 0.15915494309189535f * is2 * (float)gamma;
    k.gc = __expf(-0.5f * Rf * Rf * is2) * k.g;
    return k;
}

// sg.s0, sg.s1 hold zA, zB (SegRec). Returns tg and the tangent dcf; the contribution is tg dcf.
__device__ float hit_math_f32(const HitConst &hc, const GaussK &gk, const SubSeg &sg, int32_t ip, int32_t jp, int32_t kp,
                              float *dcf) {
    const double *coe = hc.coe;
    double X[3] = {mesh(ip, sg), mesh(jp, sg), mesh(kp, sg)};
    double a[3] = {X[0] - sg.ci0[0], X[1] - sg.ci0[1], X[2] - sg.ci0[2]};
    double t0 = (a[0] * sg.tci0[0] + a[1] * sg.tci0[1]) + a[2] * sg.tci0[2];
    double z = sg.s0 + sg.s1 * t0;
    // spline point and tangent together (Horner with derivative, FP64 FMA; the derivative's first
    // step is the leading coefficient), point wrapped to [-pi, pi) as in coo_curve_spline_5th
    double cz3[3], dc[3], xcd[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double *a = coe + 6 * c;
        double dv = a[5], pv = __builtin_fma(a[5], z, a[4]);
#pragma unroll
        for (int k = 3; k >= 0; k--) {
            dv = __builtin_fma(dv, z, pv);
            pv = __builtin_fma(pv, z, a[k]);
        }
        cz3[c] = pv;
        dc[c] = dv;
    }
    // the wrap of the spline point and cminus's correction, each as one fma with an integer
    // multiple of 2 pi (fma(-1, 2 pi, c) rounds like c - 2 pi, and a zero multiple is exact)
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int k1 = cz3[c] >= kPi ? 1 : (cz3[c] < -kPi ? -1 : 0);
        const double d = X[c] - __builtin_fma(-(double)k1, kPi2, cz3[c]);
        const int k2 = d > kPi ? 1 : (d < -kPi ? -1 : 0);
        xcd[c] = __builtin_fma(-(double)k2, kPi2, d);
    }
    float xc[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        xc[c] = (float)xcd[c];
        dcf[c] = (float)dc[c];
    }
    {
#pragma clang fp contract(fast)
        float rho2 = xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2];
        float indc = __builtin_amdgcn_rsqf(dcf[0] * dcf[0] + dcf[1] * dcf[1] + dcf[2] * dcf[2]);
        float tpg = __builtin_fmaf(__builtin_amdgcn_exp2f(rho2 * gk.e), gk.g, -gk.gc);
        return tpg * indc;
    }
}

// The reference's candidate test (tube.f90:803-818) for one extended cell. rho < R is decided
// on the squared distance unless it lies within 1e-14 (relative) of R^2, where the rounded
// square root is compared as in the reference. The reference's first test, 0 < k <= nzp for
// k = mod(k_plus + nz - 1, nz) + 1, holds for every extended index in [1 - N, 2N], and the
// binning keeps every box in that range, so it is not repeated here.
__device__ __forceinline__ bool pair_hits_fast(const SubSeg &sg, int32_t ip, int32_t jp, int32_t kp, double R2lo,
                                               double R2hi) {
    double X0 = mesh(ip, sg), X1 = mesh(jp, sg), X2 = mesh(kp, sg);
    double a0 = X0 - sg.ci0[0], a1 = X1 - sg.ci0[1], a2 = X2 - sg.ci0[2];
    double t0 = (a0 * sg.tci0[0] + a1 * sg.tci0[1]) + a2 * sg.tci0[2];
    if (!(t0 >= 0.0)) return false;
    double t1 = ((X0 - sg.ci1[0]) * sg.tci1[0] + (X1 - sg.ci1[1]) * sg.tci1[1]) + (X2 - sg.ci1[2]) * sg.tci1[2];
    if (!(t1 < 0.0)) return false;
    double v = ((p2(a0) + p2(a1)) + p2(a2)) - p2(t0);
    if (v >= 0.0 && v < R2lo) return true;
    if (!(v < R2hi)) return false;
    return sqrt(v) < sg.R;
}

// Element i of a 3-vector by selects (a runtime index into a private array would live in scratch).
template <typename T>
__device__ __forceinline__ T pick3(const T *v, int i) {
    return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

template <bool F64ACC>
struct AccT {
    using type = unsigned long long;
};
template <>
struct AccT<true> {
    using type = double;
};

// Evaluates queue entries [qn - cnt, qn) on lanes 0..cnt-1 and adds them to the brick.
// LDS stride between the accumulator's components (the largest brick)
constexpr int kCells = kBrickCells;

// The sub-segment of the lane's pair, as the candidate test and the hit evaluation need it. It
// stays in the owning lane's registers; other lanes read it with cross-lane shuffles.
// FP32 mode keeps, in place of s0 and s1, the constants of z(t0) = zA + zB t0, the section's
// spline parameter as a function of t0 = (X - ci0) . tci0 (ndci0 is then unused).
struct SegRec {
    double ci0[3], ci1[3], tci0[3], tci1[3];
    double s0, s1, ndci0, R;
    int32_t sec, tube;
    GaussK gk;   // FP32 mode
};

// Column walk of one (sub-segment, brick image) pair, kept in LDS for the column stage. Columns
// run along the chord's dominant axis a; on column (b, c) the plane t0 = 0 sits at a-index
// (relative to lo_a) P0 - B0 db - C0 dc and the plane t1 = 0 at P1 - B1 db - C1 dc (db, dc
// relative to lo_b, lo_c). Every cell that can pass the reference's tests lies on a column
// whose crossing with the plane t0 = 0 is within Rp of ci0 (cells): with (ub, uc) the (b, c)
// position of ci0, (B0 Db + C0 Dc)^2 + Db^2 + Dc^2 < Rp^2 for Db = db - ub, Dc = dc - uc (the row
// stage's ellipse, whose constants stay in the pair's registers: RowBand).
// Single precision: these only select candidates, with margins m0, m1 of at least three times
// the worst-case rounding of each plane position over the box; the reference's FP64 tests decide.
struct ColWalk {
    float P0, B0, C0, P1, B1, C1;
    float m0, m1;
    int16_t lo_a, lo_b, lo_c, a0, a1;   // a0..a1: the pair's a-range in the brick image
    int16_t tag;                        // the candidates' tag (HitQ)
    int8_t up0, up1, use1;
};

// Margin of a single-precision plane position P - B db - C dc over db, dc in [0, W].
__device__ __forceinline__ float plane_margin(double P, double B, double C, double Wb, double Wc) {
    return (float)(1e-5 + 3.0e-7 * (fabs(P) + fabs(B) * Wb + fabs(C) * Wc));
}

// A row of columns of one pair: columns b = bl .. bl + len - 1 of row c.
struct RowSeg {
    uint32_t w0, w1;   // c | bl << 16 (16-bit two's complement each), len | pair << 8
    __device__ int32_t c() const { return (int32_t)(int16_t)(w0 & 0xffffu); }
    __device__ int32_t bl() const { return (int32_t)w0 >> 16; }
    __device__ int32_t len() const { return (int32_t)(w1 & 0xffu); }
    __device__ int32_t pair() const { return (int32_t)((w1 >> 8) & 0xffu); }
};
__device__ __forceinline__ RowSeg row_seg(int32_t c, int32_t bl, int32_t len, int32_t pair) {
    return RowSeg{((uint32_t)c & 0xffffu) | ((uint32_t)bl << 16), (uint32_t)len | ((uint32_t)pair << 8)};
}

// A candidate cell in the axes of its pair's column walk: a along the column, b and c across it
// (extended indices, stored + 32768 so that a column step adds to them without borrowing from
// the neighbouring field), and tag = owning lane | column axis << 6 | the pair's brick image on
// x, y, z (0..2 each) << 8, 10, 12: w0 = (a + 32768) | tag << 16, w1 = (b + 32768) | (c + 32768)
// << 16. The x, y, z indices are formed when the candidate is tested, once per candidate instead
// of once per queue round, and the cell within the brick from the image without a wrap.
struct HitQ {
    uint32_t w0, w1;
};
constexpr int32_t kIdxBias = 32768;

// 1/x from the hardware reciprocal refined by one Newton step (within a few ulp, never exactly
// rounded): only for quantities that feed single-precision bounds carrying their own margins.
// x must be nonzero and finite.
__device__ __forceinline__ double rcp_nr(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
}

// Per-pair constants of the row stage (registers, lane = pair). Each end-plane condition is a
// bound on db that is linear in dc; the ellipse's centre is linear in Dc = dc - uc and its squared
// half-width quadratic. The constants are formed in FP64 from the single-precision walk and
// rounded once; a row then costs two plane lines, one square root and the clipping.
struct RowBand {
    float aL, bL, aU, bU;          // bounds on db: db >= aL + bL dc, db <= aU + bU dc (plane lines)
    float cm, cp, g, k0, k1, uc;   // ellipse: db in [cm + g Dc - hw, cp + g Dc + hw], hw^2 = k0 - k1 Dc^2
    int32_t b0r, b1r;              // the pair's b-range relative to lo_b
    int32_t lo_b, lo_c;
};

// One plane's condition B db >= Q - C dc (s = +1) or B db <= Q - C dc (s = -1) as a line
// al + be dc bounding db, widened by 1e-6 of its largest magnitude over rows |dc| <= dmax plus
// 1e-5 (covers the rounding of al, be and of the fma that evaluates it; Q carries twice the
// column stage's margin). With B = 0 it is a condition on dc alone, applied to the row range c0..c1.
__device__ __forceinline__ void plane_line(double Q, double B, double C, double s, double dmax, int32_t lo_c,
                                           float &al, float &be, bool &lower, int32_t &c0, int32_t &c1) {
    if (B != 0.0) {
        const double iB = rcp_nr(B), a = Q * iB, b = -C * iB;
        const double e = 1e-6 * (fabs(a) + fabs(b) * dmax) + 1e-5;
        lower = (B > 0.0) == (s > 0.0);
        al = (float)(lower ? a - e : a + e);
        be = (float)b;
    } else {   // s C dc >= s Q
        lower = true;
        al = -1e9f;
        be = 0.0f;
        const double Cs = s * C, Qs = s * Q;
        if (Cs == 0.0) {
            if (Qs > 0.0) c1 = c0 - 1;
        } else {
            const double x = fmin(fmax(Qs / Cs, -1e6), 1e6), e = 1e-6 * fabs(x) + 1e-5;
            if (Cs > 0.0) c0 = max(c0, lo_c + (int32_t)ceil(x - e));
            else c1 = min(c1, lo_c + (int32_t)floor(x + e));
        }
    }
}

// Row constants of a pair whose walk is w, b-range b0..b1 and rows c0..c1 (narrowed here where an
// end plane is parallel to the b-axis). (ub, uc): the chord's crossing of the plane t0 = 0 relative
// to (lo_b, lo_c); Rp2: the squared ellipse radius, already widened by 1e-5 relative (which covers
// the rounding of the half-width).
__device__ __forceinline__ RowBand row_setup(const ColWalk &w, double ub, double uc, double Rp2, int32_t b0,
                                             int32_t b1, int32_t &c0, int32_t &c1) {
    RowBand rb;
    rb.lo_b = w.lo_b;
    rb.lo_c = w.lo_c;
    rb.b0r = b0 - w.lo_b;
    rb.b1r = b1 - w.lo_b;
    const double dmax = fmax(fabs((double)(c0 - w.lo_c)), fabs((double)(c1 - w.lo_c)));
    const double A0 = w.a0 - w.lo_a, A1 = w.a1 - w.lo_a;
    float al0, be0, al1 = -1e9f, be1 = 0.0f;
    bool lo0, lo1 = true;
    {   // plane 0: up0 needs B0 db >= P0 - C0 dc - 2 m0 - A1, else B0 db <= P0 - C0 dc + 2 m0 - A0
        const double m2 = 2.0 * (double)w.m0;
        const double Q = w.up0 ? (double)w.P0 - m2 - A1 : (double)w.P0 + m2 - A0;
        plane_line(Q, w.B0, w.C0, w.up0 ? 1.0 : -1.0, dmax, w.lo_c, al0, be0, lo0, c0, c1);
    }
    if (w.use1) {   // plane 1: up1 needs B1 db <= P1 - C1 dc + 2 m1 - A0, else B1 db >= P1 - C1 dc - 2 m1 - A1
        const double m2 = 2.0 * (double)w.m1;
        const double Q = w.up1 ? (double)w.P1 + m2 - A0 : (double)w.P1 - m2 - A1;
        plane_line(Q, w.B1, w.C1, w.up1 ? -1.0 : 1.0, dmax, w.lo_c, al1, be1, lo1, c0, c1);
    }
    // one line per side: plane 0's on its side, plane 1's on the other if it bounds that side
    // (where both bound the same side only plane 0's is kept, a looser band); none: -/+1e9
    rb.aL = lo0 ? al0 : (lo1 ? al1 : -1e9f);
    rb.bL = lo0 ? be0 : (lo1 ? be1 : 0.0f);
    rb.aU = !lo0 ? al0 : (!lo1 ? al1 : 1e9f);
    rb.bU = !lo0 ? be0 : (!lo1 ? be1 : 0.0f);
    // ellipse (1 + B0^2) Db^2 + 2 B0 C0 Db Dc + (1 + C0^2) Dc^2 < Rp2, Db = db - ub: for each Dc,
    // Db = -B0 C0 Dc / eb +- sqrt(Rp2 / eb - (1 + B0^2 + C0^2) Dc^2 / eb^2) with eb = 1 + B0^2
    const double B0 = w.B0, C0 = w.C0, eb = 1.0 + B0 * B0, ieb = rcp_nr(eb);
    rb.k0 = (float)(Rp2 * ieb);
    rb.k1 = (float)((eb + C0 * C0) * ieb * ieb);
    rb.g = (float)(-B0 * C0 * ieb);
    rb.cm = (float)(ub - 1e-3);
    rb.cp = (float)(ub + 1e-3);
    rb.uc = (float)uc;
    return rb;
}

// b-range of row ic whose columns can hold candidates: the slab interval reaches a-indices
// a0..a1 (each plane condition is linear in dc) and the column crosses the plane t0 = 0 inside
// the ellipse of radius Rp. A superset, clipped to b0..b1; false if empty.
__device__ __forceinline__ bool row_band(const RowBand &rb, int32_t ic, int32_t &bl, int32_t &bh) {
    const float dcn = (float)(ic - rb.lo_c);
    const float Dc = dcn - rb.uc;
    const float h2 = __builtin_fmaf(-rb.k1 * Dc, Dc, rb.k0);
    const float hw = __builtin_amdgcn_sqrtf(fmaxf(h2, 0.0f));
    const float lo = fmaxf(__builtin_fmaf(rb.bL, dcn, rb.aL), __builtin_fmaf(rb.g, Dc, rb.cm) - hw);
    const float hi = fminf(__builtin_fmaf(rb.bU, dcn, rb.aU), __builtin_fmaf(rb.g, Dc, rb.cp) + hw);
    // the b-range clip in integers (ceil(max(x, b)) = max(ceil(x), b) for an integer b)
    bl = rb.lo_b + max((int32_t)ceilf(lo), rb.b0r);
    bh = rb.lo_b + min((int32_t)floorf(hi), rb.b1r);
    return h2 >= 0.0f && bl <= bh;
}

// The a-range bounds of one row segment's columns: on row dc the plane positions are linear in
// db, so each plane gives one line q - B db that bounds a from below or from above, with the
// margin already applied (q = (P - C dc) -/+ m, rounded twice before the fma that evaluates it:
// within 1.8e-7 of |P| + |B| Wb + |C| Wc, under the margin's 3e-7). One lower and one upper line
// are kept: the end planes usually bound a from opposite sides; where both bound it from the
// same side (the chord turns by more than 90 degrees) only plane 0's line is kept there, a looser
// bound that can only add candidates. An unused slot holds B = 0 and q = -/+1e9. The lower bound
// is stored minus one, so that a column's count is floor(hi) - ceil(lo - 1) (a rounding of lo - 1
// can only lower it by one more: one extra candidate).
struct ColLines {
    float qL, BL;      // lower bound, minus one
    float qU, BU;      // upper bound
    int32_t A0, A1;    // the pair's a-range, relative to lo_a (A0 minus one)
    int32_t lo_a1, lo_b;        // lo_a + 1 (a-index of the count's origin), lo_b
};

__device__ __forceinline__ ColLines col_lines(const ColWalk &w, int32_t ic) {
    ColLines L;
    const float dcn = (float)(ic - w.lo_c);
    const float p0 = __builtin_fmaf(-w.C0, dcn, w.P0), p1 = __builtin_fmaf(-w.C1, dcn, w.P1);
    const bool up0 = w.up0, lo1 = w.use1 && !w.up1, hi1 = w.use1 && w.up1;
    if (up0) {   // plane 0 bounds a from below
        L.qL = (p0 - w.m0) - 1.0f;
        L.BL = w.B0;
        L.qU = hi1 ? p1 + w.m1 : 1e9f;
        L.BU = hi1 ? w.B1 : 0.0f;
    } else {     // from above
        L.qU = p0 + w.m0;
        L.BU = w.B0;
        L.qL = lo1 ? (p1 - w.m1) - 1.0f : -1e9f;
        L.BL = lo1 ? w.B1 : 0.0f;
    }
    L.A0 = w.a0 - w.lo_a - 1;
    L.A1 = w.a1 - w.lo_a;
    L.lo_a1 = w.lo_a + 1;
    L.lo_b = w.lo_b;
    return L;
}

// Candidate a-range of the column at db = ib - lo_b of a row segment: returns the count (at most
// one more than the cells between the bounds, never fewer), the first index minus lo_a1 in *ia0.
__device__ __forceinline__ int32_t column_range(const ColLines &L, float db, int32_t *ia0) {
    const int32_t flo = max((int32_t)ceilf(__builtin_fmaf(-L.BL, db, L.qL)), L.A0);
    const int32_t fhi = min((int32_t)floorf(__builtin_fmaf(-L.BU, db, L.qU)), L.A1);
    *ia0 = flo;
    return max(fhi - flo, 0);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The sub-segment of lane `own` (every lane of the wave must call this).
template <bool FAST>
__device__ __forceinline__ void fetch_owner(const SegRec &me, int own, const Run &ri, SubSeg &sg, GaussK &gk) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        sg.ci0[c] = __shfl(me.ci0[c], own, 64);
        sg.ci1[c] = __shfl(me.ci1[c], own, 64);
        sg.tci0[c] = __shfl(me.tci0[c], own, 64);
        sg.tci1[c] = __shfl(me.tci1[c], own, 64);
    }
    sg.s0 = __shfl(me.s0, own, 64);
    sg.s1 = __shfl(me.s1, own, 64);
    if (!FAST) sg.ndci0 = __shfl(me.ndci0, own, 64);
    sg.R = __shfl(me.R, own, 64);
    sg.sec = __shfl(me.sec, own, 64);
    if (FAST) {
        gk.e = __shfl(me.gk.e, own, 64);
        gk.g = __shfl(me.gk.g, own, 64);
        gk.gc = __shfl(me.gk.gc, own, 64);
    } else {
        sg.tube = __shfl(me.tube, own, 64);
    }
    sg.N = ri.N;
    sg.dx = ri.dx;
    sg.start = ri.start;
    sg.field_off = ri.field_off;
}

// LDS slot of brick cell (x, y, z) in each component's accumulator: rows along x, with x rotated by
// y + z within its row. A batch of hits from walks along x shares nearly one x (the end-plane slab
// is thinner than a cell); the rotation spreads those over the LDS banks, which a plain x-fastest
// layout would map to the same two banks.
__device__ __forceinline__ int acc_slot(int x, int y, int z, int lbx, int lby) {
    return ((x + y + z) & ((1 << lbx) - 1)) + (((z << lby) + y) << lbx);
}

// Per-wave work state of the deposition pipeline: pairs (lanes) -> row segments -> columns ->
// candidates (tested and deposited 64 at a time with every lane busy).
template <bool F64>
struct Wave {
    using Acc = typename AccT<F64>::type;
    const DepArgs *A;
    const Run *ri;
    Acc *acc;
    ColWalk *walks;     // [64]
    RowSeg *segs;       // [128]
    HitQ *candq;        // [128]
    int ox, oy, oz;
    int lane;
    int segn, candn;
    uint32_t *stat;               // LDS: candidates tested, hits, overflows of this wave
    uint32_t ncols;               // columns walked by this lane
};

// Tests candidates [candn - cnt, candn) with the reference predicate and adds the contribution of
// each hit to the brick (most candidates are hits, so they are evaluated in place rather than
// compacted again). In FP32 mode the hit's spline and tube constants are loaded before the test.
template <bool F64, bool FAST>
__device__ __forceinline__ void stage_candidates(Wave<F64> &W, const SegRec &me, int cnt) {
    const int lane = W.lane;
    const HitQ hq = W.candq[W.candn - cnt + (lane < cnt ? lane : 0)];
    wave_sync_lds();
    struct {
        int32_t ip, jp, kp, owner, tag;
    } h;
    {
        const int32_t tag = (int32_t)(hq.w0 >> 16), ax = (tag >> 6) & 3;
        h.tag = tag;
        const int32_t a = (int32_t)(hq.w0 & 0xffffu) - kIdxBias, b = (int32_t)(hq.w1 & 0xffffu) - kIdxBias,
                      c = (int32_t)(hq.w1 >> 16) - kIdxBias;
        h.owner = tag & 63;
        h.ip = ax == 0 ? a : (ax == 1 ? c : b);
        h.jp = ax == 0 ? b : (ax == 1 ? a : c);
        h.kp = ax == 0 ? c : (ax == 1 ? b : a);
    }
    W.candn = __builtin_amdgcn_readfirstlane(W.candn - cnt);
    SubSeg sg;
    GaussK gk;
    fetch_owner<FAST>(me, lane < cnt ? h.owner : lane, *W.ri, sg, gk);
    bool hit = false, ovf = false;
    if (lane < cnt) {
        HitConst hc;
        if (FAST) {
            hit_const_load(W.A->g, sg, hc);
            // keeps the spline loads here, ahead of the test, rather than sunk into the hit branch
            asm volatile("" ::: "memory");
        }
        const double R2 = sg.R * sg.R;
        hit = pair_hits_fast(sg, h.ip, h.jp, h.kp, R2 * (1.0 - 1e-14), R2 * (1.0 + 1e-14));
        if (hit) {
            // the cell within the brick: extended index - (origin + 1) - (image - 1) N
            const int N = W.ri->N;
            const int x = h.ip - (W.ox + 1 - N) - (int)__umul24(__builtin_amdgcn_ubfe((uint32_t)h.tag, 8, 2), N);
            const int y = h.jp - (W.oy + 1 - N) - (int)__umul24(__builtin_amdgcn_ubfe((uint32_t)h.tag, 10, 2), N);
            const int z = h.kp - (W.oz + 1 - N) - (int)__umul24(__builtin_amdgcn_ubfe((uint32_t)h.tag, 12, 2), N);
            const int cell = acc_slot(x, y, z, W.ri->lbd[0], W.ri->lbd[1]);
            if (FAST) {
                // fixed point: (tg 2^k) dcf is 2^k (tg dcf) in single precision (a power-of-two
                // scale), rounded to an integer; the 64-bit integer is assembled from its halves
                float dcf[3];
                const float tgs = hit_math_f32(hc, gk, sg, h.ip, h.jp, h.kp, dcf) * (float)W.ri->fx_scale;
                const float lim = (float)W.A->fx_limit;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    float s = tgs * dcf[c];
                    if (!(fabsf(s) < lim)) { ovf = true; s = 0.0f; }
                    const double r = (double)__builtin_rintf(s);
                    const double hi = __builtin_floor(r * 0x1p-32);
                    const double lo = __builtin_fma(hi, -0x1p32, r);
                    const unsigned long long fx =
                        ((unsigned long long)(uint32_t)(int32_t)hi << 32) | (unsigned long long)(uint32_t)lo;
                    atomicAdd((unsigned long long *)W.acc + c * kCells + cell, fx);
                }
            } else {
                double v[3];
                hit_math_f64(W.A->g, sg, h.ip, h.jp, h.kp, v);
                if (F64) {
#pragma unroll
                    for (int c = 0; c < 3; c++) unsafeAtomicAdd((double *)W.acc + c * kCells + cell, v[c]);
                } else {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        double s = v[c] * W.ri->fx_scale;
                        if (!(fabs(s) < W.A->fx_limit)) { ovf = true; s = 0.0; }
                        long long fx = (long long)rint(s);
                        atomicAdd((unsigned long long *)W.acc + c * kCells + cell, (unsigned long long)fx);
                    }
                }
            }
        }
    }
    // the wave's counts live in LDS, so that no loop around this stage carries them in registers
    const uint32_t nh = (uint32_t)__popcll(__ballot(hit)), no = (uint32_t)__popcll(__ballot(ovf));
    if (lane == 0) {
        __hip_atomic_fetch_add(W.stat + 0, (uint32_t)cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_fetch_add(W.stat + 1, nh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        if (no) __hip_atomic_fetch_add(W.stat + 2, no, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
}

// Columns walked per row segment and pass (Run::colcap): a longer segment's remainder is queued
// again, so that a pass is not held for its longest segment while the other lanes idle.

// Walks the columns of row segments [segn - cnt, segn), one segment per lane and one column of
// it per step, at most colcap columns each (the remainders go back to the segment queue, which
// has room: at most 63 segments stay queued plus at most 64 remainders); candidates go to the
// candidate queue.
template <bool F64, bool FAST>
__device__ __forceinline__ void stage_columns(Wave<F64> &W, const SegRec &me, int cnt) {
    const int lane = W.lane;
    int32_t len = 0, ib = 0, ic = 0, own = 0, wtag = 0;
    ColLines L{};   // idle lanes: zero lines, their counts are masked
    if (lane < cnt) {
        const RowSeg r = W.segs[W.segn - cnt + lane];
        len = r.len();
        ib = r.bl();
        ic = r.c();
        own = r.pair();
        const ColWalk w = W.walks[own];
        wtag = w.tag;
        L = col_lines(w, ic);
    }
    // the candidate's first word is column_range's index plus wb (a = index + lo_a1; no carry:
    // a + bias < 2^16)
    const uint32_t wb = ((uint32_t)wtag << 16) + (uint32_t)(kIdxBias + L.lo_a1);
    W.segn = __builtin_amdgcn_readfirstlane(W.segn - cnt);
    const int32_t cap = W.ri->colcap;
    const int32_t bl0 = ib, steps = len < cap ? len : cap;
    // the candidate's second word for column ib: (b + bias) | (c + bias) << 16
    uint32_t w1 = (uint32_t)(ib + kIdxBias) | ((uint32_t)(ic + kIdxBias) << 16);
    // two columns per lane and step (about a third of the columns hold a candidate, so the
    // queue rounds of both are shared)
    float db = (float)(ib - L.lo_b);   // exact: |db| < 2^24
    for (int32_t rem = steps; __ballot(rem > 0) != 0ull; rem -= 2, db += 2.0f, w1 += 2) {
        int32_t ia0, ia1;
        int32_t nc0 = column_range(L, db, &ia0), nc1 = column_range(L, db + 1.0f, &ia1);
        nc0 = rem > 0 ? nc0 : 0;
        nc1 = rem > 1 ? nc1 : 0;
        const int32_t nc = nc0 + nc1, ia1s = ia1 - nc0;
        W.candn = __builtin_amdgcn_readfirstlane(W.candn);   // a scalar through the rounds
        unsigned long long m;
        for (int u = 0; (m = __ballot(u < nc)) != 0ull; u++) {
            if (u < nc) {
                const bool second = u >= nc0;
                HitQ q;
                q.w0 = (uint32_t)((second ? ia1s : ia0) + u) + wb;
                q.w1 = second ? w1 + 1u : w1;
                W.candq[lanes_below(m, (uint32_t)W.candn)] = q;
            }
            W.candn = __builtin_amdgcn_readfirstlane(W.candn + __popcll(m));
            wave_sync_lds();
            if (W.candn >= 64) stage_candidates<F64, FAST>(W, me, 64);
        }
    }
    const bool more = len > cap;
    const unsigned long long mr = __ballot(more);
    if (more) W.segs[lanes_below(mr, (uint32_t)W.segn)] =
        row_seg(ic, bl0 + cap, len - cap, own);
    W.segn = __builtin_amdgcn_readfirstlane(W.segn + __popcll(mr));
    wave_sync_lds();
}

template <bool F64, bool FAST>
__global__ __launch_bounds__(kDepThreads) void k_deposit(DepArgs A) {
    using Acc = typename AccT<F64>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    Acc *acc = (Acc *)smem;                                                  // 3 * kCells
    unsigned char *wbase = smem + sizeof(Acc) * 3 * kCells;
    constexpr size_t kWaveBytes = sizeof(ColWalk) * 64 + sizeof(RowSeg) * 128 + sizeof(HitQ) * 128;
    unsigned long long *s_misc = (unsigned long long *)(wbase + kWaveBytes * kDepWaves);  // brick, cursor
    uint32_t *s_stat = (uint32_t *)(s_misc + 2);                                         // [kDepWaves][4]
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    Wave<F64> W;
    W.A = &A;
    W.acc = acc;
    {
        unsigned char *p = wbase + kWaveBytes * wv;
        W.walks = (ColWalk *)p;
        p += sizeof(ColWalk) * 64;
        W.segs = (RowSeg *)p;
        p += sizeof(RowSeg) * 128;
        W.candq = (HitQ *)p;
    }
    W.lane = lane;
    W.segn = W.candn = 0;
    W.ncols = 0;
    W.stat = s_stat + 4 * wv;
    if (lane < 4) W.stat[lane] = 0u;
    const Geo &g = A.g;
    for (;;) {
        if (threadIdx.x == 0) {
            s_misc[0] = (unsigned long long)atomicAdd(A.next, 1u);
            s_misc[1] = 0ull;
        }
        __syncthreads();
        const int64_t bi = (int64_t)__builtin_amdgcn_readfirstlane((unsigned)s_misc[0]);
        if (bi >= A.b.nbricks) break;
        const int4 bt = A.b.brick_tab[bi];
        const Run ri = g.runs[bt.x];
        W.ri = &ri;
        const double inv_dx = 1.0 / ri.dx;   // cells per unit length
        const int N = ri.N;
        const int bdx = ri.bd[0], bdy = ri.bd[1], bdz = ri.bd[2], lbx = ri.lbd[0], lby = ri.lbd[1];
        const int ox = bt.y * bdx, oy = bt.z * bdy, oz = bt.w * bdz;
        const int ex = N - ox < bdx ? N - ox : bdx;
        const int ey = N - oy < bdy ? N - oy : bdy;
        const int ez = N - oz < bdz ? N - oz : bdz;
        const int ncell = bdx * bdy * bdz;
        W.ox = ox;
        W.oy = oy;
        W.oz = oz;
        for (int i = threadIdx.x; i < ncell; i += kDepThreads) {
#pragma unroll
            for (int c = 0; c < 3; c++) acc[c * kCells + i] = (Acc)0;
        }
        __syncthreads();
        const int64_t p0 = A.b.brick_off[bi], p1 = A.b.brick_off[bi + 1];
        // The brick's list starts with its H pairs of wide boxes (each a large share of the brick's
        // work; neighbours along a tube sit next to each other). They come first, in Kh chunks whose
        // chunk k holds pairs k, k + Kh, k + 2 Kh, ... (lane j: pair k + j Kh), so that they spread
        // over the waves instead of holding one wave while the others wait at the brick's barrier:
        // Kh is ceil(H / 64) rounded up to a multiple of the workgroup's waves (every wave gets wide
        // work; a chunk's lanes share it through the wave's queues), at most H. The other pairs follow
        // in runs of 64 (neighbouring thin sub-segments of about the same cost share a wave). Chunks
        // are claimed one step ahead, and the next chunk's pair index and image read while the
        // current one is processed.
        const int64_t P = p1 - p0, H = A.b.brick_wide[bi];
        const int64_t Kw = (((H + 63) >> 6) + kDepWaves - 1) / kDepWaves * kDepWaves;
        const int64_t Kh = H < Kw ? H : Kw, K = Kh + ((P - H + 63) >> 6);
        auto pair_of = [&](int64_t k) -> int64_t {
            if (k < Kh) {
                const int64_t i = k + (int64_t)lane * Kh;
                return i < H ? i : P;
            }
            return H + ((k - Kh) << 6) + lane;
        };
        unsigned long long next = 0;
        if (lane == 0) next = atomicAdd(&s_misc[1], 1ull);
        next = __shfl(next, 0, 64);
        uint32_t q_next = 0, warm = 0;
        int32_t code_next = 0, s_next = 0, l_next = 0;
        BinRec rc_next{};
        int64_t pn = pair_of((int64_t)next);
        if ((int64_t)next < K && pn < P) {
            q_next = A.b.pairs[p0 + pn];
            code_next = A.b.pair_img[p0 + pn];
            s_next = g.sub_sec[q_next];
            l_next = g.sub_tube[q_next];
            rc_next = A.rec[q_next];
        }
        for (;;) {
            const unsigned long long chunk = next;
            if ((int64_t)chunk >= K) break;
            const int64_t pi = pn;
            const uint32_t q_cur = q_next;
            const int32_t code = code_next, s_cur = s_next, l_cur = l_next;
            const BinRec rc_cur = rc_next;
            if (lane == 0) next = atomicAdd(&s_misc[1], 1ull);
            next = __shfl(next, 0, 64);
            pn = pair_of((int64_t)next);
            const bool have_next = (int64_t)next < K && pn < P;
            if (have_next) {
                q_next = A.b.pairs[p0 + pn];
                code_next = A.b.pair_img[p0 + pn];
            }
            // ---- stage 0 (lane = pair): the sub-segment to registers, its column walk to LDS ----
            SegRec me;
            RowBand rb;
            int32_t b0 = 0, b1 = -1, c0 = 0, c1 = -1;
            bool have = pi < P;
            {
                SubSeg sg;
                ColWalk w;
                asm volatile("" ::"v"(warm));
                if (have) subseg_load_pair(g, ri, (int64_t)q_cur, s_cur, l_cur, rc_cur, sg);
                if (have) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        me.ci0[c] = sg.ci0[c];
                        me.ci1[c] = sg.ci1[c];
                        me.tci0[c] = sg.tci0[c];
                        me.tci1[c] = sg.tci1[c];
                    }
                    me.s0 = sg.s0;
                    me.s1 = sg.s1;
                    me.ndci0 = sg.ndci0;
                    if (FAST) {
                        // zeta = (s1 t0 - s0 t1) / ndci0 with t1 = t0 - L, L = (ci1 - ci0) . tci0, so
                        // z = (zeta - czl) LD = zA + zB t0
                        const double L = ((sg.ci1[0] - sg.ci0[0]) * sg.tci0[0] + (sg.ci1[1] - sg.ci0[1]) * sg.tci0[1]) +
                                         (sg.ci1[2] - sg.ci0[2]) * sg.tci0[2];
                        const double LD = g.Len_Disc[sg.tube], ind = 1.0 / sg.ndci0;
                        me.s0 = (sg.s0 * L * ind - g.czl[sg.sec]) * LD;
                        me.s1 = (sg.s1 - sg.s0) * LD * ind;
                        me.gk = gauss_consts(g.sigma[sg.tube], g.gamma[sg.tube], sg.R);
                    }
                    me.R = sg.R;
                    me.sec = sg.sec;
                    me.tube = sg.tube;
                    // the pair's cells on each axis: the box within the brick's periodic image
                    const int32_t img[3] = {code % 3, (code / 3) % 3, code / 9};
                    const int32_t org[3] = {ox, oy, oz}, ext[3] = {ex, ey, ez};
                    int32_t clo[3], chi[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const int32_t sh = (img[d] - 1) * N;
                        clo[d] = sg.lo[d] > org[d] + 1 + sh ? sg.lo[d] : org[d] + 1 + sh;
                        chi[d] = sg.hi[d] < org[d] + ext[d] + sh ? sg.hi[d] : org[d] + ext[d] + sh;
                    }
                    const int ax = dominant_axis(sg.tci0);
                    const int bx = (ax + 1) % 3, cx = (ax + 2) % 3;
                    w.tag = (int16_t)(lane | (ax << 6) | (img[0] << 8) | (img[1] << 10) | (img[2] << 12));
                    w.lo_a = (int16_t)pick3(sg.lo, ax);
                    w.lo_b = (int16_t)pick3(sg.lo, bx);
                    w.lo_c = (int16_t)pick3(sg.lo, cx);
                    w.a0 = (int16_t)pick3(clo, ax);
                    w.a1 = (int16_t)pick3(chi, ax);
                    b0 = pick3(clo, bx);
                    b1 = pick3(chi, bx);
                    c0 = pick3(clo, cx);
                    c1 = pick3(chi, cx);
                    const double t0a = pick3(sg.tci0, ax), it0a = rcp_nr(t0a);
                    const double ua = (pick3(sg.ci0, ax) - sg.start) * inv_dx + 1.0 - w.lo_a;
                    const double ub = (pick3(sg.ci0, bx) - sg.start) * inv_dx + 1.0 - w.lo_b;
                    const double uc = (pick3(sg.ci0, cx) - sg.start) * inv_dx + 1.0 - w.lo_c;
                    const double B0 = pick3(sg.tci0, bx) * it0a, C0 = pick3(sg.tci0, cx) * it0a;
                    const double Wb = pick3(sg.hi, bx) - pick3(sg.lo, bx), Wc = pick3(sg.hi, cx) - pick3(sg.lo, cx);
                    w.B0 = (float)B0;
                    w.C0 = (float)C0;
                    w.P0 = (float)((ua + B0 * ub) + C0 * uc);
                    w.m0 = plane_margin((ua + B0 * ub) + C0 * uc, B0, C0, Wb, Wc);
                    w.m1 = 0.0f;
                    w.up0 = t0a > 0.0;
                    const double t1a = pick3(sg.tci1, ax);
                    w.use1 = fabs(t1a) >= 0.05;
                    w.P1 = w.B1 = w.C1 = 0.0f;
                    w.up1 = 1;
                    // cells between the planes within R of the chord line lie at most
                    // (ndci0 + R tan(delta)) from the plane t0 = 0 (delta: angle between the plane
                    // normals); their columns cross that plane within Rp of ci0
                    const double cd = (sg.tci0[0] * sg.tci1[0] + sg.tci0[1] * sg.tci1[1]) + sg.tci0[2] * sg.tci1[2];
                    double Rp = 1e30;
                    if (cd > 0.2) {
                        const double tand = sqrt(fmax(1.0 - cd * cd, 0.0)) * rcp_nr(cd);
                        Rp = (sg.R + (sg.ndci0 + sg.R * tand) * sqrt(fmax(1.0 - t0a * t0a, 0.0)) * fabs(it0a)) * inv_dx;
                    }
                    if (w.use1) {
                        const double va = (pick3(sg.ci1, ax) - sg.start) * inv_dx + 1.0 - w.lo_a;
                        const double vb = (pick3(sg.ci1, bx) - sg.start) * inv_dx + 1.0 - w.lo_b;
                        const double vc = (pick3(sg.ci1, cx) - sg.start) * inv_dx + 1.0 - w.lo_c;
                        const double it1a = rcp_nr(t1a);
                        const double B1 = pick3(sg.tci1, bx) * it1a, C1 = pick3(sg.tci1, cx) * it1a;
                        w.B1 = (float)B1;
                        w.C1 = (float)C1;
                        w.P1 = (float)((va + B1 * vb) + C1 * vc);
                        w.m1 = plane_margin((va + B1 * vb) + C1 * vc, B1, C1, Wb, Wc);
                        w.up1 = t1a > 0.0;
                    }
                    W.walks[lane] = w;
                    rb = row_setup(w, ub, uc, fmin(Rp * Rp * (1.0 + 1e-5) + 1e-3, 1e30), b0, b1, c0, c1);
                } else {
#pragma unroll
                    for (int c = 0; c < 3; c++) me.ci0[c] = me.ci1[c] = me.tci0[c] = me.tci1[c] = 0.0;
                    me.s0 = me.s1 = me.ndci0 = me.R = 0.0;
                    me.gk = GaussK{0.0f, 0.0f, 0.0f};
                    me.sec = me.tube = 0;
                }
            }
            wave_sync_lds();
            if (have_next) {   // q_next has arrived by now
                s_next = g.sub_sec[q_next];
                l_next = g.sub_tube[q_next];
                rc_next = A.rec[q_next];
                // one word of the next pair's sub-points, so that their cache lines are on their way
                // while this chunk runs (the value is only waited for at the next set-up)
                warm = ((const uint32_t *)(g.ci + 3 * (int64_t)q_next))[0];
            }
            // ---- stage 1 (lane = pair): one row of the pair per step, non-empty bands queued ----
            for (int32_t ic = c0; __ballot(have && ic <= c1); ic++) {
                int32_t bl = 0, bh = -1;
                const bool sgok = have && ic <= c1 && row_band(rb, ic, bl, bh);
                W.ncols += sgok ? (uint32_t)(bh - bl + 1) : 0u;
                const unsigned long long m = __ballot(sgok);
                if (sgok) W.segs[lanes_below(m, (uint32_t)W.segn)] =
                    row_seg(ic, bl, bh - bl + 1, lane);
                W.segn = __builtin_amdgcn_readfirstlane(W.segn + __popcll(m));
                wave_sync_lds();
                while (W.segn >= 64) stage_columns<F64, FAST>(W, me, 64);
            }
            // the registers of these pairs are replaced by the next chunk: flush every stage
            while (W.segn > 0) stage_columns<F64, FAST>(W, me, W.segn);
            if (W.candn > 0) stage_candidates<F64, FAST>(W, me, W.candn);
        }
        __syncthreads();
        // the brick's cells are owned by this workgroup: plain stores (or read-add-store), once
        const int64_t N3 = (int64_t)N * N * N;
        for (int i = threadIdx.x; i < ncell; i += kDepThreads) {
            const int x = i & (bdx - 1), y = (i >> lbx) & (bdy - 1), z = i >> (lbx + lby);
            if (x < ex && y < ey && z < ez) {
                const int64_t idx = ri.field_off + (ox + x) + (int64_t)N * ((oy + y) + (int64_t)N * (oz + z));
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    if (F64) {
                        double *o = (double *)A.omega + idx + c * N3;
                        const double v = (double)acc[c * kCells + acc_slot(x, y, z, lbx, lby)];
                        *o = A.accumulate ? *o + v : v;
                    } else {
                        float *o = (float *)A.omega + idx + c * N3;
                        const double v = (double)(long long)acc[c * kCells + acc_slot(x, y, z, lbx, lby)] * ri.fx_inv;
                        *o = (float)(A.accumulate ? (double)*o + v : v);
                    }
                }
            }
        }
        __syncthreads();
    }
    // counters: one atomic per wave (the column counts summed over the lanes first)
    unsigned long long ncols = W.ncols;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ncols += __shfl_xor(ncols, o, 64);
    wave_sync_lds();
    if (lane == 0) {
        atomicAdd(A.counters + 0, ncols);
        atomicAdd(A.counters + 1, (unsigned long long)W.stat[0]);
        atomicAdd(A.counters + 2, (unsigned long long)W.stat[1]);
        if (W.stat[2]) atomicAdd(A.counters + 3, (unsigned long long)W.stat[2]);
    }
}

template <bool F64, bool FAST>
static hipError_t deposit_t(wt_ctx *c, DepArgs &A) {
    hipStream_t st = c->stream;
    using Acc = typename AccT<F64>::type;
    const size_t wave_bytes = sizeof(ColWalk) * 64 + sizeof(RowSeg) * 128 + sizeof(HitQ) * 128;
    const size_t lds = sizeof(Acc) * 3 * kCells + wave_bytes * kDepWaves + 2 * sizeof(unsigned long long) +
                       4 * sizeof(uint32_t) * kDepWaves;
    const void *fn = (const void *)k_deposit<F64, FAST>;
    WT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int cus = 256, per_cu = 0;
    WT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    WT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kDepThreads, lds));
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    int64_t blocks = (int64_t)cus * per_cu;
    if (blocks > A.b.nbricks) blocks = A.b.nbricks;
    if (blocks < 1) blocks = 1;
    kt_begin(c, kSlotDeposit);
    hipLaunchKernelGGL((k_deposit<F64, FAST>), dim3((unsigned)blocks), dim3(kDepThreads), lds, st, A);
    kt_end(c, kSlotDeposit);
    return hipGetLastError();
}

// The deposit kernel with counters zeroed on the device (context-owned, 5 words).
static hipError_t deposit_kernels(wt_ctx *c, bool accumulate) {
    hipStream_t st = c->stream;
    const bool f64 = c->precision == 64;
    if (!c->dep_cnt) {
        if (c->capturing) return hipErrorInvalidValue;     // the eager run allocates them
        WT_HIP(hipMalloc((void **)&c->dep_cnt, 5 * sizeof(unsigned long long)));
    }
    unsigned long long *cnt = c->dep_cnt;
    WT_HIP(fill_zero(cnt, 5 * sizeof(unsigned long long), st));
    DepArgs A{};
    A.g = c->geo;
    A.r = c->run;
    A.b = c->bins;
    A.rec = c->binws.rec;
    A.omega = c->omega;
    A.counters = cnt;
    A.next = (unsigned int *)(cnt + 4);
    A.accumulate = accumulate ? 1 : 0;
    if (!f64) {
        A.fx_limit = ldexp(1.0, 62) / 64.0;
    }
    bool fast = !f64;   // FP32 core only for constant core radius and no twist in every instance
    for (const Run &r : c->runs_host)
        if (r.eta != 0.0 || r.sigma_ratio != 1.0) fast = false;
    return f64 ? deposit_t<true, false>(c, A) : fast ? deposit_t<false, true>(c, A) : deposit_t<false, false>(c, A);
}

hipError_t launch_deposit(wt_ctx *c, bool accumulate) {
    hipStream_t st = c->stream;
    WT_HIP(deposit_kernels(c, accumulate));
    unsigned long long h[4];
    WT_HIP(hipMemcpyAsync(h, c->dep_cnt, sizeof h, hipMemcpyDeviceToHost, st));
    WT_HIP(hipStreamSynchronize(st));
    c->dep_stats.columns = (int64_t)h[0];
    c->dep_stats.candidates = (int64_t)h[1];
    c->dep_stats.hits = (int64_t)h[2];
    c->dep_stats.overflow = (int32_t)(h[3] > 0);
    c->dep_stats.fixed_point_shift = c->precision == 64 ? 0 : c->fixed_shift;
    c->dep_stats.pairs = c->bins.npairs;
    return hipSuccess;
}

// The deposit of the last binning again (graph capture): the counters stay on the device, where
// they equal the eager run's for the same tubes.
hipError_t replay_deposit(wt_ctx *c) { return deposit_kernels(c, false); }

}  // namespace wt
