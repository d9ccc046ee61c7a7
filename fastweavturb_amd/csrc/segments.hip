// Geometry of the woven field on the GPU, FP64 (compiled with -ffp-contract=off):
//   points   -> cze, Len_Disc, ck, ckk                   (tube.f90:392-515)
//   sections -> quintic spline coefficients, ntf, dtf   (tube.f90:543-642,
//                                                         subroutine.f90:440-488)
//   sub-points with the seam fix                         (tube.f90:706-723)
//   per sub-segment bounding box -> column work items    (tube.f90:766-798)
// Every value is produced by the same IEEE operation sequence as the reference statement it
// restates, so integers (ntf, boxes) and coordinates match the FP64 oracle bit for bit.
#include "wt_device.h"

namespace wt {

// Fill out[i] = l for every i in [base[l], base[l+1]); one block per tube.
__global__ void k_fill_owner(const int64_t *__restrict__ base, int32_t n, int32_t *__restrict__ out) {
    int l = blockIdx.x;
    if (l >= n) return;
    for (int64_t i = base[l] + threadIdx.x; i < base[l + 1]; i += blockDim.x) out[i] = l;
}

__device__ __forceinline__ double wrapped_absdiff(double a, double b) {
    double d = fabs(a - b);
    if (d > kPi) d = d - kPi2;
    return d;
}

// tube.f90:419-435: chord to the previous point; at the first point, the closing chord.
__global__ void k_chords(Geo g) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= g.npoint_all) return;
    int l = g.pt_tube[p];
    int64_t b = g.pt_base[l];
    int64_t e = g.pt_base[l + 1] - 1;
    int64_t a = (p == b) ? e : p - 1;
    int64_t c = (p == b) ? b : p;
    // reference: first point pairs (cx(1), cx(npoint)); others (cx(i), cx(i-1))
    double dx = wrapped_absdiff(g.cx[c], g.cx[a]);
    double dy = wrapped_absdiff(g.cy[c], g.cy[a]);
    double dz = wrapped_absdiff(g.cz[c], g.cz[a]);
    g.chord[p] = sqrt((p2(dx) + p2(dy)) + p2(dz));
}

// tube.f90:418-437 sequential cumulative sum, one thread per tube (kept sequential so the
// rounding matches the reference), then Len_Disc and length.
// The chords are read 16 at a time into registers so that the serial chain waits on one memory
// round trip per 16 additions.
__global__ void k_cumsum(Geo g, double *__restrict__ tot_out) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= g.nline) return;
    int64_t b = g.pt_base[l], e = g.pt_base[l + 1];
    double acc = 0.0;
    g.cze[b] = 0.0;
    constexpr int kChunk = 16;
    for (int64_t i0 = b + 1; i0 < e; i0 += kChunk) {
        double buf[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; k++) buf[k] = i0 + k < e ? g.chord[i0 + k] : 0.0;
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            if (i0 + k < e) {
                acc = acc + buf[k];
                g.cze[i0 + k] = acc;
            }
        }
    }
    double tot = acc + g.chord[b];
    double LD = tot / kPi2;
    g.Len_Disc[l] = LD;
    g.length[l] = LD * kPi2;
    tot_out[l] = tot;
}

// tube.f90:437: cze = cze*pi2/(cze(npoint) + dist)
__global__ void k_cze_norm(Geo g, const double *__restrict__ tot) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= g.npoint_all) return;
    g.cze[p] = g.cze[p] * kPi2 / tot[g.pt_tube[p]];
}

// tube.f90:439-466 (ck, with min-image wrap) and 468-489 (ckk, no wrap)
template <bool SECOND>
__global__ void k_derivative(Geo g) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= g.npoint_all) return;
    int l = g.pt_tube[p];
    int64_t b = g.pt_base[l], e = g.pt_base[l + 1] - 1;
    int64_t ip = (p == e) ? b : p + 1;
    int64_t im = (p == b) ? e : p - 1;
    double ds;
    if (p == b) ds = g.cze[b + 1] - g.cze[e] + kPi2;
    else if (p == e) ds = g.cze[b] - g.cze[e - 1] + kPi2;
    else ds = g.cze[p + 1] - g.cze[p - 1];
    double LD = g.Len_Disc[l];
    if (!SECOND) {
        double d[3] = {g.cx[ip] - g.cx[im], g.cy[ip] - g.cy[im], g.cz[ip] - g.cz[im]};
#pragma unroll
        for (int c = 0; c < 3; c++) if (d[c] > kPi) d[c] = d[c] - kPi2;
#pragma unroll
        for (int c = 0; c < 3; c++) if (d[c] < -kPi) d[c] = d[c] + kPi2;
#pragma unroll
        for (int c = 0; c < 3; c++) g.ck[3 * p + c] = d[c] / (ds * LD);
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) g.ckk[3 * p + c] = (g.ck[3 * ip + c] - g.ck[3 * im + c]) / (ds * LD);
    }
}

// get_spline_5th_coe (subroutine.f90:440-461)
__device__ __forceinline__ void spline5_coe(double *coe, double zr, double xl, double xr, double vl,
                                            double vr, double al, double ar) {
    double xpl = xl, xpr = xr;
    if (xl - xr > kPi) xpr = xr + kPi2;
    else if (xl - xr < -kPi) xpl = xl + kPi2;
    double d = xpr - xpl;
    coe[0] = xl;
    coe[1] = vl;
    coe[2] = al / 2.0;
    coe[3] = ((20.0 * d - (12.0 * vl + 8.0 * vr) * zr) - (3.0 * al - ar) * p2(zr)) / (2.0 * p3(zr));
    coe[4] = ((-30.0 * d + (16.0 * vl + 14.0 * vr) * zr) + (3.0 * al - 2.0 * ar) * p2(zr)) / (2.0 * p4(zr));
    coe[5] = ((12.0 * d - 6.0 * (vl + vr) * zr) - (al - ar) * p2(zr)) / (2.0 * p5(zr));
}

// Section i of a tube (0-based, i = n_sec is the periodic ghost): end points and their
// parameters with the ghost rule of tube.f90:571-597.
struct SecEnds { int64_t p0, p1; double cze0, cze1; };
__device__ __forceinline__ SecEnds section_ends(const Geo &g, int l, int64_t s) {
    int64_t b = g.pt_base[l];
    int64_t n_sec = g.pt_base[l + 1] - b;
    int64_t i = s - g.sec_base[l];
    SecEnds e;
    auto pl = [&](int64_t k, int64_t *pt, double *cz) {
        if (k < n_sec) { *pt = b + k; *cz = g.cze[b + k]; }
        else if (k == n_sec) { *pt = b; *cz = kPi2; }
        else { *pt = b + 1; *cz = kPi2 + g.cze[b + 1]; }
    };
    pl(i, &e.p0, &e.cze0);
    pl(i + 1, &e.p1, &e.cze1);
    return e;
}

// curve_spline_5th (subroutine.f90:463-488) with abscissae cze_pl*Len_Disc (tube.f90:598)
// One section per thread; the block's coefficients (18 per section, contiguous in memory) are
// staged in LDS (rows padded to 19 doubles) and written out with consecutive threads on
// consecutive doubles.
constexpr int kSplineBlock = 256;
__global__ __launch_bounds__(kSplineBlock) void k_spline(Geo g) {
    __shared__ double st[kSplineBlock * 19];
    const int64_t s0 = blockIdx.x * (int64_t)kSplineBlock;
    const int64_t s = s0 + threadIdx.x;
    if (s < g.nsec_all) {
        int l = g.sec_tube[s];
        SecEnds e = section_ends(g, l, s);
        double LD = g.Len_Disc[l];
        double zr = e.cze1 * LD - e.cze0 * LD;
        const double *X[3] = {g.cx, g.cy, g.cz};
        double coe[18];
#pragma unroll
        for (int c = 0; c < 3; c++)
            spline5_coe(coe + 6 * c, zr, X[c][e.p0], X[c][e.p1], g.ck[3 * e.p0 + c], g.ck[3 * e.p1 + c],
                        g.ckk[3 * e.p0 + c], g.ckk[3 * e.p1 + c]);
#pragma unroll
        for (int k = 0; k < 18; k++) st[19 * threadIdx.x + k] = coe[k];
        g.czl[s] = e.cze0;
        g.czr[s] = e.cze1;
    }
    __syncthreads();
    const int64_t ns = g.nsec_all - s0 < kSplineBlock ? g.nsec_all - s0 : kSplineBlock;
    double *out = g.coe + 18 * s0;
    for (int i = threadIdx.x; i < 18 * ns; i += kSplineBlock) {
        const int t = i / 18;
        out[i] = st[19 * t + (i - 18 * t)];
    }
}

// Integrand of the turning angle: |c' x c''| / |c'|^2 (tube.f90:629-639).
__device__ __forceinline__ double turn_rate(double zz, const double *coe) {
    double dc[3], ddc[3], cr[3];
    spline_d1(zz, coe, dc);
    double ndc = norm3(dc);
    spline_d2(zz, coe, ddc);
    cr[0] = dc[1] * ddc[2] - dc[2] * ddc[1];
    cr[1] = dc[2] * ddc[0] - dc[0] * ddc[2];
    cr[2] = dc[0] * ddc[1] - dc[1] * ddc[0];
    double ncr = norm3(cr);
    return ncr / p2(ndc);
}

// tube.f90:625-642 exactly: ntf_pre+1 samples summed in order, ntf = ceiling(ntf_times*theta/pi2)
// + 1, dtf = (cze_pl(t_i+1)-cze_pl(t_i))/ntf. Returns false if theta is not finite.
__device__ __forceinline__ bool theta_exact(const Geo &g, const Run &, int64_t s) {
    int l = g.sec_tube[s];
    const Run &r = g.runs[g.tube_inst[l]];
    double LD = g.Len_Disc[l];
    double coe[18];
#pragma unroll
    for (int k = 0; k < 18; k++) coe[k] = g.coe[18 * s + k];
    double czl = g.czl[s], czr = g.czr[s];
    double dtfp = (czr - czl) / (1.0 * r.ntf_pre);
    double theta = 0.0;
    for (int tj = 1; tj <= r.ntf_pre + 1; tj++) {
        double sp = czl + dtfp * (double)(tj - 1);
        double zz = (sp - czl) * LD;
        double dc[3], ddc[3], cr[3];
        spline_d1(zz, coe, dc);
        double ndc = norm3(dc);
        spline_d2(zz, coe, ddc);
        cr[0] = dc[1] * ddc[2] - dc[2] * ddc[1];
        cr[1] = dc[2] * ddc[0] - dc[0] * ddc[2];
        cr[2] = dc[0] * ddc[1] - dc[1] * ddc[0];
        double ncr = norm3(cr);
        theta = theta + ncr / p2(ndc) * dtfp * LD;
    }
    double q = ceil((double)r.ntf_times * theta / kPi2);
    bool ok = q >= 0.0 && q < 1.0e7;
    int32_t ntf = ok ? (int32_t)q + 1 : 1;
    g.ntf[s] = ntf;
    g.dtf[s] = (czr - czl) / (1.0 * ntf);
    return ok;
}

__global__ void k_theta(Geo g, Run r, int32_t *__restrict__ err) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= g.nsec_all) return;
    if (!theta_exact(g, r, s)) atomicOr(err, 1);
}

// Fast turning angle with a certified error bound. Only the integer
// ntf = ceiling(ntf_times*theta/pi2) + 1 leaves this stage (dtf is (czr-czl)/ntf), so theta itself
// need not be reproduced bit for bit: the reference's sum of ntf_pre+1 samples of
// |c' x c''| / |c'|^2 is re-evaluated in FP32 at the same sample points, from polynomial
// coefficients in t = z/zM on [0, 1] (c' x c'' has degree 7, |c'|^2 degree 8; coefficients
// formed in FP64). The samples are taken in symmetric pairs t = 1/2 +- s: each polynomial is
// shifted to p(1/2 + s) = E(s^2) + s O(s^2) in FP64 (Taylor shift) and its even and odd parts
// rounded once to FP32, so one evaluation of E and O in s^2 gives both samples. The error of every
// sample is bounded by gamma * sum (1+k)|d_k| 2^-k over the shifted coefficients d_k (|s| <= 1/2;
// this is never more than sum (1+n)|a_n| over the unshifted ones), plus the rounding of the
// coefficients, of s and of the sqrt/reciprocal. A section is decided here when the interval
// [x - dx, x + dx] around x = ntf_times*theta/pi2 contains no integer; every other section is
// listed and re-evaluated with the reference's exact FP64 sum (k_theta_listed), so ntf and dtf
// always equal the reference's.
__device__ __forceinline__ float horner8(const float *a, int n, float t) {
    float v = a[n];
    for (int k = n - 1; k >= 0; k--) v = __builtin_fmaf(v, t, a[k]);
    return v;
}

__global__ __launch_bounds__(128) void k_theta_fast(Geo g, Run, int32_t *__restrict__ redo,
                                                    unsigned int *__restrict__ nredo) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= g.nsec_all) return;
    int l = g.sec_tube[s];
    const Run &r = g.runs[g.tube_inst[l]];
    const double LD = g.Len_Disc[l];
    double coe[18];
#pragma unroll
    for (int k = 0; k < 18; k++) coe[k] = g.coe[18 * s + k];
    const double czl = g.czl[s], czr = g.czr[s];
    const int M = r.ntf_pre;
    const double dtfp = (czr - czl) / (1.0 * M);
    const double dz = dtfp * LD;
    const double zM = dz * (double)M;
    // c'(z) = sum_m b[c][m] t^m, c''(z) = sum_m e[c][m] t^m with z = zM t
    double b[3][5], e[3][4];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double zp = 1.0;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            b[c][m] = (double)(m + 1) * coe[6 * c + m + 1] * zp;
            if (m < 4) e[c][m] = (double)((m + 2) * (m + 1)) * coe[6 * c + m + 2] * zp;
            zp *= zM;
        }
    }
    // Taylor shift by 1/2 in FP64, then even/odd parts rounded to FP32
    auto shift_half = [](double *a, int n) {
        for (int i = 0; i < n; i++)
            for (int k = n - 1; k >= i; k--) a[k] += 0.5 * a[k + 1];
    };
    float wE[3][4], wO[3][4], gE[5], gO[4];
    double Wb2 = 0.0, Gb = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        double wd[8];
#pragma unroll
        for (int n = 0; n < 8; n++) {
            double v = 0.0;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                int mm = n - m;
                if (mm >= 0 && mm < 4) v += b[c1][m] * e[c2][mm] - b[c2][m] * e[c1][mm];
            }
            wd[n] = v;
        }
        shift_half(wd, 7);
        double wb = 0.0, sc2 = 1.0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float f = (float)wd[k];
            if (k & 1) wO[c][k >> 1] = f; else wE[c][k >> 1] = f;
            wb += (double)(1 + k) * fabs((double)f) * sc2;
            sc2 *= 0.5;
        }
        Wb2 += wb * wb;
    }
    {
        double gd[9];
#pragma unroll
        for (int n = 0; n < 9; n++) {
            double v = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    int mm = n - m;
                    if (mm >= 0 && mm < 5) v += b[c][m] * b[c][mm];
                }
            gd[n] = v;
        }
        shift_half(gd, 8);
        double sc2 = 1.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const float f = (float)gd[k];
            if (k & 1) gO[k >> 1] = f; else gE[k >> 1] = f;
            Gb += (double)(1 + k) * fabs((double)f) * sc2;
            sc2 *= 0.5;
        }
    }
    double acc = 0.0;           // sum of f
    float gmin = 3.0e38f;       // smallest denominator
    auto fval = [](float wx, float wy, float wz, float gg) {
        const float n2 = __builtin_fmaf(wx, wx, __builtin_fmaf(wy, wy, wz * wz));
        return __builtin_amdgcn_sqrtf(n2) * __builtin_amdgcn_rcpf(gg);
    };
    // samples j = 0..M at t = j/M: pairs t = 1/2 +- s with s = (2i + o)/(2M), plus t = 1/2 for even M;
    // a pair's two values are added in FP32 before the FP64 sum, and the smaller of its two
    // denominators, fl(gE + gO) and fl(gE - gO), is fl(gE - |gO|) (rounding is monotonic)
    const bool odd = (M & 1) != 0;
    const int npair = odd ? (M + 1) / 2 : M / 2, o = odd ? 1 : 2;
    const float h = 1.0f / (float)(2 * M);
    if (!odd) {
        acc += (double)fval(wE[0][0], wE[1][0], wE[2][0], gE[0]);
        gmin = gE[0];
    }
#pragma unroll 2
    for (int i = 0; i < npair; i++) {
        const float sv = (float)(2 * i + o) * h, v = sv * sv;
        float wp[3], wm[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float E = horner8(wE[c], 3, v), O = horner8(wO[c], 3, v);
            wp[c] = __builtin_fmaf(sv, O, E);
            wm[c] = __builtin_fmaf(-sv, O, E);
        }
        const float gEv = horner8(gE, 4, v), gOv = horner8(gO, 3, v);
        const float gp = __builtin_fmaf(sv, gOv, gEv), gm = __builtin_fmaf(-sv, gOv, gEv);
        const float fp = fval(wp[0], wp[1], wp[2], gp), fm = fval(wm[0], wm[1], wm[2], gm);
        acc += (double)(fp + fm);
        gmin = fminf(gmin, fminf(gp, gm));
    }
    // error bound of the FP32 sum against the reference's FP64 sum (both times dz): per sample the
    // numerator's error kw Wb over the denominator and the denominator's relative error kg Gb / g,
    // summed with sum 1/g <= (M + 1) / gmin and sum f/g <= acc / gmin; the roundings of sqrt, the
    // reciprocal and the product (6u) and of the pair sums (u)
    const double u = 5.9604644775390625e-08;             // 2^-24
    const double zrel = 4.0e-15 * LD / fmax(zM, 1e-300);  // the reference's sample points
    const double kw = 16.0 * u + zrel, kg = 20.0 * u + zrel;
    const double Wb = sqrt(Wb2);
    const double ig = 1.0 / (double)gmin;
    const double sum_b = 1.1 * (kw * Wb * (double)(M + 1) * ig + 7.0 * u * acc + kg * Gb * acc * ig) * 1.01;
    const double theta = acc * dz;
    const double err = sum_b * dz + 1.0e-9 * theta;
    const double sc = (double)r.ntf_times / kPi2;
    const double x = sc * theta;
    const double dx = sc * err + 1.0e-12 * x + 1.0e-12;
    const double lo = floor(x - dx), hi = floor(x + dx);
    bool ok = isfinite(x) && isfinite(dx) && M > 0 && zM > 0.0 && (double)gmin > 1000.0 * u * Gb &&
              x - dx > 0.0 && lo == hi && x + dx < 1.0e7;
    if (ok) {
        int32_t ntf = (int32_t)lo + 2;  // ceiling(x_ref) + 1 for x_ref in (lo, lo + 1)
        g.ntf[s] = ntf;
        g.dtf[s] = (czr - czl) / (1.0 * ntf);
    } else {
        unsigned int k = atomicAdd(nredo, 1u);
        redo[k] = (int32_t)s;
    }
}

__global__ void k_theta_listed(Geo g, Run r, const int32_t *__restrict__ redo,
                               const unsigned int *__restrict__ nredo, int32_t *__restrict__ err) {
    unsigned int n = *nredo;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (!theta_exact(g, r, redo[i])) atomicOr(err, 1);
}

// The sub-point q = sp_base[s] + j of section s of tube l.
__device__ __forceinline__ void subpoint_at(const Geo &g, int64_t q, int32_t s, int l, int32_t j) {
    bool real = (int64_t)s < g.sec_base[l + 1] - 1;
    int32_t src = s;
    int32_t jj = j;
    if (real && j == g.ntf[s]) { src = s + 1; jj = 0; }
    double LD = g.Len_Disc[l];
    double czl = g.czl[src];
    double sv = czl + g.dtf[src] * (double)jj;
    double coe[18];
#pragma unroll
    for (int k = 0; k < 18; k++) coe[k] = g.coe[18 * (int64_t)src + k];
    double c3[3];
    spline_pos((sv - czl) * LD, coe, c3);
#pragma unroll
    for (int c = 0; c < 3; c++) g.ci[3 * q + c] = c3[c];
}

// tube.f90:706-713 sub-points; tube.f90:719-723 seam: the last sub-point of a real section is
// the first sub-point of the next section (evaluated with that section's spline at 0).
// Each wave takes 64 consecutive sections (lane i: section s0 + i) and then their sub-points, 64
// at a time with consecutive lanes on consecutive sub-points: a sub-point's section is the last
// lane whose first sub-point is not after it (a binary search over the lanes), so sub_sec,
// sub_tube and ci are written in order.
constexpr int kSubBlock = 256;
__global__ __launch_bounds__(kSubBlock) void k_subpoints(Geo g) {
    const int lane = threadIdx.x & 63;
    const int64_t s0 = (blockIdx.x * (int64_t)(kSubBlock / 64) + (threadIdx.x >> 6)) * 64;
    if (s0 >= g.nsec_all) return;   // whole wave
    const int64_t qa = g.sp_base[s0];
    const int64_t se = s0 + 64 < g.nsec_all ? s0 + 64 : g.nsec_all;
    const int64_t nq = g.sp_base[se] - qa;
    const int32_t rel = s0 + lane < se ? (int32_t)(g.sp_base[s0 + lane] - qa) : INT32_MAX;
    for (int64_t k0 = 0; k0 < nq; k0 += 64) {
        const int32_t qr = (int32_t)(k0 + lane);
        int32_t pos = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const int32_t v = __shfl(rel, pos + step, 64);
            if (pos + step < 64 && v <= qr) pos += step;
        }
        const int32_t first = __shfl(rel, pos, 64);   // every lane takes part in the shuffles
        if (qr >= nq) continue;
        const int64_t q = qa + qr;
        const int32_t s = (int32_t)(s0 + pos);
        const int32_t j = qr - first;
        const int l = g.sec_tube[s];
        g.sub_sec[q] = s;
        g.sub_tube[q] = l;
        subpoint_at(g, q, s, l, j);
    }
}

// Per sub-point: number of grid columns of its sub-segment's box (0 if it starts none).
// Columns run along the chord's dominant axis. Also counts the work items (column blocks).
__global__ void k_get_bins(Geo g, Run r, int32_t *__restrict__ lohi, double *__restrict__ R) {
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= g.nsub_all) return;
    SubSeg sg;
    if (subseg_load(g, r, q, sg)) {
#pragma unroll
        for (int c = 0; c < 3; c++) { lohi[6 * q + c] = sg.lo[c]; lohi[6 * q + 3 + c] = sg.hi[c]; }
        R[q] = sg.R;
    } else {
#pragma unroll
        for (int c = 0; c < 6; c++) lohi[6 * q + c] = 0;
        R[q] = -1.0;
    }
}

static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
static hipError_t dalloc(wt_ctx *c, std::vector<void *> &list, T **p, size_t count) {
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) return e;
    list.push_back(v);
    *p = (T *)v;
    return hipSuccess;
}

static void dfree(std::vector<void *> &list) {
    for (void *p : list) (void)hipFree(p);
    list.clear();
}

// The kernels of the geometry stage on buffers already laid out (the stage's own or, in a
// replay, those of the last eager run): no allocation, copy to the host or synchronisation.
static hipError_t segment_kernels(wt_ctx *c, bool replay) {
    Geo &g = c->geo;
    SegWs &w = c->segws;
    hipStream_t st = c->stream;
    const int bs = 256;
    // per point
    WT_HIP(fill_zero(g.cze, sizeof(double) * g.npoint_all, st));
    hipLaunchKernelGGL(k_fill_owner, dim3(g.nline), dim3(256), 0, st, g.pt_base, g.nline, g.pt_tube);
    hipLaunchKernelGGL(k_fill_owner, dim3(g.nline), dim3(256), 0, st, g.sec_base, g.nline, g.sec_tube);
    hipLaunchKernelGGL(k_chords, dim3(grid_for(g.npoint_all, bs)), dim3(bs), 0, st, g);
    hipLaunchKernelGGL(k_cumsum, dim3(grid_for(g.nline, 64)), dim3(64), 0, st, g, w.tot);
    hipLaunchKernelGGL(k_cze_norm, dim3(grid_for(g.npoint_all, bs)), dim3(bs), 0, st, g, w.tot);
    hipLaunchKernelGGL(k_derivative<false>, dim3(grid_for(g.npoint_all, bs)), dim3(bs), 0, st, g);
    hipLaunchKernelGGL(k_derivative<true>, dim3(grid_for(g.npoint_all, bs)), dim3(bs), 0, st, g);
    // per section
    hipLaunchKernelGGL(k_spline, dim3(grid_for(g.nsec_all, kSplineBlock)), dim3(kSplineBlock), 0, st, g);
    WT_HIP(fill_zero(w.err, sizeof(int32_t), st));
    kt_begin(c, kSlotTheta);
    if (c->theta_exact_all) {
        hipLaunchKernelGGL(k_theta, dim3(grid_for(g.nsec_all, 128)), dim3(128), 0, st, g, c->run, w.err);
    } else {
        WT_HIP(fill_zero(w.nredo, sizeof(unsigned int), st));
        hipLaunchKernelGGL(k_theta_fast, dim3(grid_for(g.nsec_all, 128)), dim3(128), 0, st, g, c->run, w.redo, w.nredo);
        hipLaunchKernelGGL(k_theta_listed, dim3(1024), dim3(64), 0, st, g, c->run, (const int32_t *)w.redo,
                           (const unsigned int *)w.nredo, w.err);
    }
    kt_end(c, kSlotTheta);
    // sub-point layout
    WT_HIP(exclusive_scan_i32_to_i64(g.ntf, 1, g.sp_base, g.nsec_all, st, w.scan));
    if (!replay) return hipGetLastError();
    kt_begin(c, kSlotSubpoints);
    hipLaunchKernelGGL(k_subpoints, dim3(grid_for(g.nsec_all, kSubBlock)), dim3(kSubBlock), 0, st, g);
    kt_end(c, kSlotSubpoints);
    return hipGetLastError();
}

hipError_t launch_segments(wt_ctx *c) {
    Geo &g = c->geo;
    SegWs &w = c->segws;
    hipStream_t st = c->stream;
    dfree(c->seg_allocs);
    w = SegWs{};
    WT_HIP(dalloc(c, c->seg_allocs, &w.tot, g.nline));
    WT_HIP(dalloc(c, c->seg_allocs, &w.err, 1));
    WT_HIP(dalloc(c, c->seg_allocs, &w.redo, g.nsec_all));
    WT_HIP(dalloc(c, c->seg_allocs, &w.nredo, 1));
    WT_HIP(dalloc(c, c->seg_allocs, &w.scan, scan_scratch_elems(g.nsec_all)));
    WT_HIP(segment_kernels(c, false));
    if (!c->theta_exact_all)
        WT_HIP(hipMemcpyAsync(&c->theta_redo, w.nredo, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    int32_t herr = 0;
    WT_HIP(hipMemcpyAsync(&herr, w.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    int64_t nsub = 0;
    WT_HIP(hipMemcpyAsync(&nsub, g.sp_base + g.nsec_all, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    WT_HIP(hipStreamSynchronize(st));
    if (herr) return hipErrorInvalidValue;
    // sub-point indices are 32-bit in the brick pair lists
    if (nsub >= (int64_t)UINT32_MAX) return hipErrorInvalidImage;
    g.nsub_all = nsub;
    WT_HIP(dalloc(c, c->seg_allocs, &g.sub_sec, nsub));
    WT_HIP(dalloc(c, c->seg_allocs, &g.sub_tube, nsub));
    WT_HIP(dalloc(c, c->seg_allocs, &g.ci, 3 * nsub));
    kt_begin(c, kSlotSubpoints);
    hipLaunchKernelGGL(k_subpoints, dim3(grid_for(g.nsec_all, kSubBlock)), dim3(kSubBlock), 0, st, g);
    kt_end(c, kSlotSubpoints);
    return hipGetLastError();
}

// The geometry stage again on the tubes and the layout of the last launch_segments (every size
// is the same for the same tubes): kernels only, for recording into a graph.
hipError_t replay_segments(wt_ctx *c) { return segment_kernels(c, true); }

hipError_t launch_get_bins(wt_ctx *c, int32_t *d_lohi, double *d_R) {
    hipLaunchKernelGGL(k_get_bins, dim3(grid_for(c->geo.nsub_all, 256)), dim3(256), 0, c->stream,
                       c->geo, c->run, d_lohi, d_R);
    return hipGetLastError();
}

}  // namespace wt

namespace wt {

// Counters of wt_seg_stats: max ntf, min dtf (all sections), sum of ntf (all sections) and
// number of sub-segments (real sections only). Each block writes one partial row; one block
// combines the rows (no contended atomics).
struct SegAcc {
    unsigned long long mx, sum, nseg, dmin;
};

__device__ __forceinline__ void seg_acc_combine(SegAcc &a, const SegAcc &b) {
    a.mx = b.mx > a.mx ? b.mx : a.mx;
    a.sum += b.sum;
    a.nseg += b.nseg;
    a.dmin = b.dmin < a.dmin ? b.dmin : a.dmin;
}

__device__ __forceinline__ SegAcc seg_acc_block(SegAcc v) {
    __shared__ SegAcc part[4];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        SegAcc w{__shfl_xor(v.mx, o, 64), __shfl_xor(v.sum, o, 64), __shfl_xor(v.nseg, o, 64),
                 __shfl_xor(v.dmin, o, 64)};
        seg_acc_combine(v, w);
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    SegAcc r = part[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) seg_acc_combine(r, part[w]);
    return r;
}

__global__ __launch_bounds__(256) void k_seg_stats(Geo g, SegAcc *__restrict__ rows) {
    SegAcc v{0, 0, 0, ~0ull};
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < g.nsec_all;
         s += (int64_t)gridDim.x * blockDim.x) {
        int l = g.sec_tube[s];
        unsigned long long n = (unsigned long long)g.ntf[s];
        SegAcc w{n, n, (int64_t)s < g.sec_base[l + 1] - 1 ? n : 0ull,
                 (unsigned long long)__double_as_longlong(g.dtf[s])};  // dtf > 0: bit order = value order
        seg_acc_combine(v, w);
    }
    SegAcc r = seg_acc_block(v);
    if (threadIdx.x == 0) rows[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void k_seg_stats_rows(const SegAcc *__restrict__ rows, int nrows,
                                                        SegAcc *__restrict__ out) {
    SegAcc v{0, 0, 0, ~0ull};
    for (int i = threadIdx.x; i < nrows; i += blockDim.x) seg_acc_combine(v, rows[i]);
    SegAcc r = seg_acc_block(v);
    if (threadIdx.x == 0) *out = r;
}

hipError_t launch_seg_stats(wt_ctx *c, wt_seg_stats *o) {
    const int blocks = 1024;
    SegAcc *d = nullptr, h{};
    WT_HIP(hipMallocAsync((void **)&d, sizeof(SegAcc) * (blocks + 1), c->stream));
    hipLaunchKernelGGL(k_seg_stats, dim3(blocks), dim3(256), 0, c->stream, c->geo, d + 1);
    hipLaunchKernelGGL(k_seg_stats_rows, dim3(1), dim3(256), 0, c->stream, (const SegAcc *)(d + 1), blocks, d);
    WT_HIP(hipGetLastError());
    WT_HIP(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c->stream));
    WT_HIP(hipFreeAsync(d, c->stream));
    WT_HIP(hipStreamSynchronize(c->stream));
    o->ntf_max = (int32_t)h.mx;
    o->ntf_sum = (int64_t)h.sum;
    o->nseg_all = (int64_t)h.nseg;
    union { unsigned long long u; double d; } bits;
    bits.u = h.dmin;
    o->dtf_min = bits.d;
    return hipSuccess;
}

}  // namespace wt
