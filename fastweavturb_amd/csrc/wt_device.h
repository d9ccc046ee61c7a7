// Device helpers shared by the geometry, binning and deposition kernels.
// The FP64 helpers keep the reference's operation order so that, compiled with
// -ffp-contract=off, they round exactly like the reference's statements.
#pragma once

#include "wt_internal.h"

namespace wt {

__device__ __forceinline__ double p2(double x) { return x * x; }
__device__ __forceinline__ double p3(double x) { return (x * x) * x; }
__device__ __forceinline__ double p4(double x) { double y = x * x; return y * y; }
__device__ __forceinline__ double p5(double x) { double y = x * x; return (y * y) * x; }

// norm_zeta (subroutine.f90:793-797)
__device__ __forceinline__ double norm3(const double *v) {
    return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}
// cminus (subroutine.f90:695-712): minimum-image r - l
__device__ __forceinline__ void cminus(const double *r, const double *l, double *d) {
#pragma unroll
    for (int c = 0; c < 3; c++) d[c] = r[c] - l[c];
#pragma unroll
    for (int c = 0; c < 3; c++) if (d[c] > kPi) d[c] = d[c] - kPi2;
#pragma unroll
    for (int c = 0; c < 3; c++) if (d[c] < -kPi) d[c] = d[c] + kPi2;
}

// sigma_func (subroutine.f90:164-173); omega is the integer floor(length/pi + 0.5) * n_T
// (for smax == smin the reference's expression is smin + 0 * finite = smin exactly)
__device__ __forceinline__ double sigma_func(double zeta, double length, double smax, double smin,
                                             int n_T) {
    if (smax == smin) return smin;
    double om = (double)((long long)floor(length / kPi + 0.5) * (long long)n_T);
    return smin + (smax - smin) * (sin(om * zeta) + 1.0) / 2.0;
}

// coo_curve_spline_5th (subroutine.f90:490-516), coe[c*6 + j]
__device__ __forceinline__ void spline_pos(double z, const double *coe, double *c3) {
    double z2 = p2(z), z3 = p3(z), z4 = p4(z), z5 = p5(z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double *a = coe + 6 * c;
        c3[c] = ((((a[5] * z5 + a[4] * z4) + a[3] * z3) + a[2] * z2) + a[1] * z) + a[0];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) if (c3[c] >= kPi) c3[c] = c3[c] - kPi2;
#pragma unroll
    for (int c = 0; c < 3; c++) if (c3[c] < -kPi) c3[c] = c3[c] + kPi2;
}
// d1_coo_curve_spline_5th (subroutine.f90:518-537)
__device__ __forceinline__ void spline_d1(double z, const double *coe, double *d) {
    double z2 = p2(z), z3 = p3(z), z4 = p4(z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double *a = coe + 6 * c;
        d[c] = ((((5.0 * a[5]) * z4 + (4.0 * a[4]) * z3) + (3.0 * a[3]) * z2) + (2.0 * a[2]) * z) + a[1];
    }
}
// d2_coo_curve_spline_5th (subroutine.f90:539-558)
__device__ __forceinline__ void spline_d2(double z, const double *coe, double *d) {
    double z2 = p2(z), z3 = p3(z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double *a = coe + 6 * c;
        d[c] = (((20.0 * a[5]) * z3 + (12.0 * a[4]) * z2) + (6.0 * a[3]) * z) + 2.0 * a[2];
    }
}

__device__ __forceinline__ int find_tube(const int64_t *base, int nline, int64_t i) {
    int lo = 0, hi = nline - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// One sub-segment of the deposition loop (tube.f90:766-798), reconstructed from the
// sub-points. Sub-point q starts the sub-segment (section sec, local index j).
struct SubSeg {
    int32_t sec, tube, j;
    double ci0[3], ci1[3];   // after the mark1 unwrap
    double tci0[3], tci1[3];
    double ndci0, s0, s1, R;
    int32_t lo[3], hi[3];
    // grid of the tube's instance
    int32_t N;
    double dx, start;
    int64_t field_off;
};

// Chord unit vector and length between sub-points a and b (tube.f90:740-750).
__device__ __forceinline__ void chord(const double *ci, int64_t b, int64_t a, double *t, double *nd) {
    double d[3];
    cminus(ci + 3 * b, ci + 3 * a, d);
    double n = norm3(d);
    *nd = n;
#pragma unroll
    for (int c = 0; c < 3; c++) t[c] = d[c] / n;
}

// Returns false when q starts no deposited sub-segment (ghost section, last sub-point of a
// section, or a tube outside line_s..line_e).
__device__ __forceinline__ bool subseg_load(const Geo &g, const Run &, int64_t q, SubSeg &o) {
    int32_t s = g.sub_sec[q];
    int32_t l = g.sec_tube[s];
    int32_t j = (int32_t)(q - g.sp_base[s]);
    if ((int64_t)s == g.sec_base[l + 1] - 1) return false;
    if (j >= g.ntf[s]) return false;
    const Run &r = g.runs[g.tube_inst[l]];
    if (l - r.tube0 < r.line_s || l - r.tube0 > r.line_e) return false;
    o.N = r.N;
    o.dx = r.dx;
    o.start = r.start;
    o.field_off = r.field_off;
    o.sec = s;
    o.tube = l;
    o.j = j;
    const double *ci = g.ci;
    double nd1;
    chord(ci, q + 1, q, o.tci0, &o.ndci0);
    if (j + 1 < g.ntf[s]) chord(ci, q + 2, q + 1, o.tci1, &nd1);
    else {
        int64_t f = g.sp_base[s + 1];
        chord(ci, f + 1, f, o.tci1, &nd1);
    }
    double czl = g.czl[s], dtf = g.dtf[s];
    o.s0 = czl + dtf * (double)j;
    o.s1 = czl + dtf * (double)(j + 1);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o.ci0[c] = ci[3 * q + c];
        o.ci1[c] = ci[3 * (q + 1) + c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {  // mark1 (tube.f90:779-785)
        if (o.ci1[c] - o.ci0[c] > kPi) o.ci0[c] = o.ci0[c] + kPi2;
        else if (o.ci1[c] - o.ci0[c] < -kPi) o.ci1[c] = o.ci1[c] + kPi2;
    }
    double smin = g.sigma[l], smax = r.sigma_ratio * smin, len = g.length[l];
    int nT = g.n_T[l];
    double sig = sigma_func(o.s0, len, smax, smin, nT);
    double R = r.Rtube_ratio * sig;
    sig = sigma_func(o.s1, len, smax, smin, nT);
    double R2 = r.Rtube_ratio * sig;
    o.R = (R2 > R) ? R2 : R;
#pragma unroll
    for (int c = 0; c < 3; c++) {  // tube.f90:793-798
        double mn = o.ci0[c] < o.ci1[c] ? o.ci0[c] : o.ci1[c];
        double mx = o.ci0[c] > o.ci1[c] ? o.ci0[c] : o.ci1[c];
        o.lo[c] = (int32_t)floor((mn - r.start - o.R) / r.dx + 1.0);
        o.hi[c] = (int32_t)ceil((mx - r.start + o.R) / r.dx + 1.0);
    }
    return true;
}

// subseg_load with a shorter chain of dependent loads (same result): the sub-point's section and
// tube come from sub_sec and sub_tube together, and the four sub-points q..q+3 that its chords can
// use are read at once (a section's last sub-segment takes its second chord from the next
// section's first two sub-points, q+2 and q+3 there; indices past the end are clamped, their
// values unused).
__device__ __forceinline__ bool subseg_load_fast(const Geo &g, int64_t q, SubSeg &o) {
    const int32_t s = g.sub_sec[q];
    const int32_t l = g.sub_tube[q];
    double p[12];
    const int64_t last = g.nsub_all - 1;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int64_t qk = q + k <= last ? q + k : last;
#pragma unroll
        for (int c = 0; c < 3; c++) p[3 * k + c] = g.ci[3 * qk + c];
    }
    const int32_t j = (int32_t)(q - g.sp_base[s]);
    const int32_t nt = g.ntf[s];
    if ((int64_t)s == g.sec_base[l + 1] - 1) return false;
    if (j >= nt) return false;
    const Run &r = g.runs[g.tube_inst[l]];
    if (l - r.tube0 < r.line_s || l - r.tube0 > r.line_e) return false;
    o.N = r.N;
    o.dx = r.dx;
    o.start = r.start;
    o.field_off = r.field_off;
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
    double czl = g.czl[s], dtf = g.dtf[s];
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
    double smin = g.sigma[l], smax = r.sigma_ratio * smin, len = g.length[l];
    int nT = g.n_T[l];
    double sig = sigma_func(o.s0, len, smax, smin, nT);
    double R = r.Rtube_ratio * sig;
    sig = sigma_func(o.s1, len, smax, smin, nT);
    double R2 = r.Rtube_ratio * sig;
    o.R = (R2 > R) ? R2 : R;
#pragma unroll
    for (int c = 0; c < 3; c++) {  // tube.f90:793-798
        double mn = o.ci0[c] < o.ci1[c] ? o.ci0[c] : o.ci1[c];
        double mx = o.ci0[c] > o.ci1[c] ? o.ci0[c] : o.ci1[c];
        o.lo[c] = (int32_t)floor((mn - r.start - o.R) / r.dx + 1.0);
        o.hi[c] = (int32_t)ceil((mx - r.start + o.R) / r.dx + 1.0);
    }
    return true;
}

// Mesh coordinate of an extended 1-based index (tube.f90:331-339).
__device__ __forceinline__ double mesh(int32_t ip, const SubSeg &sg) {
    return ((double)ip - 1.0) * sg.dx + sg.start;
}

// Fortran MOD wrap of an extended index (tube.f90:803, 870-871); <= 0 means "outside".
__device__ __forceinline__ int32_t wrap_idx(int32_t ip, int32_t N) { return (ip + N - 1) % N + 1; }

// The reference's candidate test for one extended cell (tube.f90:803-818).
__device__ __forceinline__ bool pair_hits(const SubSeg &sg, int32_t ip, int32_t jp, int32_t kp) {
    int32_t k = wrap_idx(kp, sg.N);
    if (!(k > 0 && k <= sg.N)) return false;
    double X0 = mesh(ip, sg), X1 = mesh(jp, sg), X2 = mesh(kp, sg);
    double a0 = X0 - sg.ci0[0], a1 = X1 - sg.ci0[1], a2 = X2 - sg.ci0[2];
    double t0 = (a0 * sg.tci0[0] + a1 * sg.tci0[1]) + a2 * sg.tci0[2];
    if (!(t0 >= 0.0)) return false;
    double t1 = ((X0 - sg.ci1[0]) * sg.tci1[0] + (X1 - sg.ci1[1]) * sg.tci1[1]) +
                (X2 - sg.ci1[2]) * sg.tci1[2];
    if (!(t1 < 0.0)) return false;
    double rho = sqrt(((p2(a0) + p2(a1)) + p2(a2)) - p2(t0));
    return rho < sg.R;
}

__device__ __forceinline__ int dominant_axis(const double *t) {
    double a0 = fabs(t[0]), a1 = fabs(t[1]), a2 = fabs(t[2]);
    if (a0 >= a1 && a0 >= a2) return 0;
    return (a1 >= a2) ? 1 : 2;
}

}  // namespace wt
