// Spatial binning of the deposition work (tube.f90:793-798 boxes).
//
// The grid of each instance is cut into bricks of Run::bd[0] x bd[1] x bd[2] cells (partial
// bricks at the upper edge when a brick edge does not divide N). A sub-segment is paired with every brick that its
// integer bounding box (in extended indices, wrapped periodically) overlaps, once per periodic
// image of the brick that the box meets, unless that part of the box lies entirely outside the
// slab between the sub-segment's two end planes or beyond the reach of its hits (pair_meets_slab). The deposition
// kernel then owns one brick at a time and accumulates all of its pairs in LDS, so no global
// atomics touch the field.
//
// Pairs are built in two passes over the sub-segments: count per brick, exclusive scan, fill.
// Both passes combine the per-brick atomics of a wavefront (consecutive sub-segments of a tube
// mostly share their bricks), so a brick counter sees about one atomic per wave.
// Compiled with -ffp-contract=off: the boxes are the reference's FP64 expressions.
#include <algorithm>

#include "wt_device.h"

namespace wt {

__device__ __forceinline__ uint64_t lanemask_lt64() {
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Adds 1 to cnt[key] for every lane with `active`; one atomic per distinct key in the wave, none
// of them waited for.
__device__ __forceinline__ void wave_count_add(uint32_t *cnt, int64_t key, bool active) {
    uint64_t pending = __ballot(active);
    const int lane = (int)__lane_id();
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int64_t lk = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(key >> 32), leader) << 32) |
                                     (uint32_t)__builtin_amdgcn_readlane((int)key, leader));
        const uint64_t m = __ballot(active && key == lk) & pending;
        if (lane == leader) atomicAdd(cnt + lk, (uint32_t)__popcll(m));
        pending &= ~m;
    }
}

// Adds 1 to cnt[key] for every lane with `active` and returns the counter value before this lane's
// increment (its slot). The wave's lanes are grouped by key first; then every group's leader adds
// the group's size at once, so the wave waits for one round of returned atomics, not one per key.
__device__ __forceinline__ uint32_t wave_slot_add(uint32_t *cnt, int64_t key, bool active) {
    uint64_t pending = __ballot(active);
    const int lane = (int)__lane_id();
    int my_leader = lane;
    uint64_t my_m = 0;
    uint32_t lead_n = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int64_t lk = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(key >> 32), leader) << 32) |
                                     (uint32_t)__builtin_amdgcn_readlane((int)key, leader));
        const uint64_t m = __ballot(active && key == lk) & pending;
        if ((m >> lane) & 1ull) {
            my_leader = leader;
            my_m = m;
        }
        if (lane == leader) lead_n = (uint32_t)__popcll(m);
        pending &= ~m;
    }
    uint32_t base = 0;
    if (lead_n) base = atomicAdd(cnt + key, lead_n);
    base = (uint32_t)__shfl((int)base, my_leader, 64);
    return base + (uint32_t)__popcll(my_m & lanemask_lt64());
}

// Bricks of one axis met by the extended index range [lo, hi] (lo >= 1 - N, hi <= 2N): for each
// periodic image k in {-1, 0, 1}, the part of [lo, hi] in [1 + kN, N + kN] covers the wrapped
// bricks f[k] .. l[k] (empty when f > l). A brick met through two images gives two pairs.
struct AxisImages {
    int32_t f[3], l[3];
    __device__ int32_t count() const {
        int32_t n = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) n += l[k] >= f[k] ? l[k] - f[k] + 1 : 0;
        return n;
    }
    // e-th (brick, image) of the axis; image returned as 0..2 for k = -1..1
    __device__ void at(int32_t e, int32_t &brick, int32_t &img) const {
        img = 0;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            int32_t n = l[k] >= f[k] ? l[k] - f[k] + 1 : 0;
            if (img == k && e >= n) { e -= n; img = k + 1; }
        }
        brick = f[img] + e;
    }
};

__device__ __forceinline__ AxisImages axis_images(int32_t lo, int32_t hi, int32_t N, int32_t lbd) {
    AxisImages r;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int32_t s = (k - 1) * N;
        const int32_t a = lo > 1 + s ? lo : 1 + s, b = hi < N + s ? hi : N + s;
        r.f[k] = 0;
        r.l[k] = -1;
        if (a <= b) {
            r.f[k] = (a - s - 1) >> lbd;
            r.l[k] = (b - s - 1) >> lbd;
        }
    }
    return r;
}

// Steps through the (brick, image) entries of one axis in the order of AxisImages::at.
struct AxisWalk {
    int32_t bi, im;
    __device__ void start(const AxisImages &a) {
        im = a.l[0] >= a.f[0] ? 0 : (a.l[1] >= a.f[1] ? 1 : 2);
        bi = a.f[im];
    }
    // past the axis's last entry it leaves bi beyond l[2]; callers restart the axis instead
    __device__ void next(const AxisImages &a) {
        if (++bi <= a.l[im] || im == 2) return;
        im = (im == 0 && a.l[1] >= a.f[1]) ? 1 : 2;
        bi = a.f[im];
    }
};

// Whether the part of the sub-segment's box inside brick b (axis bricks bi[c], images im[c]) can
// hold a cell with t0 >= 0 and t1 < 0 (tube.f90:806-813). t0 and t1 are linear in the cell, so
// over that part their extremes sit at corners picked by the signs of tci0 and tci1; the bricks
// where even the largest t0 is negative or the smallest t1 is not, beyond 1e-12 (well above the
// rounding of either in the reference's expression for any cell), receive no contribution from
// the sub-segment and get no pair. Both extremes are sums of one term per axis (slab_terms),
// added in the order x, y, z.
// A hit X also lies within reach of ci0: with rho < R its distance from the chord line and t0 its
// offset along tci0, t1 = (t0 - l) cos(delta) + rho sin(delta) (e . f) < 0 (delta: the angle
// between tci0 and tci1, l = ndci0) gives 0 <= t0 < l + R tan(delta), so
// |X - ci0|^2 = rho^2 + t0^2 < R^2 + (l + R tan(delta))^2. Brick parts whose nearest point is
// farther get no pair either (the squared distance is a sum of one term per axis, the bound
// widened by 1e-9 relative and 1e-12 absolute; no bound when the chords turn by 60 degrees or more).
__device__ __forceinline__ void slab_terms(const SubSeg &sg, int c, int32_t bi, int32_t im, int32_t bd, double &h0,
                                           double &l1, double &d2) {
    const int32_t org = bi * bd, ext = sg.N - org < bd ? sg.N - org : bd;
    const int32_t sh = (im - 1) * sg.N;
    const int32_t a = sg.lo[c] > org + 1 + sh ? sg.lo[c] : org + 1 + sh;
    const int32_t b = sg.hi[c] < org + ext + sh ? sg.hi[c] : org + ext + sh;
    const double xa = mesh(a, sg), xb = mesh(b, sg);
    h0 = ((sg.tci0[c] > 0.0 ? xb : xa) - sg.ci0[c]) * sg.tci0[c];
    l1 = ((sg.tci1[c] > 0.0 ? xa : xb) - sg.ci1[c]) * sg.tci1[c];
    const double d = fmax(fmax(xa - sg.ci0[c], sg.ci0[c] - xb), 0.0);
    d2 = d * d;
}

__device__ __forceinline__ bool slab_met(double t0, double t1) { return t0 >= -1e-12 && t1 < 1e-12; }

// Squared reach of the sub-segment's hits from ci0 (above).
__device__ __forceinline__ double hit_reach2(const SubSeg &sg) {
    const double cd = (sg.tci0[0] * sg.tci1[0] + sg.tci0[1] * sg.tci1[1]) + sg.tci0[2] * sg.tci1[2];
    if (!(cd > 0.5)) return INFINITY;
    const double tand = sqrt(fmax(1.0 - cd * cd, 0.0)) / cd;
    const double e = sg.ndci0 + sg.R * tand;
    return (sg.R * sg.R + e * e) * (1.0 + 1e-9) + 1e-12;
}

__device__ __forceinline__ bool pair_meets_slab(const SubSeg &sg, const int32_t *bi, const int32_t *im,
                                                const int32_t *bd, double reach2) {
    double h0[3], l1[3], d2[3];
#pragma unroll
    for (int c = 0; c < 3; c++) slab_terms(sg, c, bi[c], im[c], bd[c], h0[c], l1[c], d2[c]);
    return slab_met((h0[0] + h0[1]) + h0[2], (l1[0] + l1[1]) + l1[2]) && (d2[0] + d2[1]) + d2[2] < reach2;
}

// A sub-segment with more pairs than this (a box of many bricks: the widest cores at high N) is
// expanded by a whole wave in k_pairs_big instead of one lane, so that no lane runs thousands of
// steps while the rest of its wave waits.
constexpr int32_t kBigPairs = 256;

// The walk over one sub-segment's (brick, image) entries, x fastest as AxisImages::at orders
// them. With SLAB, the y and z slab terms are kept and change only when the walk leaves a row.
template <bool SLAB>
struct EntryWalk {
    AxisWalk wk[3];
    double h0y = 0.0, l1y = 0.0, h0z = 0.0, l1z = 0.0, d2y = 0.0, d2z = 0.0, reach2 = 0.0;
    int32_t jx = 0, jy = 0;
    int64_t rowkey = 0;
    __device__ void start(const AxisImages *ab, const SubSeg &sg, const int32_t *bd, int64_t brick0, int32_t nbx,
                          int32_t nby) {
#pragma unroll
        for (int c = 0; c < 3; c++) wk[c].start(ab[c]);
        if (SLAB) {
            reach2 = hit_reach2(sg);
            slab_terms(sg, 1, wk[1].bi, wk[1].im, bd[1], h0y, l1y, d2y);
            slab_terms(sg, 2, wk[2].bi, wk[2].im, bd[2], h0z, l1z, d2z);
        }
        rowkey = brick0 + ((int64_t)wk[2].bi * nby + wk[1].bi) * nbx;
    }
    // the current entry's brick key, image code and (SLAB) slab test; then steps to the next entry
    // (t: the current entry's index, n: the number of entries)
    __device__ bool take(const AxisImages *ab, const SubSeg &sg, const int32_t *bd, int64_t brick0, int32_t nbx,
                         int32_t nby, int32_t nx, int32_t ny, int32_t t, int32_t n, int64_t &key, int32_t &code) {
        bool met = true;
        if (SLAB) {
            double h0x, l1x, d2x;
            slab_terms(sg, 0, wk[0].bi, wk[0].im, bd[0], h0x, l1x, d2x);
            met = slab_met((h0x + h0y) + h0z, (l1x + l1y) + l1z) && (d2x + d2y) + d2z < reach2;
        }
        key = rowkey + wk[0].bi;
        code = wk[0].im + 3 * wk[1].im + 9 * wk[2].im;
        if (++jx < nx) {
            wk[0].next(ab[0]);
        } else if (t + 1 < n) {
            jx = 0;
            wk[0].start(ab[0]);
            if (++jy < ny) {
                wk[1].next(ab[1]);
            } else {
                jy = 0;
                wk[1].start(ab[1]);
                wk[2].next(ab[2]);
                if (SLAB) slab_terms(sg, 2, wk[2].bi, wk[2].im, bd[2], h0z, l1z, d2z);
            }
            if (SLAB) slab_terms(sg, 1, wk[1].bi, wk[1].im, bd[1], h0y, l1y, d2y);
            rowkey = brick0 + ((int64_t)wk[2].bi * nby + wk[1].bi) * nbx;
        }
        return met;
    }
};

__device__ __forceinline__ int32_t wave_max(int32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const int32_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Count pass: counts pairs per brick, flags boxes outside [1-N, 2N], lists the sub-segments with
// more than kBigPairs pairs in big[0 .. *nbig) (entries beyond big_cap are not stored: the host
// then grows the list and repeats the pass) and writes every sub-point's BinRec.
__global__ __launch_bounds__(256) void k_pairs_count(Geo g, Run r0, uint32_t *__restrict__ cnt,
                                                     BinRec *__restrict__ rec, int32_t *__restrict__ err,
                                                     uint32_t *__restrict__ big, uint32_t *__restrict__ nbig,
                                                     uint32_t big_cap) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool in = q < g.nsub_all;
    SubSeg sg;
    bool ok = in && subseg_load_fast(g, q, sg);
    AxisImages ab[3];
    int32_t n = 0, nx = 1, ny = 1;
    int64_t brick0 = 0;
    int32_t nbx = 1, nby = 1, bd[3] = {1, 1, 1};
    if (ok) {
        const Run &r = g.runs[g.tube_inst[sg.tube]];
        brick0 = r.brick0;
        nbx = r.nb[0];
        nby = r.nb[1];
#pragma unroll
        for (int c = 0; c < 3; c++) bd[c] = r.bd[c];
#pragma unroll
        for (int c = 0; c < 3; c++)
            if (sg.lo[c] < 1 - sg.N || sg.hi[c] > 2 * sg.N) ok = false;
        if (!ok) {
            atomicOr(err, 2);
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) ab[c] = axis_images(sg.lo[c], sg.hi[c], sg.N, r.lbd[c]);
            nx = ab[0].count();
            ny = ab[1].count();
            n = nx * ny * ab[2].count();
            if (n > kBigPairs) {
                const uint32_t k = atomicAdd(nbig, 1u);
                if (k < big_cap) big[k] = (uint32_t)q;
                n = 0;
            }
        }
    }
    const int32_t nmax = wave_max(n);
    EntryWalk<true> ew;
    if (n > 0) ew.start(ab, sg, bd, brick0, nbx, nby);
    uint32_t mask = 0;
    for (int32_t t = 0; t < nmax; t++) {
        bool act = t < n;
        int64_t key = 0;
        int32_t code = 0;
        if (act) act = ew.take(ab, sg, bd, brick0, nbx, nby, nx, ny, t, n, key, code);
        if (act && t < kRecMaskBits) mask |= 1u << t;
        wave_count_add(cnt, key, act);
    }
    if (in) {
        BinRec o;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o.lo[c] = (int16_t)(ok ? sg.lo[c] : 1);
            o.hi[c] = (int16_t)(ok ? sg.hi[c] : 0);
        }
        o.mask = mask;
        rec[q] = o;
    }
}

// Fill pass: writes every pair (sub-point index, image code kx + 3ky + 9kz) into its brick's list,
// after the brick's wide[key] pairs of wide boxes (k_pairs_big), which lead it: the deposition
// spreads those expensive pairs over its waves before it takes the rest in runs. The boxes and
// (up to kRecMaskBits entries) the slab tests come from the count pass's records: such a lane
// steps through the set bits of its mask only, forming each entry's brick and image from its
// index; a box of more entries walks all of them and repeats its slab tests. The two kinds run as
// two launches (SLOW), so that the first, which has nearly all the pairs, keeps few registers
// and many waves in flight for its returned atomics.
template <bool SLOW>
__global__ __launch_bounds__(256) void k_pairs_fill(Geo g, Run r0, uint32_t *__restrict__ cnt,
                                                    const BinRec *__restrict__ rec, const int64_t *__restrict__ off,
                                                    const uint32_t *__restrict__ wide, uint32_t *__restrict__ pairs,
                                                    uint8_t *__restrict__ pair_img) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    BinRec rc;
    int32_t n = 0, nx = 1, ny = 1;
    int64_t brick0 = 0;
    int32_t nbx = 1, nby = 1, bd[3] = {1, 1, 1};
    AxisImages ab[3];
    if (q < g.nsub_all) {
        rc = rec[q];
        if (rc.lo[0] <= rc.hi[0]) {
            const Run &r = g.runs[g.tube_inst[g.sub_tube[q]]];
            brick0 = r.brick0;
            nbx = r.nb[0];
            nby = r.nb[1];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                bd[c] = r.bd[c];
                ab[c] = axis_images(rc.lo[c], rc.hi[c], r.N, r.lbd[c]);
            }
            nx = ab[0].count();
            ny = ab[1].count();
            n = nx * ny * ab[2].count();
            if (n > kBigPairs) n = 0;   // k_pairs_big's
        }
    }
    const bool slow = n > kRecMaskBits;
    if (SLOW != slow) n = 0;   // the other launch's
    SubSeg sg;
    if (SLOW && n > 0) subseg_load_fast(g, q, sg);   // a paired sub-segment: it loads
    uint32_t m = (!SLOW && n > 0) ? rc.mask : 0u;
    const int32_t nmax = wave_max(SLOW ? n : __popc(m));
    EntryWalk<true> ews;     // slow lanes: slab tests again
    if (SLOW && n > 0) ews.start(ab, sg, bd, brick0, nbx, nby);
    // entry e = jx + nx (jy + ny jz): quotients by nx, ny from single-precision reciprocals
    // ((e + 1/2) / nx is at least 1/64 from an integer for e < 32)
    const float inx = 1.0f / (float)nx, iny = 1.0f / (float)ny;
    for (int32_t t = 0; t < nmax; t++) {
        bool act = false;
        int64_t key = 0;
        int32_t code = 0;
        if (SLOW) {
            act = t < n && ews.take(ab, sg, bd, brick0, nbx, nby, nx, ny, t, n, key, code);
        } else if (m != 0u) {
            const int32_t e = __builtin_ctz(m);
            m &= m - 1u;
            const int32_t q1 = (int32_t)__builtin_floorf(((float)e + 0.5f) * inx);
            const int32_t q2 = (int32_t)__builtin_floorf(((float)q1 + 0.5f) * iny);
            int32_t bx, by, bz, ix, iy, iz;
            ab[0].at(e - q1 * nx, bx, ix);
            ab[1].at(q1 - q2 * ny, by, iy);
            ab[2].at(q2, bz, iz);
            key = brick0 + ((int64_t)bz * nby + by) * nbx + bx;
            code = ix + 3 * iy + 9 * iz;
            act = true;
        }
        const uint32_t slot = wave_slot_add(cnt, key, act);
        if (act) {
            const int64_t at = off[key] + wide[key] + slot;
            pairs[at] = (uint32_t)q;
            pair_img[at] = (uint8_t)code;
        }
    }
}

// The sub-segments listed by k_pairs, one wave each (grid-stride over the list): lane t takes
// pairs t, t + 64, ... of the box (pairs of one box are distinct bricks or images). The count pass
// adds each to its brick's total (cnt) and wide-pair count (wide); the fill pass writes it at the
// slot its brick's wide-pair counter (wide: zeroed slots) gives, at the head of the brick's list.
template <bool FILL>
__global__ __launch_bounds__(256) void k_pairs_big(Geo g, Run r0, uint32_t *__restrict__ cnt,
                                                   uint32_t *__restrict__ wide, const int64_t *__restrict__ off,
                                                   uint32_t *__restrict__ pairs,
                                                   uint8_t *__restrict__ pair_img, const uint32_t *__restrict__ big,
                                                   const uint32_t *__restrict__ nbig, uint32_t big_cap) {
    const uint32_t nb = *nbig < big_cap ? *nbig : big_cap;
    const int lane = threadIdx.x & 63;
    for (uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nb; w += gridDim.x * (blockDim.x >> 6)) {
        const int64_t q = big[w];
        SubSeg sg;
        if (!subseg_load_fast(g, q, sg)) continue;   // listed sub-segments always load
        const Run &r = g.runs[g.tube_inst[sg.tube]];
        AxisImages ab[3];
#pragma unroll
        for (int c = 0; c < 3; c++) ab[c] = axis_images(sg.lo[c], sg.hi[c], sg.N, r.lbd[c]);
        const int32_t nx = ab[0].count(), ny = ab[1].count(), n = nx * ny * ab[2].count();
        const double reach2 = hit_reach2(sg);
        for (int32_t t = lane; t < n; t += 64) {
            int32_t bi[3], im[3];
            ab[0].at(t % nx, bi[0], im[0]);
            ab[1].at((t / nx) % ny, bi[1], im[1]);
            ab[2].at(t / (nx * ny), bi[2], im[2]);
            if (!pair_meets_slab(sg, bi, im, r.bd, reach2)) continue;
            const int64_t key = r.brick0 + ((int64_t)bi[2] * r.nb[1] + bi[1]) * r.nb[0] + bi[0];
            if (FILL) {
                const uint32_t slot = atomicAdd(wide + key, 1u);   // wslot: the wide pairs lead the list
                pairs[off[key] + slot] = (uint32_t)q;
                pair_img[off[key] + slot] = (uint8_t)(im[0] + 3 * im[1] + 9 * im[2]);
            } else {
                atomicAdd(cnt + key, 1u);
                atomicAdd(wide + key, 1u);
            }
        }
    }
}

static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
static hipError_t balloc(wt_ctx *c, T **p, size_t count) {
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) return e;
    c->bin_allocs.push_back(v);
    *p = (T *)v;
    return hipSuccess;
}

// Count pass of the binning: pairs per brick, their offsets, the wide-box list.
static hipError_t bin_count_kernels(wt_ctx *c) {
    Geo &g = c->geo;
    Bins &b = c->bins;
    BinWs &w = c->binws;
    hipStream_t st = c->stream;
    WT_HIP(fill_zero(w.cnt, sizeof(uint32_t) * b.nbricks, st));
    WT_HIP(fill_zero(b.brick_wide, sizeof(uint32_t) * b.nbricks, st));
    WT_HIP(fill_zero(w.err, sizeof(int32_t), st));
    WT_HIP(fill_zero(w.nbig, sizeof(uint32_t), st));
    kt_begin(c, kSlotBinCount);
    hipLaunchKernelGGL(k_pairs_count, dim3(grid_for(g.nsub_all, 256)), dim3(256), 0, st, g, c->run, w.cnt, w.rec,
                       w.err, w.big, w.nbig, w.big_cap);
    hipLaunchKernelGGL(k_pairs_big<false>, dim3(w.big_grid), dim3(256), 0, st, g, c->run, w.cnt, b.brick_wide,
                       (const int64_t *)nullptr, (uint32_t *)nullptr, (uint8_t *)nullptr, (const uint32_t *)w.big,
                       (const uint32_t *)w.nbig, w.big_cap);
    kt_end(c, kSlotBinCount);
    WT_HIP(hipGetLastError());
    return exclusive_scan_i32_to_i64((const int32_t *)w.cnt, 0, b.brick_off, b.nbricks, st, w.scan);
}

// Fill pass: the (sub-point, image) pairs of every brick at its offsets.
static hipError_t bin_fill_kernels(wt_ctx *c) {
    Geo &g = c->geo;
    Bins &b = c->bins;
    BinWs &w = c->binws;
    hipStream_t st = c->stream;
    WT_HIP(fill_zero(w.cnt, sizeof(uint32_t) * b.nbricks, st));
    WT_HIP(fill_zero(w.wslot, sizeof(uint32_t) * b.nbricks, st));
    kt_begin(c, kSlotBinFill);
    // the boxes of more than 32 entries first: their pairs, costlier than most, then follow the
    // wide boxes' at the head of each brick's list, and the cheap pairs fill the list's end
    hipLaunchKernelGGL(k_pairs_fill<true>, dim3(grid_for(g.nsub_all, 256)), dim3(256), 0, st, g, c->run, w.cnt,
                       (const BinRec *)w.rec, (const int64_t *)b.brick_off, (const uint32_t *)b.brick_wide, b.pairs,
                       b.pair_img);
    hipLaunchKernelGGL(k_pairs_fill<false>, dim3(grid_for(g.nsub_all, 256)), dim3(256), 0, st, g, c->run, w.cnt,
                       (const BinRec *)w.rec, (const int64_t *)b.brick_off, (const uint32_t *)b.brick_wide, b.pairs,
                       b.pair_img);
    hipLaunchKernelGGL(k_pairs_big<true>, dim3(w.big_grid), dim3(256), 0, st, g, c->run, w.cnt, w.wslot,
                       (const int64_t *)b.brick_off, b.pairs, b.pair_img, (const uint32_t *)w.big,
                       (const uint32_t *)w.nbig, w.big_cap);
    kt_end(c, kSlotBinFill);
    return hipGetLastError();
}

hipError_t launch_bins(wt_ctx *c) {
    Geo &g = c->geo;
    Bins &b = c->bins;
    BinWs &w = c->binws;
    hipStream_t st = c->stream;
    for (void *p : c->bin_allocs) (void)hipFree(p);
    c->bin_allocs.clear();
    w = BinWs{};
    // brick table: (instance, bx, by, bz) of every brick, instances in arena order
    int64_t nbricks = 0;
    for (int i = 0; i < c->ninst(); i++) {
        const Run &r = c->runs_host[i];
        nbricks = std::max<int64_t>(nbricks, r.brick0 + (int64_t)r.nb[0] * r.nb[1] * r.nb[2]);
    }
    b.nbricks = nbricks;
    std::vector<int4> tab((size_t)nbricks);
    for (int i = 0; i < c->ninst(); i++) {
        const Run &r = c->runs_host[i];
        for (int z = 0; z < r.nb[2]; z++)
            for (int y = 0; y < r.nb[1]; y++)
                for (int x = 0; x < r.nb[0]; x++)
                    tab[(size_t)(r.brick0 + ((int64_t)z * r.nb[1] + y) * r.nb[0] + x)] = make_int4(i, x, y, z);
    }
    WT_HIP(balloc(c, &b.brick_tab, nbricks));
    WT_HIP(balloc(c, &b.brick_off, nbricks + 1));
    WT_HIP(balloc(c, &w.cnt, nbricks));
    WT_HIP(balloc(c, &w.rec, g.nsub_all));
    WT_HIP(balloc(c, &w.wslot, nbricks));
    WT_HIP(balloc(c, &b.brick_wide, nbricks));
    WT_HIP(balloc(c, &w.err, 1));
    WT_HIP(balloc(c, &w.scan, scan_scratch_elems(nbricks)));
    // list of wide boxes: sized for one sub-segment in 64 (at least 1024), grown if the count pass
    // finds more
    w.big_cap = (uint32_t)std::min<int64_t>(std::max<int64_t>(g.nsub_all / 64, 1024), g.nsub_all + 1);
    WT_HIP(balloc(c, &w.big, w.big_cap));
    WT_HIP(balloc(c, &w.nbig, 1));
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    w.big_grid = (unsigned)(4 * cus);
    WT_HIP(hipMemcpyAsync(b.brick_tab, tab.data(), sizeof(int4) * nbricks, hipMemcpyHostToDevice, st));
    int64_t npairs = 0;
    for (int attempt = 0;; attempt++) {
        WT_HIP(bin_count_kernels(c));
        int32_t herr = 0;
        uint32_t hbig = 0;
        WT_HIP(hipMemcpyAsync(&npairs, b.brick_off + nbricks, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        WT_HIP(hipMemcpyAsync(&herr, w.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        WT_HIP(hipMemcpyAsync(&hbig, w.nbig, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        WT_HIP(hipStreamSynchronize(st));
        if (herr) return hipErrorInvalidValue;
        if (hbig <= w.big_cap) break;
        if (attempt > 0) return hipErrorUnknown;   // the list was sized from the first count
        // too many wide boxes for the list: grow it and count again
        w.big_cap = hbig;
        WT_HIP(balloc(c, &w.big, w.big_cap));
    }
    b.npairs = npairs;
    // the pair list is kept between calls and grown when needed
    if ((size_t)npairs > c->pairs_cap) {
        if (c->pairs_buf) (void)hipFree(c->pairs_buf);
        c->pairs_buf = nullptr;
        c->pairs_cap = 0;
        size_t cap = (size_t)npairs + (size_t)npairs / 8 + 1024;
        WT_HIP(hipMalloc((void **)&c->pairs_buf, (sizeof(uint32_t) + sizeof(uint8_t)) * cap));
        c->pairs_cap = cap;
    }
    b.pairs = c->pairs_buf;
    b.pair_img = (uint8_t *)(c->pairs_buf + c->pairs_cap);
    return bin_fill_kernels(c);
}

// The binning again on the sub-segments and buffers of the last launch_bins (the same tubes give
// the same counts): kernels only, for recording into a graph.
hipError_t replay_bins(wt_ctx *c) {
    WT_HIP(bin_count_kernels(c));
    return bin_fill_kernels(c);
}

}  // namespace wt
