// C ABI of the woven-turbulence generator (declarations and reference citations in
// include/wovengpu.h). Owns device memory, the HIP stream, the instances of a batch and the
// call-order checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "wt_internal.h"

namespace wt {
int64_t nek_of(int N) {
    float v = (float)(3 * N * N) / 12.0f;  // get_spectrum: int(sqrt((nx^2+ny^2+nz^2)/12.0))
    return (int64_t)std::sqrt(v);
}

bool dep_tiles_valid(const int32_t t[4]) {
    for (int a = 0; a < 3; a++)
        if (t[a] < kBrickMin || t[a] > kBrickMax || (t[a] & (t[a] - 1))) return false;
    return (int64_t)t[0] * t[1] * t[2] <= kBrickCells && t[3] >= 1 && t[3] <= kColCapMax;
}

void dep_tiles_of(const wt_ctx *c, int N, int32_t out[4]) {
    for (int k = 0; k < 4; k++) out[k] = kDefaultTiles[k];
    for (size_t i = 0; i + 5 <= c->dep_tiles.size(); i += 5)
        if (c->dep_tiles[i] == N)
            for (int k = 0; k < 4; k++) out[k] = c->dep_tiles[i + 1 + k];
}
}  // namespace wt

static thread_local char g_create_err[256] = "";

static int set_err(wt_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else snprintf(g_create_err, sizeof g_create_err, "%s", buf);
    return code;
}

static int hip_err(wt_ctx *c, hipError_t e, const char *where) {
    return set_err(c, e == hipErrorOutOfMemory ? WT_ENOMEM : WT_EDEVICE, "%s: HIP error %d (%s)", where, (int)e,
                   hipGetErrorString(e));
}

#define CHECK_CTX(c)                                   \
    do {                                               \
        if (!(c)) return WT_EINVAL;                    \
        (c)->err.clear();                              \
        hipError_t _s = hipSetDevice((c)->device);     \
        if (_s != hipSuccess) return hip_err(c, _s, "hipSetDevice"); \
    } while (0)

#define HIPCALL(c, expr, where)                        \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return hip_err(c, _e, where); \
    } while (0)

static size_t mem_cap_bytes() {
    const char *s = getenv("WT_DEVICE_MEM_CAP");  // test knob: forces WT_ENOMEM without a GPU-sized input
    if (!s || !*s) return 0;
    return (size_t)strtoull(s, nullptr, 10);
}

static size_t field_bytes(int64_t field_elems, int64_t spec_elems, int precision) {
    size_t es = precision == 64 ? 8 : 4;
    return 2 * (size_t)field_elems * es + (size_t)spec_elems * 2 * es;
}

template <typename T>
static hipError_t alloc(T **p, size_t n) {
    return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

static void free_list(std::vector<void *> &l) {
    for (void *p : l) (void)hipFree(p);
    l.clear();
}

static int pos_of(const wt_ctx *c, int inst) {
    for (int p = 0; p < c->ninst(); p++)
        if (c->order[p] == inst) return p;
    return -1;
}

extern "C" {

const char *wt_last_error(const wt_ctx *c) { return c ? c->err.c_str() : g_create_err; }

int32_t wt_nek(int32_t n) { return (int32_t)wt::nek_of(n); }

int32_t wt_sp_separations(int32_t n, int32_t *r, int32_t cap) {
    // subroutine.f90:78, 97-98 in FP32: lg2L = floor(log(real(nx))/log(2.0)+0.01),
    // r = floor(sqrt(2.0)**lg2r + 0.5), lg2r = 0..2*lg2L
    if (n < 2) return WT_EINVAL;
    int lg2L = (int)std::floor(std::log((float)n) / std::log(2.0f) + 0.01f);
    int nr = 2 * lg2L + 1;
    if (!r) return nr;
    if (cap < nr) return WT_ERANGE;
    float s2 = std::sqrt(2.0f);
    for (int lr = 0; lr < nr; lr++) {
        float v = 1.0f;
        for (int t = 0; t < lr; t++) v = v * s2;
        r[lr] = (int32_t)std::floor(v + 0.5f);
    }
    return nr;
}

void wt_destroy(wt_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_list(c->bin_allocs);
    free_list(c->seg_allocs);
    free_list(c->geo_allocs);
    if (c->spec_exec) (void)hipGraphExecDestroy(c->spec_exec);
    if (c->spec_graph) (void)hipGraphDestroy(c->spec_graph);
    if (c->gen_exec) (void)hipGraphExecDestroy(c->gen_exec);
    if (c->gen_graph) (void)hipGraphDestroy(c->gen_graph);
    void *bufs[] = {c->omega, c->u, c->spec, c->pairs_buf, c->partials, c->scal, c->E_w, c->E_u, c->runs_dev,
                    c->sf_items, c->dep_cnt};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (void *p : c->tw_tab)
        if (p) (void)hipFree(p);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &pair : c->kev)
        for (auto &ev : pair)
            if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int wt_create_batch(wt_ctx **out, int32_t ninst, const int32_t *n, int32_t precision, int32_t device) {
    if (!out) return set_err(nullptr, WT_EINVAL, "wt_create: ctx pointer is NULL");
    *out = nullptr;
    if (ninst <= 0 || !n) return set_err(nullptr, WT_EINVAL, "need at least one instance");
    if (precision != 32 && precision != 64) return set_err(nullptr, WT_EINVAL, "precision must be 32 or 64");
    for (int i = 0; i < ninst; i++)
        if (!wt::fft_size_supported(n[i]))
            return set_err(nullptr, WT_ERANGE,
                           "N=%d unsupported: need even 4 <= N <= 1024 with factors 2 and 3 only", n[i]);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return set_err(nullptr, WT_EDEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return set_err(nullptr, WT_EINVAL, "device %d out of range (%d)", device, ndev);
    wt_ctx *c = new (std::nothrow) wt_ctx();
    if (!c) return set_err(nullptr, WT_ENOMEM, "host allocation failed");
    c->device = device;
    c->precision = precision;
    c->inst.resize(ninst);
    c->order.resize(ninst);
    std::iota(c->order.begin(), c->order.end(), 0);
    std::stable_sort(c->order.begin(), c->order.end(), [&](int a, int b) { return n[a] < n[b]; });
    // arena placement in order of N, so that equal-N instances are contiguous
    int64_t fo = 0, so = 0, eo = 0;
    for (int pos = 0; pos < ninst; pos++) {
        wt::Instance &I = c->inst[c->order[pos]];
        int64_t N = n[c->order[pos]];
        I.N = (int)N;
        I.field_off = fo;
        I.spec_off = so;
        I.nek = (int)wt::nek_of((int)N);
        I.E_off = eo;
        fo += 3 * N * N * N;
        so += wt::spec_elems_of((int)N, precision);
        eo += I.nek;
    }
    c->field_elems = fo;
    c->spec_elems = so;
    c->E_elems = eo;
    c->N = c->inst[0].N;
    size_t cap = mem_cap_bytes();
    if (cap && field_bytes(fo, so, precision) > cap) {
        size_t need = field_bytes(fo, so, precision);
        delete c;
        return set_err(nullptr, WT_ENOMEM, "fields need %zu bytes, above WT_DEVICE_MEM_CAP=%zu", need, cap);
    }
    auto fail = [&](hipError_t er, const char *w) {
        int code = er == hipErrorOutOfMemory ? WT_ENOMEM : WT_EDEVICE;
        set_err(nullptr, code, "%s: %s", w, hipGetErrorString(er));
        wt_destroy(c);
        return code;
    };
    if ((e = hipSetDevice(device)) != hipSuccess) return fail(e, "hipSetDevice");
    if ((e = wt::sf_scratch_init(device)) != hipSuccess) return fail(e, "scratch address");
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    for (auto &ev : c->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(e, "event");
    for (int s = 0; s < wt::kNumSlots; s++)
        for (int k = 0; k < 2; k++)
            if ((e = hipEventCreate(&c->kev[s][k])) != hipSuccess) return fail(e, "event");
    const size_t es = precision == 64 ? 8 : 4;
    if ((e = hipMalloc(&c->omega, fo * es)) != hipSuccess) return fail(e, "omega");
    if ((e = hipMalloc(&c->u, fo * es)) != hipSuccess) return fail(e, "u");
    if ((e = hipMalloc(&c->spec, so * 2 * es)) != hipSuccess) return fail(e, "spectrum buffer");
    // per grid size: twiddles exp(-2 pi i k / N) in the compute precision, then the
    // digit-reversal table
    for (int pos = 0; pos < ninst; pos++) {
        int N = c->inst[c->order[pos]].N;
        if (std::find(c->tw_N.begin(), c->tw_N.end(), N) != c->tw_N.end()) continue;
        std::vector<int32_t> rev;
        wt::fft_reversal(N, rev, true);
        size_t twb = 2 * (size_t)N * es + (size_t)N * sizeof(int32_t);
        std::vector<unsigned char> host(twb);
        for (int k = 0; k < N; k++) {
            long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)N;
            double cr = (double)cosl(ang), ci = (double)sinl(ang);
            if (precision == 64) {
                ((double *)host.data())[2 * k] = cr;
                ((double *)host.data())[2 * k + 1] = ci;
            } else {
                ((float *)host.data())[2 * k] = (float)cr;
                ((float *)host.data())[2 * k + 1] = (float)ci;
            }
        }
        memcpy(host.data() + 2 * (size_t)N * es, rev.data(), (size_t)N * sizeof(int32_t));
        void *d = nullptr;
        if ((e = hipMalloc(&d, twb)) != hipSuccess) return fail(e, "twiddles");
        c->tw_N.push_back(N);
        c->tw_tab.push_back(d);
        if ((e = hipMemcpy(d, host.data(), twb, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "twiddles");
    }
    if ((e = hipMalloc((void **)&c->scal, sizeof(wt::DeviceScalars) * ninst)) != hipSuccess) return fail(e, "scalars");
    if ((e = hipMemset(c->scal, 0, sizeof(wt::DeviceScalars) * ninst)) != hipSuccess) return fail(e, "scalars");
    if ((e = hipMalloc((void **)&c->E_w, sizeof(double) * (eo + 1))) != hipSuccess) return fail(e, "spectra");
    if ((e = hipMalloc((void **)&c->E_u, sizeof(double) * (eo + 1))) != hipSuccess) return fail(e, "spectra");
    if ((e = hipMemset(c->omega, 0, fo * es)) != hipSuccess) return fail(e, "omega");
    if ((e = hipMemset(c->u, 0, fo * es)) != hipSuccess) return fail(e, "u");
    if ((e = hipMalloc((void **)&c->runs_dev, sizeof(wt::Run) * ninst)) != hipSuccess) return fail(e, "runs");
    {
        const char *te = getenv("WT_THETA_EXACT");
        c->theta_exact_all = te && te[0] == '1';
    }
    *out = c;
    return WT_OK;
}

int wt_create(wt_ctx **out, int32_t n, int32_t precision, int32_t device) {
    return wt_create_batch(out, 1, &n, precision, device);
}

int wt_select_instance(wt_ctx *c, int32_t inst) {
    CHECK_CTX(c);
    if (inst < 0 || inst >= c->ninst()) return set_err(c, WT_EINVAL, "instance %d out of range (%d)", inst, c->ninst());
    c->cur = inst;
    c->N = c->inst[inst].N;
    return WT_OK;
}

int wt_set_params(wt_ctx *c, const wt_params *p) {
    CHECK_CTX(c);
    if (!p) return set_err(c, WT_EINVAL, "params is NULL");
    wt::Instance &I = c->inst[c->cur];
    if (p->N != I.N) return set_err(c, WT_EINVAL, "params N=%d differs from the instance's N=%d", p->N, I.N);
    if (p->ntf_times < 1 || p->ntf_pre < 1) return set_err(c, WT_EINVAL, "ntf_times and ntf_pre must be >= 1");
    if (!(p->Rtube_ratio > 0) || !(p->dzeta_sup > 0) || !(p->sigma_ratio > 0) || !std::isfinite(p->eta_t))
        return set_err(c, WT_EINVAL, "Rtube_ratio, dzeta_sup and sigma_ratio must be positive, eta_t finite");
    if (p->line_s < 1 || p->line_e < p->line_s) return set_err(c, WT_EINVAL, "need 1 <= line_s <= line_e");
    I.params = *p;
    I.have_params = true;
    c->tubes_dirty = true;
    c->have_segments = c->have_bins = false;
    return WT_OK;
}

int wt_set_tubes(wt_ctx *c, int32_t nline, const int32_t *npoint, const double *sigma, const double *gamma,
                 const int32_t *n_T, const int32_t *if_filter, const double *cx, const double *cy, const double *cz) {
    CHECK_CTX(c);
    if (nline <= 0 || !npoint || !sigma || !gamma || !n_T || !cx || !cy || !cz)
        return set_err(c, WT_EINVAL, "wt_set_tubes: need nline > 0 and non-NULL arrays");
    int64_t P = 0;
    for (int l = 0; l < nline; l++) {
        if (npoint[l] < 3) return set_err(c, WT_EINVAL, "tube %d has %d points; need at least 3", l + 1, npoint[l]);
        if (!(sigma[l] > 0) || !std::isfinite(sigma[l]) || !std::isfinite(gamma[l]))
            return set_err(c, WT_EINVAL, "tube %d: sigma must be positive and finite, Gamma finite", l + 1);
        P += npoint[l];
    }
    for (int64_t i = 0; i < P; i++)
        if (!std::isfinite(cx[i]) || !std::isfinite(cy[i]) || !std::isfinite(cz[i]))
            return set_err(c, WT_EINVAL, "non-finite centerline coordinate at point %lld", (long long)i + 1);
    wt::Instance &I = c->inst[c->cur];
    I.npoint.assign(npoint, npoint + nline);
    I.n_T.assign(n_T, n_T + nline);
    I.if_filter.assign(nline, 0);
    if (if_filter)
        for (int l = 0; l < nline; l++) I.if_filter[l] = if_filter[l] != 0;
    I.sigma.assign(sigma, sigma + nline);
    I.gamma.assign(gamma, gamma + nline);
    I.cx.assign(cx, cx + P);
    I.cy.assign(cy, cy + P);
    I.cz.assign(cz, cz + P);
    I.have_tubes = true;
    c->tubes_dirty = true;
    c->have_segments = c->have_bins = false;
    return WT_OK;
}

// Concatenates every instance's tubes into the device tube table.
static int upload_tubes(wt_ctx *c) {
    const int ni = c->ninst();
    for (int i = 0; i < ni; i++)
        if (!c->inst[i].have_params || !c->inst[i].have_tubes)
            return set_err(c, WT_ESTATE, "instance %d needs wt_set_params and wt_set_tubes", i);
    int64_t nline = 0, P = 0;
    for (auto &I : c->inst) {
        I.tube0 = (int32_t)nline;
        I.ntube = (int32_t)I.npoint.size();
        nline += I.ntube;
        P += (int64_t)I.cx.size();
        if (I.params.line_e > I.ntube)
            return set_err(c, WT_EINVAL, "line_e=%d exceeds the %d tubes of its instance", I.params.line_e, I.ntube);
    }
    if (nline > INT32_MAX / 2) return set_err(c, WT_ERANGE, "too many tubes");
    std::vector<int64_t> pt(nline + 1), sec(nline + 1);
    std::vector<int32_t> npt, nT, tinst;
    std::vector<double> sig, gam, xs, ys, zs;
    npt.reserve(nline);
    xs.reserve(P);
    ys.reserve(P);
    zs.reserve(P);
    std::vector<double> bound(ni, 0.0);
    for (int i = 0; i < ni; i++) {
        const wt::Instance &I = c->inst[i];
        for (int l = 0; l < I.ntube; l++) {
            npt.push_back(I.npoint[l]);
            nT.push_back(I.n_T[l]);
            sig.push_back(I.sigma[l]);
            gam.push_back(I.gamma[l]);
            tinst.push_back(i);
            bound[i] += std::fabs(I.gamma[l]) / (wt::kPi2 * I.sigma[l] * I.sigma[l]);
        }
        xs.insert(xs.end(), I.cx.begin(), I.cx.end());
        ys.insert(ys.end(), I.cy.begin(), I.cy.end());
        zs.insert(zs.end(), I.cz.begin(), I.cz.end());
    }
    pt[0] = sec[0] = 0;
    for (int64_t l = 0; l < nline; l++) {
        pt[l + 1] = pt[l] + npt[l];
        sec[l + 1] = sec[l] + npt[l] + 1;
    }
    const int64_t S = sec[nline];
    if (S >= INT32_MAX) return set_err(c, WT_ERANGE, "%lld sections (points + one per tube) exceed 2^31 - 1", (long long)S);
    size_t cap = mem_cap_bytes();
    size_t need = (size_t)P * 8 * 13 + (size_t)S * 8 * 24;
    if (cap && field_bytes(c->field_elems, c->spec_elems, c->precision) + need > cap)
        return set_err(c, WT_ENOMEM, "tube data needs %zu bytes beyond the fields, above WT_DEVICE_MEM_CAP", need);
    // per-instance run table
    c->runs_host.resize(ni);
    int64_t brick0 = 0;
    for (int i = 0; i < ni; i++) {
        const wt::Instance &I = c->inst[i];
        const wt_params &p = I.params;
        wt::Run &r = c->runs_host[i];
        r.N = I.N;
        r.line_s = p.line_s - 1;
        r.line_e = p.line_e - 1;
        r.tube0 = I.tube0;
        r.sigma_ratio = p.sigma_ratio;
        r.eta = p.eta_t;
        r.Rtube_ratio = p.Rtube_ratio;
        r.dzeta_sup = p.dzeta_sup;
        r.ntf_times = p.ntf_times;
        r.ntf_pre = p.ntf_pre;
        r.dx = wt::kPi2 / I.N;  // lx/nx (tube.f90:274)
        r.start = -wt::kPi;
        r.field_off = I.field_off;
        int32_t tiles[4];
        wt::dep_tiles_of(c, I.N, tiles);
        for (int a = 0; a < 3; a++) {
            r.bd[a] = tiles[a];
            r.lbd[a] = __builtin_ctz((unsigned)tiles[a]);
            r.nb[a] = (I.N + tiles[a] - 1) / tiles[a];
        }
        r.colcap = tiles[3];
        r.brick0 = brick0;
        brick0 += (int64_t)r.nb[0] * r.nb[1] * r.nb[2];
        // fixed-point scale of the FP32 accumulation: |omega| <= sum Gamma/(2 pi sigma^2) over the
        // instance's tubes times a generous factor for the curvature and twist terms
        const double B = bound[i] * 64.0 + 1.0;
        const int shift = (int)std::floor(60.0 - std::log2(B));
        r.fx_shift = shift < 8 ? 8 : (shift > 44 ? 44 : shift);
        r.fx_scale = std::ldexp(1.0, r.fx_shift);
        r.fx_inv = std::ldexp(1.0, -r.fx_shift);
    }
    c->run = c->runs_host[0];
    c->fixed_shift = c->runs_host[0].fx_shift;
    for (const wt::Run &r : c->runs_host) c->fixed_shift = std::min(c->fixed_shift, r.fx_shift);
    free_list(c->bin_allocs);
    free_list(c->seg_allocs);
    free_list(c->geo_allocs);
    c->have_segments = c->have_bins = false;
    wt::Geo g{};
    g.nline = (int32_t)nline;
    g.npoint_all = P;
    g.nsec_all = S;
    g.runs = c->runs_dev;
    auto A = [&](auto **p, size_t n) -> hipError_t {
        hipError_t e = alloc(p, n);
        if (e == hipSuccess) c->geo_allocs.push_back((void *)*p);
        return e;
    };
    int64_t *pt_base, *sec_base, *sp_base;
    double *sigd, *gamd, *xd, *yd, *zd;
    int32_t *nTd, *tid;
    HIPCALL(c, A(&pt_base, nline + 1), "alloc");
    HIPCALL(c, A(&sec_base, nline + 1), "alloc");
    HIPCALL(c, A(&sigd, nline), "alloc");
    HIPCALL(c, A(&gamd, nline), "alloc");
    HIPCALL(c, A(&nTd, nline), "alloc");
    HIPCALL(c, A(&tid, nline), "alloc");
    HIPCALL(c, A(&g.Len_Disc, nline), "alloc");
    HIPCALL(c, A(&g.length, nline), "alloc");
    HIPCALL(c, A(&xd, P), "alloc");
    HIPCALL(c, A(&yd, P), "alloc");
    HIPCALL(c, A(&zd, P), "alloc");
    HIPCALL(c, A(&g.pt_tube, P), "alloc");
    HIPCALL(c, A(&g.chord, P), "alloc");
    HIPCALL(c, A(&g.cze, P), "alloc");
    HIPCALL(c, A(&g.ck, 3 * P), "alloc");
    HIPCALL(c, A(&g.ckk, 3 * P), "alloc");
    HIPCALL(c, A(&g.sec_tube, S), "alloc");
    HIPCALL(c, A(&g.coe, 18 * S), "alloc");
    HIPCALL(c, A(&g.czl, S), "alloc");
    HIPCALL(c, A(&g.czr, S), "alloc");
    HIPCALL(c, A(&g.ntf, S), "alloc");
    HIPCALL(c, A(&g.dtf, S), "alloc");
    HIPCALL(c, A(&sp_base, S + 1), "alloc");
    g.sp_base = sp_base;
    hipStream_t st = c->stream;
    HIPCALL(c, hipMemcpyAsync(c->runs_dev, c->runs_host.data(), sizeof(wt::Run) * ni, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(pt_base, pt.data(), sizeof(int64_t) * (nline + 1), hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(sec_base, sec.data(), sizeof(int64_t) * (nline + 1), hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(sigd, sig.data(), sizeof(double) * nline, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(gamd, gam.data(), sizeof(double) * nline, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(nTd, nT.data(), sizeof(int32_t) * nline, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(tid, tinst.data(), sizeof(int32_t) * nline, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(xd, xs.data(), sizeof(double) * P, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(yd, ys.data(), sizeof(double) * P, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(zd, zs.data(), sizeof(double) * P, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipStreamSynchronize(st), "copy");
    g.pt_base = pt_base;
    g.sec_base = sec_base;
    g.sigma = sigd;
    g.gamma = gamd;
    g.n_T = nTd;
    g.tube_inst = tid;
    g.cx = xd;
    g.cy = yd;
    g.cz = zd;
    c->geo = g;
    c->pt_base = pt;
    c->sec_base = sec;
    c->tubes_dirty = false;
    return WT_OK;
}

// The whole-generation graph records the layout of one tube table; anything that re-lays out the
// geometry, the binning or the FFT tiles drops it.
static void drop_gen_graph(wt_ctx *c) {
    if (c->gen_exec) (void)hipGraphExecDestroy(c->gen_exec);
    if (c->gen_graph) (void)hipGraphDestroy(c->gen_graph);
    c->gen_exec = nullptr;
    c->gen_graph = nullptr;
}

int wt_build_segments(wt_ctx *c, wt_seg_stats *out) {
    CHECK_CTX(c);
    if (c->tubes_dirty) {
        int rc = upload_tubes(c);
        if (rc) return rc;
    }
    c->have_segments = c->have_bins = false;
    drop_gen_graph(c);
    hipError_t e = wt::launch_segments(c);
    if (e == hipErrorInvalidValue)
        return set_err(c, WT_ERANGE, "turning angle not finite (degenerate spline: repeated or cusped points)");
    if (e == hipErrorInvalidImage)
        return set_err(c, WT_ERANGE, "more than 2^32 - 2 sub-points in one context; split the batch");
    if (e != hipSuccess) return hip_err(c, e, "segments");
    HIPCALL(c, wt::launch_seg_stats(c, &c->seg_stats), "segment statistics");
    c->seg_stats.npoint_all = c->geo.npoint_all;
    c->seg_stats.nsec_all = c->geo.nsec_all;
    c->seg_stats.nsub_all = c->geo.nsub_all;
    c->seg_stats.theta_exact_sections = c->theta_exact_all ? c->geo.nsec_all : (int64_t)c->theta_redo;
    c->have_segments = true;
    if (out) *out = c->seg_stats;
    return WT_OK;
}

int wt_get_seg_stats(wt_ctx *c, wt_seg_stats *out) {
    CHECK_CTX(c);
    if (!c->have_segments) return set_err(c, WT_ESTATE, "no segments built");
    if (!out) return set_err(c, WT_EINVAL, "NULL output");
    *out = c->seg_stats;
    return WT_OK;
}

int wt_bin(wt_ctx *c) {
    CHECK_CTX(c);
    if (!c->have_segments) return set_err(c, WT_ESTATE, "wt_bin needs wt_build_segments");
    drop_gen_graph(c);
    hipError_t e = wt::launch_bins(c);
    if (e == hipErrorInvalidValue)
        return set_err(c, WT_ERANGE, "a tube's cut-off radius spans more than the periodic box");
    if (e != hipSuccess) return hip_err(c, e, "bin");
    c->have_bins = true;
    return WT_OK;
}

// Tubes line_s..line_e of instance 0 with if_filter set (the reference's loop visits only those,
// tube.f90:529, 926).
static std::vector<int> filter_tubes(const wt_ctx *c) {
    std::vector<int> f;
    if (c->ninst() != 1) return f;
    const wt::Instance &I = c->inst[0];
    for (int l = I.params.line_s - 1; l <= I.params.line_e - 1 && l < (int)I.if_filter.size(); l++)
        if (I.if_filter[l]) f.push_back(l);
    return f;
}

static bool batch_has_filters(const wt_ctx *c) {
    for (const auto &I : c->inst)
        for (int32_t v : I.if_filter)
            if (v) return true;
    return false;
}

// The reference deposits tube by tube and filters the accumulated vorticity after each tube with
// if_filter = 1, with k_c = ceiling(4/sigma) of that tube (tube.f90:926-929). Here the tubes
// between two filters deposit as one segment (binned and deposited in one launch each, added to
// the field after the first), then the field is filtered.
static int deposit_filtered(wt_ctx *c, const std::vector<int> &flt) {
    wt::Run &r = c->runs_host[0];
    const int32_t s0 = r.line_s, e0 = r.line_e;
    struct Seg { int32_t a, b; int64_t kc2; };
    std::vector<Seg> segs;
    int32_t a = s0;
    for (int l : flt) {
        const double kc = std::ceil(4.0 / c->inst[0].sigma[l]);
        if (!(kc < 46341.0)) return set_err(c, WT_ERANGE, "tube %d: k_c = ceiling(4/sigma) exceeds 46340", l + 1);
        segs.push_back({a, l, (int64_t)kc * (int64_t)kc});
        a = l + 1;
    }
    if (a <= e0) segs.push_back({a, e0, -1});
    wt_deposit_stats tot{};
    auto set_range = [&](int32_t lo, int32_t hi) -> hipError_t {
        r.line_s = lo;
        r.line_e = hi;
        c->run = r;
        hipError_t e = hipMemcpyAsync(c->runs_dev, &r, sizeof r, hipMemcpyHostToDevice, c->stream);
        return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
    };
    int rc = WT_OK;
    for (size_t i = 0; i < segs.size() && rc == WT_OK; i++) {
        hipError_t e = set_range(segs[i].a, segs[i].b);
        if (e == hipSuccess) e = wt::launch_bins(c);
        if (e == hipErrorInvalidValue) {
            rc = set_err(c, WT_ERANGE, "a tube's cut-off radius spans more than the periodic box");
            break;
        }
        if (e == hipSuccess) e = wt::launch_deposit(c, i > 0);
        if (e == hipSuccess && segs[i].kc2 >= 0) e = wt::launch_filter(c, 0, segs[i].kc2);
        if (e != hipSuccess) {
            rc = hip_err(c, e, "filtered deposit");
            break;
        }
        tot.columns += c->dep_stats.columns;
        tot.candidates += c->dep_stats.candidates;
        tot.hits += c->dep_stats.hits;
        tot.pairs += c->dep_stats.pairs;
        tot.overflow |= c->dep_stats.overflow;
        tot.fixed_point_shift = c->dep_stats.fixed_point_shift;
    }
    hipError_t e = set_range(s0, e0);
    if (rc != WT_OK) return rc;
    if (e != hipSuccess) return hip_err(c, e, "filtered deposit");
    c->dep_stats = tot;
    c->have_bins = false;   // the bins now hold the last segment only
    return WT_OK;
}

int wt_deposit(wt_ctx *c, wt_deposit_stats *out) {
    CHECK_CTX(c);
    if (batch_has_filters(c) && c->ninst() != 1)
        return set_err(c, WT_EINVAL, "if_filter = 1 needs a single-instance context");
    const std::vector<int> flt = filter_tubes(c);
    if (!flt.empty()) {
        if (!c->have_segments) return set_err(c, WT_ESTATE, "wt_deposit needs wt_build_segments");
        int rc = deposit_filtered(c, flt);
        if (rc) return rc;
    } else {
        if (!c->have_bins) return set_err(c, WT_ESTATE, "wt_deposit needs wt_bin");
        hipError_t e = wt::launch_deposit(c);
        if (e != hipSuccess) return hip_err(c, e, "deposit");
    }
    if (out) *out = c->dep_stats;
    if (c->dep_stats.overflow)
        return set_err(c, WT_ERANGE, "fixed-point accumulator overflow (2^%d scale); use precision 64", c->fixed_shift);
    c->have_omega = true;
    c->have_u = false;
    c->have_spectra = false;
    return WT_OK;
}

int wt_filter_spec(wt_ctx *c, int32_t k_c) {
    CHECK_CTX(c);
    if (k_c < 0 || k_c > 46340) return set_err(c, WT_EINVAL, "need 0 <= k_c <= 46340");
    if (!c->have_omega) return set_err(c, WT_ESTATE, "the vorticity is not resident");
    hipError_t e = wt::launch_filter(c, c->cur, (int64_t)k_c * k_c);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_err(c, e, "filter");
    c->have_u = false;
    c->have_spectra = false;
    return WT_OK;
}

int wt_get_deposit_stats(wt_ctx *c, wt_deposit_stats *out) {
    CHECK_CTX(c);
    if (!c->have_omega || !out) return set_err(c, WT_ESTATE, "no deposition has run");
    *out = c->dep_stats;
    return WT_OK;
}

static int read_scalars(wt_ctx *c, int inst, wt::DeviceScalars *s) {
    int pos = pos_of(c, inst);
    HIPCALL(c, hipMemcpyAsync(s, c->scal + pos, sizeof *s, hipMemcpyDeviceToHost, c->stream), "scalars");
    HIPCALL(c, hipStreamSynchronize(c->stream), "scalars");
    return WT_OK;
}

// Records the spectral stage of every group (x, y, z + Biot-Savart + spectra, reductions, y and x
// inverse) as a graph on the context stream. The stage only reads the context's fixed arenas, so
// the graph stays valid for the context's lifetime.
static hipError_t capture_spectral(wt_ctx *c) {
    WT_HIP(wt::spectral_reserve(c));
    WT_HIP(hipStreamSynchronize(c->stream));
    WT_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    hipError_t e = wt::launch_spectral(c, 0, -1);
    c->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return e;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return e;
    }
    c->spec_graph = g;
    c->spec_exec = x;
    return hipSuccess;
}

// Autotune cache file: one JSON object of flat integer records. FFT tiles: key
// "<arch>:<precision>:<N>" -> {"ppb", "tx", "wy", "ty", "wz", "tz"}; deposition tiles: key
// "<arch>:<precision>:dep:<N>:r<class>" -> {"bx", "by", "bz", "colcap"}, where the class is the
// rounded log2 of the instance's point-weighted mean cut-off radius in cells. Entries of other
// devices, precisions, sizes and radius classes are kept.
static const char *kTileFields[6] = {"ppb", "tx", "wy", "ty", "wz", "tz"};
static const char *kDepFields[4] = {"bx", "by", "bz", "colcap"};
using CacheEntry = std::pair<std::string, std::vector<std::pair<std::string, int>>>;

static std::vector<CacheEntry> read_tile_cache(const char *path) {
    std::vector<CacheEntry> out;
    FILE *f = fopen(path, "rb");
    if (!f) return out;
    std::string txt;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, n);
    fclose(f);
    size_t p = txt.find('{');
    while (p != std::string::npos) {
        size_t k0 = txt.find('"', p + 1);
        if (k0 == std::string::npos) break;
        size_t k1 = txt.find('"', k0 + 1);
        size_t b0 = txt.find('{', k1);
        size_t b1 = txt.find('}', b0);
        if (k1 == std::string::npos || b0 == std::string::npos || b1 == std::string::npos) break;
        CacheEntry e;
        e.first = txt.substr(k0 + 1, k1 - k0 - 1);
        const std::string body = txt.substr(b0 + 1, b1 - b0 - 1);
        for (size_t q = body.find('"'); q != std::string::npos;) {
            const size_t q1 = body.find('"', q + 1), colon = body.find(':', q1);
            int x = 0;
            if (q1 == std::string::npos || colon == std::string::npos || sscanf(body.c_str() + colon + 1, "%d", &x) != 1)
                break;
            e.second.emplace_back(body.substr(q + 1, q1 - q - 1), x);
            q = body.find('"', colon + 1);   // values are integers: the next quote opens the next name
        }
        out.push_back(e);
        p = b1;
    }
    return out;
}

// The values of `fields` in entry e, in that order; false if one is missing.
static bool cache_values(const CacheEntry &e, const char *const *fields, int nf, std::vector<int> &v) {
    v.clear();
    for (int i = 0; i < nf; i++) {
        bool found = false;
        for (const auto &kv : e.second)
            if (kv.first == fields[i]) { v.push_back(kv.second); found = true; break; }
        if (!found) return false;
    }
    return true;
}

static void cache_put(std::vector<CacheEntry> &entries, const std::string &key, const char *const *fields,
                      const int32_t *v, int nf) {
    CacheEntry e;
    e.first = key;
    for (int i = 0; i < nf; i++) e.second.emplace_back(fields[i], v[i]);
    for (auto &en : entries)
        if (en.first == key) { en = e; return; }
    entries.push_back(e);
}

static bool write_tile_cache(const char *path, const std::vector<CacheEntry> &entries) {
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    fprintf(f, "{\n");
    for (size_t i = 0; i < entries.size(); i++) {
        fprintf(f, "  \"%s\": {", entries[i].first.c_str());
        for (size_t j = 0; j < entries[i].second.size(); j++)
            fprintf(f, "%s\"%s\": %d", j ? ", " : "", entries[i].second[j].first.c_str(), entries[i].second[j].second);
        fprintf(f, "}%s\n", i + 1 < entries.size() ? "," : "");
    }
    fprintf(f, "}\n");
    fclose(f);
    return true;
}


// Re-lays out the bricks of every instance after a change of the deposition tiles: the run table
// here and on the device now, the brick table and pair lists at the next binning.
static int apply_dep_tiles(wt_ctx *c) {
    drop_gen_graph(c);
    c->have_bins = false;
    if (c->runs_host.size() != (size_t)c->ninst()) return WT_OK;   // tubes not uploaded yet
    int64_t brick0 = 0;
    for (int i = 0; i < c->ninst(); i++) {
        wt::Run &r = c->runs_host[i];
        int32_t t[4];
        wt::dep_tiles_of(c, r.N, t);
        for (int a = 0; a < 3; a++) {
            r.bd[a] = t[a];
            r.lbd[a] = __builtin_ctz((unsigned)t[a]);
            r.nb[a] = (r.N + t[a] - 1) / t[a];
        }
        r.colcap = t[3];
        r.brick0 = brick0;
        brick0 += (int64_t)r.nb[0] * r.nb[1] * r.nb[2];
    }
    c->run = c->runs_host[0];
    HIPCALL(c, hipMemcpyAsync(c->runs_dev, c->runs_host.data(), sizeof(wt::Run) * c->ninst(), hipMemcpyHostToDevice,
                              c->stream), "runs");
    HIPCALL(c, hipStreamSynchronize(c->stream), "runs");
    return WT_OK;
}

static void set_tiles_of(wt_ctx *c, int N, const int32_t t[4]) {
    for (size_t i = 0; i + 5 <= c->dep_tiles.size(); i += 5)
        if (c->dep_tiles[i] == N) {
            std::copy(t, t + 4, c->dep_tiles.begin() + i + 1);
            return;
        }
    c->dep_tiles.push_back(N);
    c->dep_tiles.insert(c->dep_tiles.end(), t, t + 4);
}

int wt_set_deposit_tiles(wt_ctx *c, int32_t N, const int32_t *tiles) {
    CHECK_CTX(c);
    if (!tiles || !wt::dep_tiles_valid(tiles))
        return set_err(c, WT_EINVAL, "deposition tiles: brick edges powers of two in [%d, %d], at most %d cells, "
                       "1 <= colcap <= %d", wt::kBrickMin, wt::kBrickMax, wt::kBrickCells, wt::kColCapMax);
    if (N < 0) return set_err(c, WT_EINVAL, "N must be >= 0");
    bool any = false;
    for (const wt::Instance &I : c->inst)
        if (N == 0 || I.N == N) {
            set_tiles_of(c, I.N, tiles);
            any = true;
        }
    if (!any) return set_err(c, WT_EINVAL, "no instance of the context has N = %d", N);
    return apply_dep_tiles(c);
}

int wt_get_deposit_tiles(wt_ctx *c, int32_t N, int32_t *tiles) {
    CHECK_CTX(c);
    if (!tiles) return set_err(c, WT_EINVAL, "NULL tiles");
    wt::dep_tiles_of(c, N > 0 ? N : c->N, tiles);
    return WT_OK;
}

// Radius class of the instances of grid size N: rounded log2 of the point-weighted mean cut-off
// radius Rtube_ratio * sigma_max in cells.
static int radius_class(const wt_ctx *c, int N) {
    double w = 0.0, s = 0.0;
    for (const wt::Instance &I : c->inst) {
        if (I.N != N) continue;
        const double ratio = std::max(1.0, I.params.sigma_ratio);
        for (size_t l = 0; l < I.sigma.size(); l++) {
            w += I.npoint[l];
            s += I.npoint[l] * I.params.Rtube_ratio * ratio * I.sigma[l] * N / wt::kPi2;
        }
    }
    return w > 0.0 ? (int)std::lround(std::log2(std::max(s / w, 1e-3))) : 0;
}

// Binning plus deposition of the current tubes with the current tiles, in ms (best of two).
static int time_bin_deposit(wt_ctx *c, double *ms) {
    hipEvent_t e0, e1;
    HIPCALL(c, hipEventCreate(&e0), "event");
    HIPCALL(c, hipEventCreate(&e1), "event");
    double best = 1e30;
    int rc = WT_OK;
    for (int rep = 0; rep < 2 && rc == WT_OK; rep++) {
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = set_err(c, WT_EDEVICE, "event"); break; }
        if ((rc = wt_bin(c)) || (rc = wt_deposit(c, nullptr))) break;
        float t = 0.0f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) { rc = set_err(c, WT_EDEVICE, "event"); break; }
        best = std::min(best, (double)t);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return rc;
}

// Chooses the deposition tiles of every grid size of the context on its tubes (or takes them from
// the cache): brick shapes of 2^9..2^12 cells, then the column cap of the fastest, one grid size
// at a time with the others kept. The defaults are timed first; another choice is kept only if it
// is more than 0.5% faster than the best so far.
static int deposit_autotune(wt_ctx *c, const std::string &pre, std::vector<CacheEntry> &entries,
                            std::string &rep) {
    static const int32_t kShapes[][3] = {{16, 16, 16}, {32, 16, 8}, {16, 32, 8}, {8, 16, 32}, {32, 32, 4},
                                          {16, 16, 8}, {8, 8, 8}};
    static const int32_t kCaps[] = {4, 8, 16};
    std::vector<int> Ns;
    for (const wt::Instance &I : c->inst)
        if (std::find(Ns.begin(), Ns.end(), I.N) == Ns.end()) Ns.push_back(I.N);
    int rc = wt_build_segments(c, nullptr);
    if (rc) return rc;
    for (int N : Ns) {
        const std::string key = pre + "dep:" + std::to_string(N) + ":r" + std::to_string(radius_class(c, N));
        std::vector<int> v;
        bool hit = false;
        for (const auto &e : entries)
            if (e.first == key && cache_values(e, kDepFields, 4, v)) {
                int32_t t[4] = {v[0], v[1], v[2], v[3]};
                if (wt::dep_tiles_valid(t)) {
                    set_tiles_of(c, N, t);
                    if ((rc = apply_dep_tiles(c))) return rc;
                    hit = true;
                }
            }
        char line[256];
        if (hit) {
            snprintf(line, sizeof line, "deposit N=%d (%s) from cache: %dx%dx%d, colcap %d\n", N,
                     key.c_str() + pre.size(), v[0], v[1], v[2], v[3]);
            rep += line;
            continue;
        }
        int32_t cur[4];
        wt::dep_tiles_of(c, N, cur);
        int32_t best[4] = {cur[0], cur[1], cur[2], cur[3]};
        double best_ms = 1e30;
        auto trial = [&](const int32_t *t) -> int {
            set_tiles_of(c, N, t);
            int r = apply_dep_tiles(c);
            double ms = 0.0;
            if (!r) r = time_bin_deposit(c, &ms);
            if (r) return r;
            if (ms < 0.995 * best_ms) {   // a candidate must beat the best so far by more than timing noise
                best_ms = ms;
                std::copy(t, t + 4, best);
            }
            return WT_OK;
        };
        const double ms_default = [&]() {
            int32_t t[4] = {wt::kDefaultTiles[0], wt::kDefaultTiles[1], wt::kDefaultTiles[2], wt::kDefaultTiles[3]};
            rc = trial(t);
            return best_ms;
        }();
        if (rc) return rc;
        for (const auto &sh : kShapes) {
            int32_t t[4] = {sh[0], sh[1], sh[2], wt::kDefaultTiles[3]};
            if (std::equal(t, t + 3, wt::kDefaultTiles)) continue;
            if ((rc = trial(t))) return rc;
        }
        const int32_t shape[3] = {best[0], best[1], best[2]};
        for (int cap : kCaps) {
            if (cap == wt::kDefaultTiles[3]) continue;
            int32_t t[4] = {shape[0], shape[1], shape[2], cap};
            if ((rc = trial(t))) return rc;
        }
        set_tiles_of(c, N, best);
        if ((rc = apply_dep_tiles(c))) return rc;
        cache_put(entries, key, kDepFields, best, 4);
        snprintf(line, sizeof line, "deposit N=%d (%s): %dx%dx%d, colcap %d: binning + deposit %.2f ms (default %.2f ms)\n",
                 N, key.c_str() + pre.size(), best[0], best[1], best[2], best[3], best_ms, ms_default);
        rep += line;
    }
    return WT_OK;
}

int wt_autotune(wt_ctx *c, const char *cache_path, char *report, int32_t report_cap) {
    CHECK_CTX(c);
    hipDeviceProp_t prop;
    HIPCALL(c, hipGetDeviceProperties(&prop, c->device), "device properties");
    std::string arch = prop.gcnArchName;
    if (arch.find(':') != std::string::npos) arch = arch.substr(0, arch.find(':'));
    const std::string pre = arch + ":" + std::to_string(c->precision) + ":";
    std::vector<CacheEntry> entries = cache_path ? read_tile_cache(cache_path) : std::vector<CacheEntry>{};
    std::vector<int32_t> cached;
    for (const auto &e : entries) {
        std::vector<int> v;
        if (e.first.compare(0, pre.size(), pre) == 0 && e.first.find("dep:") == std::string::npos &&
            cache_values(e, kTileFields, 6, v)) {
            cached.push_back(atoi(e.first.c_str() + pre.size()));
            cached.insert(cached.end(), v.begin(), v.end());
        }
    }
    std::string rep;
    hipError_t e = wt::spectral_autotune(c, cached, rep);
    if (e != hipSuccess) return hip_err(c, e, "autotune");
    // the captured stages used the old shapes
    drop_gen_graph(c);
    if (c->spec_exec) (void)hipGraphExecDestroy(c->spec_exec);
    if (c->spec_graph) (void)hipGraphDestroy(c->spec_graph);
    c->spec_exec = nullptr;
    c->spec_graph = nullptr;
    c->have_spectra = false;
    for (size_t i = 0; i + 7 <= c->fft_tiles.size(); i += 7)
        cache_put(entries, pre + std::to_string(c->fft_tiles[i]), kTileFields, &c->fft_tiles[i + 1], 6);
    bool tubes = !c->inst.empty();
    for (const wt::Instance &I : c->inst) tubes = tubes && I.have_tubes && I.have_params;
    if (tubes && !batch_has_filters(c)) {
        int rc = deposit_autotune(c, pre, entries, rep);
        if (rc) return rc;
        c->have_omega = c->have_u = false;
    }
    if (cache_path && !write_tile_cache(cache_path, entries))
        return set_err(c, WT_EINVAL, "cannot write the autotune cache %s", cache_path);
    if (report && report_cap > 0) snprintf(report, (size_t)report_cap, "%s", rep.c_str());
    return WT_OK;
}

int wt_set_graph(wt_ctx *c, int32_t enable) {
    CHECK_CTX(c);
    if (enable < 0 || enable > 2) return set_err(c, WT_EINVAL, "graph mode must be 0, 1 or 2");
    c->use_graph = enable == 1;
    c->gen_graph_mode = enable == 2;
    if (!c->gen_graph_mode) drop_gen_graph(c);
    return WT_OK;
}

// Reads the instances' scalars after a spectral stage and checks that every velocity could be
// normalised.
static int finish_spectral(wt_ctx *c, double *etot, double *up, double *enstrophy) {
    std::vector<wt::DeviceScalars> all(c->ninst());
    HIPCALL(c, hipMemcpyAsync(all.data(), c->scal, sizeof(wt::DeviceScalars) * c->ninst(), hipMemcpyDeviceToHost,
                              c->stream), "scalars");
    HIPCALL(c, hipStreamSynchronize(c->stream), "biot_savart");
    for (int p = 0; p < c->ninst(); p++)
        if (!(all[p].up > 0) || !std::isfinite(all[p].up))
            return set_err(c, WT_ERANGE, "instance %d: velocity is zero, cannot normalise (up=%g)", c->order[p],
                           all[p].up);
    const wt::DeviceScalars &s = all[pos_of(c, c->cur)];
    if (etot) *etot = s.etot_raw;
    if (up) *up = s.up;
    if (enstrophy) *enstrophy = s.enstrophy;
    c->have_u = true;
    c->have_spectra = true;
    return WT_OK;
}

int wt_biot_savart(wt_ctx *c, double *etot, double *up, double *enstrophy) {
    CHECK_CTX(c);
    if (!c->have_omega) return set_err(c, WT_ESTATE, "wt_biot_savart needs a deposited or set vorticity");
    hipError_t e;
    if (c->use_graph && c->spec_exec) {
        e = hipGraphLaunch(c->spec_exec, c->stream);
    } else {
        e = wt::launch_spectral(c, 0, -1);
        if (e == hipSuccess && c->use_graph) e = capture_spectral(c);   // replayed from the next call on
    }
    if (e != hipSuccess) return hip_err(c, e, "biot_savart");
    return finish_spectral(c, etot, up, enstrophy);
}

// Records geometry, binning, deposition and the spectral stage of the current tubes as one graph,
// on the buffers the eager run just laid out (same tubes, same sizes).
static hipError_t capture_generate(wt_ctx *c) {
    WT_HIP(wt::spectral_reserve(c));
    WT_HIP(hipStreamSynchronize(c->stream));
    WT_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    hipError_t e = wt::replay_segments(c);
    if (e == hipSuccess) e = wt::replay_bins(c);
    if (e == hipSuccess) e = wt::replay_deposit(c);
    if (e == hipSuccess) e = wt::launch_spectral(c, 0, -1);
    c->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return e;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return e;
    }
    c->gen_graph = g;
    c->gen_exec = x;
    return hipSuccess;
}

int wt_instance_scalars(wt_ctx *c, double *etot, double *up, double *enstrophy) {
    CHECK_CTX(c);
    if (!c->have_u) return set_err(c, WT_ESTATE, "no velocity resident");
    wt::DeviceScalars s;
    int rc = read_scalars(c, c->cur, &s);
    if (rc) return rc;
    if (etot) *etot = s.etot_raw;
    if (up) *up = s.up;
    if (enstrophy) *enstrophy = s.enstrophy;
    return WT_OK;
}

int wt_generate(wt_ctx *c, double *etot, double *up, double *enstrophy) {
    CHECK_CTX(c);
    if (c->tubes_dirty) {  // the upload is not part of the timed stages
        drop_gen_graph(c);
        int rc = upload_tubes(c);
        if (rc) return rc;
    }
    if (c->gen_graph_mode && c->gen_exec) {   // replay: every stage of these tubes in one graph
        HIPCALL(c, hipEventRecord(c->ev[0], c->stream), "event");
        HIPCALL(c, hipGraphLaunch(c->gen_exec, c->stream), "generate graph");
        HIPCALL(c, hipEventRecord(c->ev[4], c->stream), "event");
        HIPCALL(c, hipEventSynchronize(c->ev[4]), "generate");
        float ms = 0;
        HIPCALL(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[4]), "event");
        c->stage_ms[0] = ms;
        c->stage_ms[1] = c->stage_ms[2] = c->stage_ms[3] = -1.0;
        c->have_segments = c->have_bins = c->have_omega = true;
        return finish_spectral(c, etot, up, enstrophy);
    }
    HIPCALL(c, hipEventRecord(c->ev[0], c->stream), "event");
    int rc = wt_build_segments(c, nullptr);
    if (rc) return rc;
    HIPCALL(c, hipEventRecord(c->ev[1], c->stream), "event");
    if (filter_tubes(c).empty() && (rc = wt_bin(c))) return rc;   // a filtered deposit bins per segment
    HIPCALL(c, hipEventRecord(c->ev[2], c->stream), "event");
    if ((rc = wt_deposit(c, nullptr))) return rc;
    HIPCALL(c, hipEventRecord(c->ev[3], c->stream), "event");
    if ((rc = wt_biot_savart(c, etot, up, enstrophy))) return rc;
    HIPCALL(c, hipEventRecord(c->ev[4], c->stream), "event");
    HIPCALL(c, hipEventSynchronize(c->ev[4]), "generate");
    for (int i = 0; i < 4; i++) {
        float ms = 0;
        HIPCALL(c, hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]), "event");
        c->stage_ms[i] = ms;
    }
    if (c->gen_graph_mode && filter_tubes(c).empty()) {
        hipError_t e = capture_generate(c);
        if (e != hipSuccess) return hip_err(c, e, "generate graph capture");
    }
    return WT_OK;
}

int wt_stage_times(wt_ctx *c, double *ms, int32_t cap) {
    CHECK_CTX(c);
    if (!ms) return set_err(c, WT_EINVAL, "ms is NULL");
    const int total = 4 + wt::kNumSlots;
    HIPCALL(c, hipStreamSynchronize(c->stream), "stage times");
    for (int i = 0; i < total && i < cap; i++) {
        if (i < 4) { ms[i] = c->stage_ms[i]; continue; }
        int s = i - 4;
        float v = -1.0f;
        if (c->kvalid[s] && hipEventElapsedTime(&v, c->kev[s][0], c->kev[s][1]) != hipSuccess) v = -1.0f;
        ms[i] = v;
    }
    return total;
}

int wt_spectrum(wt_ctx *c, int32_t field, double *E, int32_t cap) {
    CHECK_CTX(c);
    const wt::Instance &I = c->inst[c->cur];
    if (!E) return set_err(c, WT_EINVAL, "E is NULL");
    if (cap < I.nek) return set_err(c, WT_ERANGE, "need room for %d shells", I.nek);
    if (field == 0) {
        if (!c->have_spectra) return set_err(c, WT_ESTATE, "the vorticity spectrum comes from wt_biot_savart");
        HIPCALL(c, hipMemcpyAsync(E, c->E_w + I.E_off, sizeof(double) * I.nek, hipMemcpyDeviceToHost, c->stream),
                "spectrum");
    } else if (field == 1) {
        if (!c->have_u) return set_err(c, WT_ESTATE, "no velocity resident");
        if (!c->have_spectra) {
            hipError_t e = wt::launch_spectral(c, 1, c->cur);
            if (e != hipSuccess) return hip_err(c, e, "spectrum");
        }
        HIPCALL(c, hipMemcpyAsync(E, c->E_u + I.E_off, sizeof(double) * I.nek, hipMemcpyDeviceToHost, c->stream),
                "spectrum");
    } else {
        return set_err(c, WT_EINVAL, "field must be 0 (vorticity) or 1 (velocity)");
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "spectrum");
    return I.nek;
}

int wt_structure_functions(wt_ctx *c, int32_t mode, int32_t pmin, int32_t pmax, int32_t nsep, const int32_t *seps,
                           double *S) {
    CHECK_CTX(c);
    if (!c->have_u) return set_err(c, WT_ESTATE, "no velocity resident");
    if (mode != 0 && mode != 1) return set_err(c, WT_EINVAL, "mode must be 0 or 1");
    if (mode == 0 && pmin != 1)
        return set_err(c, WT_EINVAL, "mode 0 (Sp_calculate) computes orders 1..pmax; pmin must be 1");
    if (pmin < 1 || pmax < pmin || pmax - pmin + 1 > 12)
        return set_err(c, WT_EINVAL, "need 1 <= pmin <= pmax, at most 12 orders");
    if (nsep <= 0 || !seps || !S) return set_err(c, WT_EINVAL, "need nsep > 0 and non-NULL arrays");
    int32_t *d_seps = nullptr;
    double *d_out = nullptr;
    HIPCALL(c, hipMalloc(&d_seps, sizeof(int32_t) * nsep), "alloc");
    hipError_t e = hipMalloc(&d_out, sizeof(double) * nsep * (pmax - pmin + 1));
    if (e == hipSuccess) e = hipMemcpyAsync(d_seps, seps, sizeof(int32_t) * nsep, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = wt::launch_sf(c, mode, pmin, pmax, nsep, d_seps, d_out);
    if (e == hipSuccess)
        e = hipMemcpyAsync(S, d_out, sizeof(double) * nsep * (pmax - pmin + 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_seps);
    if (d_out) (void)hipFree(d_out);
    if (e == hipErrorInvalidValue) return set_err(c, WT_ERANGE, "too many separations/orders for one block's LDS");
    if (e != hipSuccess) return hip_err(c, e, "structure functions");
    return WT_OK;
}

int wt_structure_functions_noalloc(wt_ctx *c, int32_t mode, int32_t pmin, int32_t pmax, int32_t nsep,
                                   const int32_t *seps_dev, double *S_dev) {
    CHECK_CTX(c);
    if (!c->have_u) return set_err(c, WT_ESTATE, "no velocity resident");
    if (mode != 0 && mode != 1) return set_err(c, WT_EINVAL, "mode must be 0 or 1");
    if (mode == 0 && pmin != 1)
        return set_err(c, WT_EINVAL, "mode 0 (Sp_calculate) computes orders 1..pmax; pmin must be 1");
    if (pmin < 1 || pmax < pmin || pmax - pmin + 1 > 12)
        return set_err(c, WT_EINVAL, "need 1 <= pmin <= pmax, at most 12 orders");
    if (nsep <= 0 || !seps_dev || !S_dev) return set_err(c, WT_EINVAL, "need nsep > 0 and non-NULL device arrays");
    const int64_t cols = (int64_t)nsep * (pmax - pmin + 1), maxc = wt::sf_noalloc_max_cols(c, mode);
    if (cols > maxc)
        return set_err(c, WT_ERANGE, "nsep * orders = %lld exceeds the %lld the fixed scratch holds at this grid size",
                       (long long)cols, (long long)maxc);
    hipError_t e = wt::launch_sf_noalloc(c, mode, pmin, pmax, nsep, seps_dev, S_dev);
    if (e == hipErrorInvalidValue) return set_err(c, WT_ERANGE, "too many separations/orders for one block's LDS");
    if (e != hipSuccess) return hip_err(c, e, "structure functions");
    return WT_OK;
}

int wt_batch_statistics(wt_ctx *c, int32_t pmin, int32_t pmax, int32_t nsep, const int32_t *seps, double *E,
                        double *S) {
    CHECK_CTX(c);
    if (!c->have_u) return set_err(c, WT_ESTATE, "no velocity resident");
    if (pmin < 1 || pmax < pmin || pmax - pmin + 1 > 12)
        return set_err(c, WT_EINVAL, "need 1 <= pmin <= pmax, at most 12 orders");
    if (nsep < 0 || (nsep > 0 && (!seps || !S)))
        return set_err(c, WT_EINVAL, "need nsep >= 0 and, when nsep > 0, separations and an output array");
    const int ni = c->ninst(), np = pmax - pmin + 1;
    if (E) {
        if (!c->have_spectra) {
            for (int i = 0; i < ni; i++) {
                hipError_t e = wt::launch_spectral(c, 1, i);
                if (e != hipSuccess) return hip_err(c, e, "spectrum");
            }
            c->have_spectra = true;
        }
        std::vector<double> all(c->E_elems);
        HIPCALL(c, hipMemcpyAsync(all.data(), c->E_u, sizeof(double) * c->E_elems, hipMemcpyDeviceToHost, c->stream),
                "spectra");
        HIPCALL(c, hipStreamSynchronize(c->stream), "spectra");
        int64_t o = 0;
        for (int i = 0; i < ni; i++) {
            const wt::Instance &I = c->inst[i];
            memcpy(E + o, all.data() + I.E_off, sizeof(double) * I.nek);
            o += I.nek;
        }
    }
    if (nsep == 0) return WT_OK;
    std::vector<int32_t> sp((size_t)ni * nsep);
    for (int pos = 0; pos < ni; pos++)
        memcpy(sp.data() + (size_t)pos * nsep, seps + (size_t)c->order[pos] * nsep, sizeof(int32_t) * nsep);
    int32_t *d_seps = nullptr;
    double *d_out = nullptr;
    HIPCALL(c, hipMallocAsync((void **)&d_seps, sizeof(int32_t) * sp.size(), c->stream), "alloc");
    hipError_t e = hipMallocAsync((void **)&d_out, sizeof(double) * (size_t)ni * np * nsep, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_seps, sp.data(), sizeof(int32_t) * sp.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = wt::launch_sf_batch(c, pmin, pmax, nsep, d_seps, d_out);
    std::vector<double> out((size_t)ni * np * nsep);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, c->stream);
    if (d_seps) (void)hipFreeAsync(d_seps, c->stream);
    if (d_out) (void)hipFreeAsync(d_out, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipErrorInvalidValue) return set_err(c, WT_ERANGE, "too many separations/orders for one block's LDS");
    if (e != hipSuccess) return hip_err(c, e, "batch statistics");
    for (int pos = 0; pos < ni; pos++)
        memcpy(S + (size_t)c->order[pos] * np * nsep, out.data() + (size_t)pos * np * nsep, sizeof(double) * np * nsep);
    return WT_OK;
}

int wt_pdf(wt_ctx *c, int32_t field, int32_t nbins, float *val, float *pdf, int64_t *counts) {
    CHECK_CTX(c);
    if (field != 0 && field != 1) return set_err(c, WT_EINVAL, "field must be 0 (vorticity) or 1 (velocity)");
    if (nbins < 1 || nbins > 8192 || !val || !pdf) return set_err(c, WT_EINVAL, "need 1 <= nbins <= 8192 and outputs");
    if (field == 1 ? !c->have_u : !c->have_omega) return set_err(c, WT_ESTATE, "the field is not resident");
    hipError_t e = wt::launch_pdf(c, field, nbins, val, pdf, counts);
    if (e != hipSuccess) return hip_err(c, e, "pdf");
    return WT_OK;
}

static void *field_ptr(wt_ctx *c, int which) {
    const wt::Instance &I = c->inst[c->cur];
    size_t N3 = (size_t)I.N * I.N * I.N, es = c->precision == 64 ? 8 : 4;
    char *base = (char *)(which < 3 ? c->omega : c->u);
    return base + ((size_t)I.field_off + (size_t)(which % 3) * N3) * es;
}

int wt_get_field(wt_ctx *c, int32_t which, void *host, int32_t host_is_fp64) {
    CHECK_CTX(c);
    if (which < 0 || which > 5 || !host) return set_err(c, WT_EINVAL, "bad field selector or NULL buffer");
    const wt::Instance &I = c->inst[c->cur];
    size_t N3 = (size_t)I.N * I.N * I.N;
    bool dev64 = c->precision == 64;
    if ((bool)host_is_fp64 == dev64) {
        HIPCALL(c, hipMemcpyAsync(host, field_ptr(c, which), N3 * (dev64 ? 8 : 4), hipMemcpyDeviceToHost, c->stream),
                "copy");
        HIPCALL(c, hipStreamSynchronize(c->stream), "copy");
        return WT_OK;
    }
    std::vector<unsigned char> tmp(N3 * (dev64 ? 8 : 4));
    HIPCALL(c, hipMemcpyAsync(tmp.data(), field_ptr(c, which), tmp.size(), hipMemcpyDeviceToHost, c->stream), "copy");
    HIPCALL(c, hipStreamSynchronize(c->stream), "copy");
    if (dev64) for (size_t i = 0; i < N3; i++) ((float *)host)[i] = (float)((double *)tmp.data())[i];
    else for (size_t i = 0; i < N3; i++) ((double *)host)[i] = (double)((float *)tmp.data())[i];
    return WT_OK;
}

int wt_set_field(wt_ctx *c, int32_t which, const void *host, int32_t host_is_fp64) {
    CHECK_CTX(c);
    if (which < 0 || which > 5 || !host) return set_err(c, WT_EINVAL, "bad field selector or NULL buffer");
    const wt::Instance &I = c->inst[c->cur];
    size_t N3 = (size_t)I.N * I.N * I.N;
    bool dev64 = c->precision == 64;
    if ((bool)host_is_fp64 == dev64) {
        HIPCALL(c, hipMemcpyAsync(field_ptr(c, which), host, N3 * (dev64 ? 8 : 4), hipMemcpyHostToDevice, c->stream),
                "copy");
    } else {
        std::vector<unsigned char> tmp(N3 * (dev64 ? 8 : 4));
        if (dev64) for (size_t i = 0; i < N3; i++) ((double *)tmp.data())[i] = (double)((const float *)host)[i];
        else for (size_t i = 0; i < N3; i++) ((float *)tmp.data())[i] = (float)((const double *)host)[i];
        HIPCALL(c, hipMemcpyAsync(field_ptr(c, which), tmp.data(), tmp.size(), hipMemcpyHostToDevice, c->stream), "copy");
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "copy");
    if (which < 3) { c->have_omega = true; c->have_u = false; }
    else c->have_u = true;
    c->have_spectra = false;
    return WT_OK;
}

int wt_get_sections(wt_ctx *c, int32_t *ntf, double *dtf) {
    CHECK_CTX(c);
    if (!c->have_segments) return set_err(c, WT_ESTATE, "no segments built");
    if (ntf)
        HIPCALL(c, hipMemcpyAsync(ntf, c->geo.ntf, sizeof(int32_t) * c->geo.nsec_all, hipMemcpyDeviceToHost, c->stream),
                "copy");
    if (dtf)
        HIPCALL(c, hipMemcpyAsync(dtf, c->geo.dtf, sizeof(double) * c->geo.nsec_all, hipMemcpyDeviceToHost, c->stream),
                "copy");
    HIPCALL(c, hipStreamSynchronize(c->stream), "copy");
    return WT_OK;
}

int32_t wt_fbb_embedding(int32_t npoint) {
    if (npoint < 3) return WT_EINVAL;
    const int64_t m = wt::fbb_embedding(npoint);
    return m > ((int64_t)1 << 20) ? WT_ERANGE : (int32_t)m;
}

int wt_fbb_centerlines(wt_ctx *c, int32_t ntube, const int32_t *npoint, const double *hurst, const int32_t *winding,
                       double wiggle, uint64_t seed, const uint64_t *keys, const double *gauss, double *cx, double *cy,
                       double *cz) {
    CHECK_CTX(c);
    if (ntube <= 0 || !npoint || !hurst || !winding || !cx || !cy || !cz)
        return set_err(c, WT_EINVAL, "need ntube > 0 and non-NULL arrays");
    if (!(wiggle >= 0.0) || !std::isfinite(wiggle)) return set_err(c, WT_EINVAL, "wiggle must be finite and >= 0");
    for (int l = 0; l < ntube; l++) {
        if (npoint[l] < 3) return set_err(c, WT_EINVAL, "tube %d has %d points (at least 3)", l, npoint[l]);
        if (wt::fbb_embedding(npoint[l]) > ((int64_t)1 << 20))
            return set_err(c, WT_ERANGE, "tube %d: %d points exceed the generator's 2^19", l, npoint[l]);
        if (!(hurst[l] > 0.0 && hurst[l] < 1.0)) return set_err(c, WT_EINVAL, "tube %d: Hurst index outside (0, 1)", l);
        if (!winding[3 * l] && !winding[3 * l + 1] && !winding[3 * l + 2])
            return set_err(c, WT_EINVAL, "tube %d: zero winding vector", l);
    }
    hipError_t e = wt::fbb_centerlines(c, ntube, npoint, hurst, winding, wiggle, nullptr, seed, keys, gauss, cx, cy, cz);
    if (e != hipSuccess) return hip_err(c, e, "fbb centerlines");
    return WT_OK;
}

int wt_fbb_reference_centerlines(wt_ctx *c, int32_t ntube, const int32_t *npoint, const double *hurst,
                                 const double *dB, uint64_t seed, const uint64_t *keys, const double *gauss, double *cx,
                                 double *cy, double *cz) {
    CHECK_CTX(c);
    if (ntube <= 0 || !npoint || !hurst || !dB || !cx || !cy || !cz)
        return set_err(c, WT_EINVAL, "need ntube > 0 and non-NULL arrays");
    for (int l = 0; l < ntube; l++) {
        if (npoint[l] < 3) return set_err(c, WT_EINVAL, "tube %d has %d points (at least 3)", l, npoint[l]);
        if (wt::fbb_embedding(wt::fbb_generated_points(npoint[l], true)) > ((int64_t)1 << 20))
            return set_err(c, WT_ERANGE, "tube %d: %d points exceed the generator's 2^19", l, npoint[l]);
        if (!(hurst[l] > 0.0 && hurst[l] < 1.0)) return set_err(c, WT_EINVAL, "tube %d: Hurst index outside (0, 1)", l);
        if (!(dB[l] >= 0.0) || !std::isfinite(dB[l])) return set_err(c, WT_EINVAL, "tube %d: dB must be finite and >= 0", l);
    }
    hipError_t e = wt::fbb_centerlines(c, ntube, npoint, hurst, nullptr, 0.0, dB, seed, keys, gauss, cx, cy, cz);
    if (e != hipSuccess) return hip_err(c, e, "fbb reference centerlines");
    return WT_OK;
}

int wt_fbb_normals(wt_ctx *c, uint64_t seed, int32_t tube, int64_t count, double *out) {
    CHECK_CTX(c);
    if (tube < 0 || count <= 0 || !out) return set_err(c, WT_EINVAL, "need tube >= 0, count > 0 and an output array");
    hipError_t e = wt::fbb_normals(c, seed, (uint64_t)tube, count, out);
    if (e != hipSuccess) return hip_err(c, e, "fbb normals");
    return WT_OK;
}

int wt_get_tube_lengths(wt_ctx *c, double *length) {
    CHECK_CTX(c);
    if (!c->have_segments) return set_err(c, WT_ESTATE, "no segments built");
    if (!length) return set_err(c, WT_EINVAL, "length is NULL");
    const wt::Instance &I = c->inst[c->cur];
    HIPCALL(c, hipMemcpyAsync(length, c->geo.length + I.tube0, sizeof(double) * I.ntube, hipMemcpyDeviceToHost,
                              c->stream), "copy");
    HIPCALL(c, hipStreamSynchronize(c->stream), "copy");
    return WT_OK;
}

int wt_get_subpoints(wt_ctx *c, double *ci_xyz, double *s) {
    CHECK_CTX(c);
    if (!c->have_segments) return set_err(c, WT_ESTATE, "no segments built");
    const int64_t NSP = c->geo.nsub_all, S = c->geo.nsec_all;
    if (ci_xyz)
        HIPCALL(c, hipMemcpyAsync(ci_xyz, c->geo.ci, sizeof(double) * 3 * NSP, hipMemcpyDeviceToHost, c->stream), "copy");
    if (s) {
        std::vector<double> czl(S), dtf(S);
        std::vector<int64_t> spb(S + 1);
        HIPCALL(c, hipMemcpyAsync(czl.data(), c->geo.czl, sizeof(double) * S, hipMemcpyDeviceToHost, c->stream), "copy");
        HIPCALL(c, hipMemcpyAsync(dtf.data(), c->geo.dtf, sizeof(double) * S, hipMemcpyDeviceToHost, c->stream), "copy");
        HIPCALL(c, hipMemcpyAsync(spb.data(), c->geo.sp_base, sizeof(int64_t) * (S + 1), hipMemcpyDeviceToHost,
                                  c->stream), "copy");
        HIPCALL(c, hipStreamSynchronize(c->stream), "copy");
        for (int64_t sec = 0; sec < S; sec++)  // s_list2 = cze_pl(t_i) + dtf(t_i)*(t_j-1) (tube.f90:708)
            for (int64_t q = spb[sec]; q < spb[sec + 1]; q++) s[q] = czl[sec] + dtf[sec] * (double)(q - spb[sec]);
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "copy");
    return WT_OK;
}

int wt_get_bins(wt_ctx *c, int32_t *lohi, double *R) {
    CHECK_CTX(c);
    if (!c->have_segments) return set_err(c, WT_ESTATE, "no segments built");
    if (!lohi || !R) return set_err(c, WT_EINVAL, "NULL buffer");
    const int64_t NSP = c->geo.nsub_all;
    int32_t *d_lohi = nullptr;
    double *d_R = nullptr;
    HIPCALL(c, hipMalloc(&d_lohi, sizeof(int32_t) * 6 * NSP), "alloc");
    hipError_t e = hipMalloc(&d_R, sizeof(double) * NSP);
    if (e == hipSuccess) e = wt::launch_get_bins(c, d_lohi, d_R);
    if (e == hipSuccess) e = hipMemcpyAsync(lohi, d_lohi, sizeof(int32_t) * 6 * NSP, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(R, d_R, sizeof(double) * NSP, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_lohi);
    if (d_R) (void)hipFree(d_R);
    if (e != hipSuccess) return hip_err(c, e, "bins");
    return WT_OK;
}

}  // extern "C"
