// C ABI of the woven-turbulence generator (declarations and reference citations in
// include/wovengpu.h). Owns device memory, the HIP stream and the call order checks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "wt_internal.h"

namespace wt {
hipError_t launch_seg_stats(wt_ctx *c, wt_seg_stats *out);
hipError_t launch_spectrum_of_u(wt_ctx *c);
bool fft_size_supported(int N);
void fft_reversal(int N, std::vector<int32_t> &rev);
int64_t nek_of(int N) {
    float v = (float)(3 * N * N) / 12.0f;  // get_spectrum: int(sqrt((nx^2+ny^2+nz^2)/12.0))
    return (int64_t)std::sqrt(v);
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
    return set_err(c, WT_EDEVICE, "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
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

static size_t field_bytes(const wt_ctx *c) {
    size_t N = (size_t)c->N, N3 = N * N * N;
    size_t es = c->precision == 64 ? 8 : 4;
    size_t b = 2 * 3 * N3 * es + 3 * N * N * (N / 2 + 1) * 2 * es;
    if (c->precision == 32) b += 3 * N3 * 8;
    return b;
}

template <typename T>
static hipError_t alloc(T **p, size_t n) {
    return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
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

int wt_create(wt_ctx **out, int32_t n, int32_t precision, int32_t device) {
    if (!out) return set_err(nullptr, WT_EINVAL, "wt_create: ctx pointer is NULL");
    *out = nullptr;
    if (precision != 32 && precision != 64) return set_err(nullptr, WT_EINVAL, "precision must be 32 or 64");
    if (!wt::fft_size_supported(n))
        return set_err(nullptr, WT_ERANGE, "N=%d unsupported: need even 4 <= N <= 1024 with factors 2 and 3 only", n);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return set_err(nullptr, WT_EDEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return set_err(nullptr, WT_EINVAL, "device %d out of range (%d)", device, ndev);
    wt_ctx *c = new (std::nothrow) wt_ctx();
    if (!c) return set_err(nullptr, WT_ENOMEM, "host allocation failed");
    c->device = device;
    c->precision = precision;
    c->N = n;
    size_t cap = mem_cap_bytes();
    if (cap && field_bytes(c) > cap) {
        delete c;
        return set_err(nullptr, WT_ENOMEM, "fields need %zu bytes, above WT_DEVICE_MEM_CAP=%zu", field_bytes(c), cap);
    }
    auto fail = [&](hipError_t er, const char *w) {
        set_err(nullptr, er == hipErrorOutOfMemory ? WT_ENOMEM : WT_EDEVICE, "%s: %s", w, hipGetErrorString(er));
        wt_destroy(c);
        return er == hipErrorOutOfMemory ? WT_ENOMEM : WT_EDEVICE;
    };
    if ((e = hipSetDevice(device)) != hipSuccess) return fail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    for (auto &ev : c->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(e, "event");
    for (int s = 0; s < wt::kNumSlots; s++)
        for (int k = 0; k < 2; k++)
            if ((e = hipEventCreate(&c->kev[s][k])) != hipSuccess) return fail(e, "event");
    const size_t N = n, N3 = N * N * N, es = precision == 64 ? 8 : 4;
    if ((e = hipMalloc(&c->omega, 3 * N3 * es)) != hipSuccess) return fail(e, "omega");
    if ((e = hipMalloc(&c->u, 3 * N3 * es)) != hipSuccess) return fail(e, "u");
    if ((e = hipMalloc(&c->spec, 3 * N * N * (N / 2 + 1) * 2 * es)) != hipSuccess) return fail(e, "spectrum buffer");
    if (precision == 32 && (e = hipMalloc((void **)&c->acc, 3 * N3 * sizeof(long long))) != hipSuccess)
        return fail(e, "accumulator");
    // twiddles exp(-2 pi i k / N) in the compute precision, then the digit-reversal table
    std::vector<int32_t> rev;
    wt::fft_reversal(n, rev);
    size_t twb = 2 * N * es + N * sizeof(int32_t);
    std::vector<unsigned char> host(twb);
    for (size_t k = 0; k < N; k++) {
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
    memcpy(host.data() + 2 * N * es, rev.data(), N * sizeof(int32_t));
    if ((e = hipMalloc(&c->twiddle, twb)) != hipSuccess) return fail(e, "twiddles");
    if ((e = hipMemcpy(c->twiddle, host.data(), twb, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "twiddles");
    if ((e = hipMalloc((void **)&c->scal, sizeof(wt::DeviceScalars))) != hipSuccess) return fail(e, "scalars");
    if ((e = hipMemset(c->scal, 0, sizeof(wt::DeviceScalars))) != hipSuccess) return fail(e, "scalars");
    c->nek = (int)wt::nek_of(n);
    if ((e = hipMalloc((void **)&c->E_w, sizeof(double) * (c->nek + 1))) != hipSuccess) return fail(e, "spectra");
    if ((e = hipMalloc((void **)&c->E_u, sizeof(double) * (c->nek + 1))) != hipSuccess) return fail(e, "spectra");
    if ((e = hipMemset(c->omega, 0, 3 * N3 * es)) != hipSuccess) return fail(e, "omega");
    if ((e = hipMemset(c->u, 0, 3 * N3 * es)) != hipSuccess) return fail(e, "u");
    {
        const char *te = getenv("WT_THETA_EXACT");
        c->theta_exact_all = te && te[0] == '1';
    }
    c->run.N = n;
    c->run.dx = wt::kPi2 / n;  // lx/nx (tube.f90:274)
    c->run.start = -wt::kPi;
    *out = c;
    return WT_OK;
}

static void free_list(std::vector<void *> &l) {
    for (void *p : l) (void)hipFree(p);
    l.clear();
}

void wt_destroy(wt_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_list(c->bin_allocs);
    free_list(c->seg_allocs);
    free_list(c->geo_allocs);
    void *bufs[] = {c->omega, c->u, c->spec, c->acc, c->partials, c->twiddle, c->scal, c->E_w, c->E_u};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &pair : c->kev)
        for (auto &ev : pair)
            if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int wt_set_params(wt_ctx *c, const wt_params *p) {
    CHECK_CTX(c);
    if (!p) return set_err(c, WT_EINVAL, "params is NULL");
    if (p->N != c->N) return set_err(c, WT_EINVAL, "params N=%d differs from the context N=%d", p->N, c->N);
    if (p->ntf_times < 1 || p->ntf_pre < 1) return set_err(c, WT_EINVAL, "ntf_times and ntf_pre must be >= 1");
    if (!(p->Rtube_ratio > 0) || !(p->dzeta_sup > 0) || !(p->sigma_ratio > 0) || !std::isfinite(p->eta_t))
        return set_err(c, WT_EINVAL, "Rtube_ratio, dzeta_sup and sigma_ratio must be positive, eta_t finite");
    if (p->line_s < 1 || p->line_e < p->line_s) return set_err(c, WT_EINVAL, "need 1 <= line_s <= line_e");
    c->params = *p;
    c->run.line_s = p->line_s - 1;
    c->run.line_e = p->line_e - 1;
    c->run.sigma_ratio = p->sigma_ratio;
    c->run.eta = p->eta_t;
    c->run.Rtube_ratio = p->Rtube_ratio;
    c->run.dzeta_sup = p->dzeta_sup;
    c->run.ntf_times = p->ntf_times;
    c->run.ntf_pre = p->ntf_pre;
    c->have_params = true;
    c->have_segments = c->have_bins = false;
    return WT_OK;
}

int wt_set_tubes(wt_ctx *c, int32_t nline, const int32_t *npoint, const double *sigma, const double *gamma,
                 const int32_t *n_T, const int32_t *if_filter, const double *cx, const double *cy, const double *cz) {
    CHECK_CTX(c);
    if (nline <= 0 || !npoint || !sigma || !gamma || !n_T || !cx || !cy || !cz)
        return set_err(c, WT_EINVAL, "wt_set_tubes: need nline > 0 and non-NULL arrays");
    std::vector<int64_t> pt(nline + 1), sec(nline + 1);
    pt[0] = sec[0] = 0;
    double bound = 0.0;
    for (int l = 0; l < nline; l++) {
        if (npoint[l] < 3) return set_err(c, WT_EINVAL, "tube %d has %d points; need at least 3", l + 1, npoint[l]);
        if (!(sigma[l] > 0) || !std::isfinite(sigma[l]) || !std::isfinite(gamma[l]))
            return set_err(c, WT_EINVAL, "tube %d: sigma must be positive and finite, Gamma finite", l + 1);
        if (if_filter && if_filter[l] != 0)
            return set_err(c, WT_EINVAL, "tube %d: if_filter=1 (filter_spec) is not supported", l + 1);
        pt[l + 1] = pt[l] + npoint[l];
        sec[l + 1] = sec[l] + npoint[l] + 1;
        bound += std::fabs(gamma[l]) / (wt::kPi2 * sigma[l] * sigma[l]);
    }
    const int64_t P = pt[nline], S = sec[nline];
    for (int64_t i = 0; i < P; i++)
        if (!std::isfinite(cx[i]) || !std::isfinite(cy[i]) || !std::isfinite(cz[i]))
            return set_err(c, WT_EINVAL, "non-finite centerline coordinate at point %lld", (long long)i + 1);
    size_t cap = mem_cap_bytes();
    size_t need = (size_t)P * 8 * 13 + (size_t)S * 8 * 24;
    if (cap && field_bytes(c) + need > cap)
        return set_err(c, WT_ENOMEM, "tube data needs %zu bytes beyond the fields, above WT_DEVICE_MEM_CAP", need);
    free_list(c->bin_allocs);
    free_list(c->seg_allocs);
    free_list(c->geo_allocs);
    c->have_tubes = c->have_segments = c->have_bins = false;
    wt::Geo g{};
    g.nline = nline;
    g.npoint_all = P;
    g.nsec_all = S;
    auto A = [&](auto **p, size_t n) -> hipError_t {
        hipError_t e = alloc(p, n);
        if (e == hipSuccess) c->geo_allocs.push_back((void *)*p);
        return e;
    };
    int64_t *pt_base, *sec_base, *sp_base;
    double *sig, *gam, *xs, *ys, *zs;
    int32_t *nT;
    HIPCALL(c, A(&pt_base, nline + 1), "alloc");
    HIPCALL(c, A(&sec_base, nline + 1), "alloc");
    HIPCALL(c, A(&sig, nline), "alloc");
    HIPCALL(c, A(&gam, nline), "alloc");
    HIPCALL(c, A(&nT, nline), "alloc");
    HIPCALL(c, A(&g.Len_Disc, nline), "alloc");
    HIPCALL(c, A(&g.length, nline), "alloc");
    HIPCALL(c, A(&xs, P), "alloc");
    HIPCALL(c, A(&ys, P), "alloc");
    HIPCALL(c, A(&zs, P), "alloc");
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
    HIPCALL(c, hipMemcpyAsync(pt_base, pt.data(), sizeof(int64_t) * (nline + 1), hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(sec_base, sec.data(), sizeof(int64_t) * (nline + 1), hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(sig, sigma, sizeof(double) * nline, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(gam, gamma, sizeof(double) * nline, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(nT, n_T, sizeof(int32_t) * nline, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(xs, cx, sizeof(double) * P, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(ys, cy, sizeof(double) * P, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipMemcpyAsync(zs, cz, sizeof(double) * P, hipMemcpyHostToDevice, st), "copy");
    HIPCALL(c, hipStreamSynchronize(st), "copy");
    g.pt_base = pt_base;
    g.sec_base = sec_base;
    g.sigma = sig;
    g.gamma = gam;
    g.n_T = nT;
    g.cx = xs;
    g.cy = ys;
    g.cz = zs;
    c->geo = g;
    c->npoint.assign(npoint, npoint + nline);
    c->pt_base = pt;
    c->sec_base = sec;
    // fixed-point scale for the FP32 accumulation: |omega| <= sum Gamma/(2 pi sigma^2) times a
    // generous factor for the curvature and twist terms
    double B = bound * 64.0 + 1.0;
    int shift = (int)std::floor(60.0 - std::log2(B));
    c->fixed_shift = shift < 8 ? 8 : (shift > 44 ? 44 : shift);
    c->have_tubes = true;
    return WT_OK;
}

int wt_build_segments(wt_ctx *c, wt_seg_stats *out) {
    CHECK_CTX(c);
    if (!c->have_params || !c->have_tubes) return set_err(c, WT_ESTATE, "wt_build_segments needs params and tubes");
    c->have_segments = c->have_bins = false;
    hipError_t e = wt::launch_segments(c);
    if (e == hipErrorInvalidValue)
        return set_err(c, WT_ERANGE, "turning angle not finite (degenerate spline: repeated or cusped points)");
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

int wt_bin(wt_ctx *c) {
    CHECK_CTX(c);
    if (!c->have_segments) return set_err(c, WT_ESTATE, "wt_bin needs wt_build_segments");
    if (c->run.line_e >= c->geo.nline)
        return set_err(c, WT_EINVAL, "line_e=%d exceeds the %d tubes", c->run.line_e + 1, c->geo.nline);
    hipError_t e = wt::launch_bins(c);
    if (e == hipErrorInvalidValue)
        return set_err(c, WT_ERANGE, "a tube's cut-off radius spans more than the periodic box");
    if (e != hipSuccess) return hip_err(c, e, "bin");
    c->have_bins = true;
    return WT_OK;
}

int wt_deposit(wt_ctx *c, wt_deposit_stats *out) {
    CHECK_CTX(c);
    if (!c->have_bins) return set_err(c, WT_ESTATE, "wt_deposit needs wt_bin");
    hipError_t e = wt::launch_deposit(c);
    if (e != hipSuccess) return hip_err(c, e, "deposit");
    if (out) *out = c->dep_stats;
    if (c->dep_stats.overflow)
        return set_err(c, WT_ERANGE, "fixed-point accumulator overflow (2^%d scale); use precision 64", c->fixed_shift);
    c->have_omega = true;
    c->have_u = false;
    c->have_spectra = false;
    return WT_OK;
}

int wt_biot_savart(wt_ctx *c, double *etot, double *up, double *enstrophy) {
    CHECK_CTX(c);
    if (!c->have_omega) return set_err(c, WT_ESTATE, "wt_biot_savart needs a deposited or set vorticity");
    hipError_t e = wt::launch_biot_savart(c);
    if (e != hipSuccess) return hip_err(c, e, "biot_savart");
    wt::DeviceScalars s;
    HIPCALL(c, hipMemcpyAsync(&s, c->scal, sizeof s, hipMemcpyDeviceToHost, c->stream), "scalars");
    HIPCALL(c, hipStreamSynchronize(c->stream), "biot_savart");
    if (!(s.up > 0) || !std::isfinite(s.up)) return set_err(c, WT_ERANGE, "velocity is zero: cannot normalise (up=%g)", s.up);
    if (etot) *etot = s.etot_raw;
    if (up) *up = s.up;
    if (enstrophy) *enstrophy = s.enstrophy;
    c->have_u = true;
    c->have_spectra = true;
    return WT_OK;
}

int wt_generate(wt_ctx *c, double *etot, double *up, double *enstrophy) {
    CHECK_CTX(c);
    HIPCALL(c, hipEventRecord(c->ev[0], c->stream), "event");
    int rc = wt_build_segments(c, nullptr);
    if (rc) return rc;
    HIPCALL(c, hipEventRecord(c->ev[1], c->stream), "event");
    if ((rc = wt_bin(c))) return rc;
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
    if (!E) return set_err(c, WT_EINVAL, "E is NULL");
    if (cap < c->nek) return set_err(c, WT_ERANGE, "need room for %d shells", c->nek);
    if (field == 0) {
        if (!c->have_spectra) return set_err(c, WT_ESTATE, "the vorticity spectrum comes from wt_biot_savart");
        HIPCALL(c, hipMemcpyAsync(E, c->E_w, sizeof(double) * c->nek, hipMemcpyDeviceToHost, c->stream), "spectrum");
    } else if (field == 1) {
        if (!c->have_u) return set_err(c, WT_ESTATE, "no velocity resident");
        if (!c->have_spectra) {
            hipError_t e = wt::launch_spectrum_of_u(c);
            if (e != hipSuccess) return hip_err(c, e, "spectrum");
        }
        HIPCALL(c, hipMemcpyAsync(E, c->E_u, sizeof(double) * c->nek, hipMemcpyDeviceToHost, c->stream), "spectrum");
    } else {
        return set_err(c, WT_EINVAL, "field must be 0 (vorticity) or 1 (velocity)");
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "spectrum");
    return c->nek;
}

int wt_structure_functions(wt_ctx *c, int32_t mode, int32_t pmin, int32_t pmax, int32_t nsep, const int32_t *seps,
                           double *S) {
    CHECK_CTX(c);
    if (!c->have_u) return set_err(c, WT_ESTATE, "no velocity resident");
    if (mode != 0 && mode != 1) return set_err(c, WT_EINVAL, "mode must be 0 or 1");
    if (mode == 0 && pmin != 1) return set_err(c, WT_EINVAL, "mode 0 (Sp_calculate) computes orders 1..pmax; pmin must be 1");
    if (pmin < 1 || pmax < pmin || pmax - pmin + 1 > 12) return set_err(c, WT_EINVAL, "need 1 <= pmin <= pmax, at most 12 orders");
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

static void *field_ptr(wt_ctx *c, int which) {
    size_t N3 = (size_t)c->N * c->N * c->N, es = c->precision == 64 ? 8 : 4;
    char *base = (char *)(which < 3 ? c->omega : c->u);
    return base + (size_t)(which % 3) * N3 * es;
}

int wt_get_field(wt_ctx *c, int32_t which, void *host, int32_t host_is_fp64) {
    CHECK_CTX(c);
    if (which < 0 || which > 5 || !host) return set_err(c, WT_EINVAL, "bad field selector or NULL buffer");
    size_t N3 = (size_t)c->N * c->N * c->N;
    bool dev64 = c->precision == 64;
    if ((bool)host_is_fp64 == dev64) {
        HIPCALL(c, hipMemcpyAsync(host, field_ptr(c, which), N3 * (dev64 ? 8 : 4), hipMemcpyDeviceToHost, c->stream), "copy");
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
    size_t N3 = (size_t)c->N * c->N * c->N;
    bool dev64 = c->precision == 64;
    if ((bool)host_is_fp64 == dev64) {
        HIPCALL(c, hipMemcpyAsync(field_ptr(c, which), host, N3 * (dev64 ? 8 : 4), hipMemcpyHostToDevice, c->stream), "copy");
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
    if (ntf) HIPCALL(c, hipMemcpyAsync(ntf, c->geo.ntf, sizeof(int32_t) * c->geo.nsec_all, hipMemcpyDeviceToHost, c->stream), "copy");
    if (dtf) HIPCALL(c, hipMemcpyAsync(dtf, c->geo.dtf, sizeof(double) * c->geo.nsec_all, hipMemcpyDeviceToHost, c->stream), "copy");
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
        HIPCALL(c, hipMemcpyAsync(spb.data(), c->geo.sp_base, sizeof(int64_t) * (S + 1), hipMemcpyDeviceToHost, c->stream), "copy");
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
