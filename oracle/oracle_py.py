"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg
of bench.py, always as the checker. The product path (fastweavturb_amd) never imports it.

Arrays follow the reference layout: a field component is N^3 doubles in Fortran order
(x fastest), returned here as numpy arrays of shape (N, N, N) indexed [z, y, x].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class Params(C.Structure):
    """or_params: the global run parameters of parameters_input.dat (tube.f90:191-217)."""
    _fields_ = [
        ("N", C.c_int32), ("n_d", C.c_int32), ("line_s", C.c_int32), ("line_e", C.c_int32),
        ("sigma_ratio", C.c_double), ("eta_t", C.c_double), ("ntf_times", C.c_int32),
        ("ntf_pre", C.c_int32), ("Rtube_ratio", C.c_double), ("dzeta_sup", C.c_double),
    ]


_P64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)
_PI32 = C.POINTER(C.c_int32)


class _Geom(C.Structure):
    _fields_ = [
        ("nline", C.c_int32), ("npoint_all", C.c_int64), ("nsec_all", C.c_int64),
        ("nsp_all", C.c_int64), ("nseg_all", C.c_int64),
        ("pt_base", _P64), ("sec_base", _P64), ("seg_base", _P64),
        ("Len_Disc", _PD), ("length", _PD),
        ("cze", _PD), ("ck", _PD), ("ckk", _PD),
        ("coe", _PD), ("cze_l", _PD), ("cze_r", _PD), ("ntf", _PI32), ("dtf", _PD),
        ("theta", _PD), ("sp_base", _P64),
        ("ci", _PD), ("s", _PD), ("tci", _PD), ("ndci", _PD),
        ("seg_sec", _PI32), ("seg_j", _PI32),
    ]


class _BinRec(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3), ("R", C.c_double),
                ("ci0", C.c_double * 3), ("ci1", C.c_double * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_last_error.restype = C.c_char_p
        L.or_sigma.restype = C.c_double
        L.or_sigma.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        _lib = L
    return _lib


def threads() -> int:
    """Threads of the oracle's parallel loops (geometry per tube, deposition per z-slab, FFT lines)."""
    return int(lib().or_threads())


def _check(rc):
    if rc < 0:
        raise RuntimeError("oracle: " + lib().or_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(_PD)


def _ip(a):
    return a.ctypes.data_as(_PI32)


def make_params(N, line_s=1, line_e=None, nline=1, n_d=None, sigma_ratio=1.0, eta_t=0.0,
                ntf_times=50, ntf_pre=1000, Rtube_ratio=3.0, dzeta_sup=1e-4) -> Params:
    return Params(N, N if n_d is None else n_d, line_s, nline if line_e is None else line_e,
                  sigma_ratio, eta_t, ntf_times, ntf_pre, Rtube_ratio, dzeta_sup)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Geometry:
    """Owns an or_geom built by or_preprocess + or_segments (tube.f90:392-751)."""

    def __init__(self, params: Params, npoint, cx, cy, cz):
        self.params = params
        self.npoint = np.ascontiguousarray(npoint, dtype=np.int32)
        self.cx = np.ascontiguousarray(cx, dtype=np.float64)
        self.cy = np.ascontiguousarray(cy, dtype=np.float64)
        self.cz = np.ascontiguousarray(cz, dtype=np.float64)
        self._g = _Geom()
        L = lib()
        try:
            _check(L.or_preprocess(C.byref(params), len(self.npoint), _ip(self.npoint),
                                   _dp(self.cx), _dp(self.cy), _dp(self.cz), C.byref(self._g)))
            _check(L.or_segments(C.byref(params), _dp(self.cx), _dp(self.cy), _dp(self.cz),
                                 C.byref(self._g)))
        except Exception:
            L.or_geom_free(C.byref(self._g))
            raise
        g = self._g
        nl = g.nline
        self.nline = nl
        self.pt_base = _arr(g.pt_base, nl + 1, np.int64)
        self.sec_base = _arr(g.sec_base, nl + 1, np.int64)
        self.seg_base = _arr(g.seg_base, nl + 1, np.int64)
        self.Len_Disc = _arr(g.Len_Disc, nl, np.float64)
        self.length = _arr(g.length, nl, np.float64)
        P, S, NSP, NSEG = g.npoint_all, g.nsec_all, g.nsp_all, g.nseg_all
        self.cze = _arr(g.cze, P, np.float64)
        self.ck = _arr(g.ck, 3 * P, np.float64).reshape(P, 3)
        self.ckk = _arr(g.ckk, 3 * P, np.float64).reshape(P, 3)
        self.coe = _arr(g.coe, 18 * S, np.float64).reshape(S, 3, 6)
        self.cze_l = _arr(g.cze_l, S, np.float64)
        self.cze_r = _arr(g.cze_r, S, np.float64)
        self.ntf = _arr(g.ntf, S, np.int32)
        self.dtf = _arr(g.dtf, S, np.float64)
        self.theta = _arr(g.theta, S, np.float64)
        self.sp_base = _arr(g.sp_base, S + 1, np.int64)
        self.ci = _arr(g.ci, 3 * NSP, np.float64).reshape(NSP, 3)
        self.s = _arr(g.s, NSP, np.float64)
        self.tci = _arr(g.tci, 3 * NSP, np.float64).reshape(NSP, 3)
        self.ndci = _arr(g.ndci, NSP, np.float64)
        self.seg_sec = _arr(g.seg_sec, NSEG, np.int32)
        self.seg_j = _arr(g.seg_j, NSEG, np.int32)
        self.nseg = NSEG

    def close(self):
        if self._g is not None:
            lib().or_geom_free(C.byref(self._g))
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bins(self, sigma_line, n_T_line):
        """or_bin: per sub-segment AABB (1-based extended indices), R and unwrapped ends."""
        sig = np.ascontiguousarray(sigma_line, dtype=np.float64)
        nT = np.ascontiguousarray(n_T_line, dtype=np.int32)
        recs = (_BinRec * max(self.nseg, 1))()
        _check(lib().or_bin(C.byref(self.params), _dp(sig), _ip(nT), C.byref(self._g), recs))
        raw = np.frombuffer(recs, dtype=np.dtype([("lo", "<i4", 3), ("hi", "<i4", 3), ("R", "<f8"),
                                                   ("ci0", "<f8", 3), ("ci1", "<f8", 3)]))
        raw = raw[: self.nseg].copy()
        self._bin_recs = recs
        return raw

    def pair_hits(self, seg, ip, jp, kp) -> bool:
        """Brute-force predicate of tube.f90:800-818 for one (sub-segment, extended cell)."""
        rec = C.pointer(self._bin_recs[seg])
        return bool(lib().or_pair_hits(C.byref(self.params), C.byref(self._g), rec,
                                       C.c_int64(int(seg)), int(ip), int(jp), int(kp)))

    def deposit(self, sigma_line, Gamma_line, n_T_line):
        """or_deposit: FP64 vorticity (3, N, N, N) [c, z, y, x] and the number of hits."""
        N = self.params.N
        om = np.zeros(3 * N ** 3, dtype=np.float64)
        sig = np.ascontiguousarray(sigma_line, dtype=np.float64)
        gam = np.ascontiguousarray(Gamma_line, dtype=np.float64)
        nT = np.ascontiguousarray(n_T_line, dtype=np.int32)
        hits = C.c_int64(0)
        _check(lib().or_deposit(C.byref(self.params), _dp(sig), _dp(gam), _ip(nT),
                                C.byref(self._g), _dp(om), C.byref(hits)))
        return om.reshape(3, N, N, N), int(hits.value)

    def deposit_planes(self, sigma_line, Gamma_line, n_T_line, kfirst, klast):
        """or_deposit_planes: the vorticity on planes z = kfirst..klast (1-based) only, as
        (3, klast - kfirst + 1, N, N) [c, z, y, x], and the deposits on those planes."""
        N = self.params.N
        om = np.zeros(3 * N ** 3, dtype=np.float64)
        sig = np.ascontiguousarray(sigma_line, dtype=np.float64)
        gam = np.ascontiguousarray(Gamma_line, dtype=np.float64)
        nT = np.ascontiguousarray(n_T_line, dtype=np.int32)
        hits = C.c_int64(0)
        _check(lib().or_deposit_planes(C.byref(self.params), _dp(sig), _dp(gam), _ip(nT), C.byref(self._g),
                                       int(kfirst), int(klast), _dp(om), C.byref(hits)))
        return om.reshape(3, N, N, N)[:, kfirst - 1:klast].copy(), int(hits.value)

    def deposit_filtered(self, sigma_line, Gamma_line, n_T_line, if_filter):
        """or_deposit_filtered: the tube loop with the per-tube spectral filters."""
        N = self.params.N
        om = np.zeros(3 * N ** 3, dtype=np.float64)
        sig = np.ascontiguousarray(sigma_line, dtype=np.float64)
        gam = np.ascontiguousarray(Gamma_line, dtype=np.float64)
        nT = np.ascontiguousarray(n_T_line, dtype=np.int32)
        flt = np.ascontiguousarray(if_filter, dtype=np.int32)
        hits = C.c_int64(0)
        _check(lib().or_deposit_filtered(C.byref(self.params), _dp(sig), _dp(gam), _ip(nT), _ip(flt),
                                         C.byref(self._g), _dp(om), C.byref(hits)))
        return om.reshape(3, N, N, N), int(hits.value)


def deposit_range_timed(params: Params, geom: "Geometry", sigma_line, Gamma_line, n_T_line, seg_lo, seg_hi):
    """Time or_deposit_range on sub-segments [seg_lo, seg_hi) into a scratch field (seconds, hits)."""
    import time
    N = params.N
    om = np.zeros(3 * N ** 3)
    sig = np.ascontiguousarray(sigma_line, dtype=np.float64)
    gam = np.ascontiguousarray(Gamma_line, dtype=np.float64)
    nT = np.ascontiguousarray(n_T_line, dtype=np.int32)
    hits = C.c_int64(0)
    t0 = time.perf_counter()
    _check(lib().or_deposit_range(C.byref(params), _dp(sig), _dp(gam), _ip(nT), C.byref(geom._g),
                                  C.c_int64(seg_lo), C.c_int64(seg_hi), _dp(om), C.byref(hits)))
    return time.perf_counter() - t0, int(hits.value)


def fft3d(a: np.ndarray, forward=True) -> np.ndarray:
    """or_fft3d on a complex (N, N, N) array [z, y, x]; forward scaled by 1/N per axis."""
    N = a.shape[0]
    buf = np.ascontiguousarray(a, dtype=np.complex128).copy()
    _check(lib().or_fft3d(N, buf.ctypes.data_as(_PD), 1 if forward else 0))
    return buf


def fft1d(a: np.ndarray, forward=True) -> np.ndarray:
    buf = np.ascontiguousarray(a, dtype=np.complex128).copy()
    _check(lib().or_fft1d(len(buf), buf.ctypes.data_as(_PD), 1 if forward else 0))
    return buf


def biot_savart(omega: np.ndarray) -> np.ndarray:
    N = omega.shape[-1]
    om = np.ascontiguousarray(omega, dtype=np.float64)
    u = np.zeros_like(om)
    _check(lib().or_biot_savart(N, _dp(om), _dp(u)))
    return u


def normalize(omega: np.ndarray, u: np.ndarray):
    """or_normalize in place; returns (Etot_raw, up, Enstrophy, Etot_normalised)."""
    N = omega.shape[-1]
    out = np.zeros(4)
    assert omega.flags.c_contiguous and u.flags.c_contiguous
    _check(lib().or_normalize(N, _dp(omega), _dp(u), _dp(out)))
    return tuple(out)


def nek(N: int) -> int:
    return int(lib().or_nek(N))


def spectrum(f: np.ndarray) -> np.ndarray:
    N = f.shape[-1]
    ff = np.ascontiguousarray(f, dtype=np.float64)
    E = np.zeros(nek(N) + 1)
    n = _check(lib().or_spectrum(N, _dp(ff), _dp(E), len(E)))
    return E[:n]


def sp_separations(N: int) -> np.ndarray:
    r = np.zeros(64, dtype=np.int32)
    n = _check(lib().or_sp_separations(N, _ip(r), 64))
    return r[:n]


def sp_calculate(u: np.ndarray, order=10):
    N = u.shape[-1]
    uu = np.ascontiguousarray(u, dtype=np.float64)
    r = np.zeros(64, dtype=np.int32)
    Sp = np.zeros(order * 64)
    n = _check(lib().or_sp_calculate(N, order, _dp(uu), _dp(Sp), _ip(r), 64))
    return Sp[: order * n].reshape(order, n), r[:n]


def sf_x(ux: np.ndarray, pmin: int, pmax: int, seps) -> np.ndarray:
    N = ux.shape[-1]
    u = np.ascontiguousarray(ux, dtype=np.float64)
    s = np.ascontiguousarray(seps, dtype=np.int32)
    S = np.zeros((pmax - pmin + 1) * len(s))
    _check(lib().or_sf_x(N, _dp(u), pmin, pmax, len(s), _ip(s), _dp(S)))
    return S.reshape(pmax - pmin + 1, len(s))


def sigma(zeta, length, smax, smin, n_T) -> float:
    return float(lib().or_sigma(zeta, length, smax, smin, n_T))


def filter_spec(omega: np.ndarray, kc: int) -> np.ndarray:
    """or_filter_spec on a copy of omega (3, N, N, N)."""
    N = omega.shape[-1]
    om = np.ascontiguousarray(omega, dtype=np.float64).copy()
    _check(lib().or_filter_spec(N, int(kc), _dp(om)))
    return om


def run(params: Params, npoint, cx, cy, cz, sigma_line, Gamma_line, n_T_line, if_filter=None):
    """or_run: the whole path; returns omega, u (3, N, N, N) normalised, and stats."""
    N = params.N
    om = np.zeros(3 * N ** 3)
    u = np.zeros(3 * N ** 3)
    st = np.zeros(6)
    npt = np.ascontiguousarray(npoint, dtype=np.int32)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (cx, cy, cz, sigma_line, Gamma_line)]
    nT = np.ascontiguousarray(n_T_line, dtype=np.int32)
    flt = None if if_filter is None else np.ascontiguousarray(if_filter, dtype=np.int32)
    _check(lib().or_run(C.byref(params), len(npt), _ip(npt), _dp(args[0]), _dp(args[1]),
                        _dp(args[2]), _dp(args[3]), _dp(args[4]), _ip(nT),
                        None if flt is None else _ip(flt), _dp(om), _dp(u), _dp(st)))
    return om.reshape(3, N, N, N), u.reshape(3, N, N, N), st
