"""Python host interface over the wovengpu C ABI (include/wovengpu.h), through ctypes.

The reference's interface for this path is its two input files plus inline Fortran
(tube.f90:392-1043) and external subroutines (subroutine.f90). `WovenGPU` keeps the same
stages and names: build segments, bin, deposit, Biot-Savart + normalisation, spectrum and
structure functions, each one call into the HIP library. There is no CPU fallback: if the
library is missing the import of the library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from .woven_io import RunParams, Tubes

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwovengpu.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "wovengpu.h")

WT_FIELD = {"wx": 0, "wy": 1, "wz": 2, "ux": 3, "uy": 4, "uz": 5}
ERRORS = {-1: "WT_EINVAL", -2: "WT_ESTATE", -3: "WT_EDEVICE", -4: "WT_ENOMEM", -5: "WT_ERANGE"}


class WovenError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Params(C.Structure):
    _fields_ = [("N", C.c_int32), ("n_d", C.c_int32), ("line_s", C.c_int32), ("line_e", C.c_int32),
                ("sigma_ratio", C.c_double), ("eta_t", C.c_double), ("ntf_times", C.c_int32),
                ("ntf_pre", C.c_int32), ("Rtube_ratio", C.c_double), ("dzeta_sup", C.c_double)]


class _SegStats(C.Structure):
    _fields_ = [("npoint_all", C.c_int64), ("nsec_all", C.c_int64), ("nsub_all", C.c_int64),
                ("nseg_all", C.c_int64), ("ntf_max", C.c_int32), ("reserved", C.c_int32),
                ("dtf_min", C.c_double), ("ntf_sum", C.c_int64), ("theta_exact_sections", C.c_int64)]


class _DepStats(C.Structure):
    _fields_ = [("columns", C.c_int64), ("candidates", C.c_int64), ("hits", C.c_int64),
                ("fixed_point_shift", C.c_int32), ("overflow", C.c_int32), ("pairs", C.c_int64)]


_lib = None


def header_functions() -> list[str]:
    """Names of every function the public header declares."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(wt_[a-z_0-9]+)\s*\(", text)))


def load_library():
    """Load libwovengpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "or python -m fastweavturb_amd.build")
    L = C.CDLL(LIB_PATH)
    vp, i32, dp = C.c_void_p, C.c_int32, C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    L.wt_last_error.restype = C.c_char_p
    L.wt_last_error.argtypes = [vp]
    L.wt_create.argtypes = [C.POINTER(vp), i32, i32, i32]
    L.wt_destroy.argtypes = [vp]
    L.wt_destroy.restype = None
    L.wt_set_params.argtypes = [vp, C.POINTER(_Params)]
    L.wt_set_tubes.argtypes = [vp, i32, ip, dp, dp, ip, ip, dp, dp, dp]
    L.wt_build_segments.argtypes = [vp, C.POINTER(_SegStats)]
    L.wt_bin.argtypes = [vp]
    L.wt_deposit.argtypes = [vp, C.POINTER(_DepStats)]
    L.wt_biot_savart.argtypes = [vp, dp, dp, dp]
    L.wt_generate.argtypes = [vp, dp, dp, dp]
    L.wt_spectrum.argtypes = [vp, i32, dp, i32]
    L.wt_nek.argtypes = [i32]
    L.wt_nek.restype = i32
    L.wt_structure_functions.argtypes = [vp, i32, i32, i32, i32, ip, dp]
    L.wt_sp_separations.argtypes = [i32, ip, i32]
    L.wt_sp_separations.restype = i32
    L.wt_get_field.argtypes = [vp, i32, vp, i32]
    L.wt_set_field.argtypes = [vp, i32, vp, i32]
    L.wt_get_sections.argtypes = [vp, ip, dp]
    L.wt_get_subpoints.argtypes = [vp, dp, dp]
    L.wt_get_tube_lengths.argtypes = [vp, dp]
    L.wt_get_bins.argtypes = [vp, ip, dp]
    L.wt_stage_times.argtypes = [vp, dp, i32]
    L.wt_create_batch.argtypes = [C.POINTER(vp), i32, ip, i32, i32]
    L.wt_select_instance.argtypes = [vp, i32]
    L.wt_instance_scalars.argtypes = [vp, dp, dp, dp]
    L.wt_get_seg_stats.argtypes = [vp, C.POINTER(_SegStats)]
    L.wt_get_deposit_stats.argtypes = [vp, C.POINTER(_DepStats)]
    L.wt_batch_statistics.argtypes = [vp, i32, i32, i32, ip, dp, dp]
    L.wt_set_graph.argtypes = [vp, i32]
    L.wt_autotune.argtypes = [vp, C.c_char_p, C.c_char_p, i32]
    L.wt_set_deposit_tiles.argtypes = [vp, i32, C.POINTER(C.c_int32)]
    L.wt_get_deposit_tiles.argtypes = [vp, i32, C.POINTER(C.c_int32)]
    L.wt_pdf.argtypes = [vp, i32, i32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int64)]
    L.wt_filter_spec.argtypes = [vp, i32]
    L.wt_fbb_centerlines.argtypes = [vp, i32, ip, dp, ip, C.c_double, C.c_uint64, C.POINTER(C.c_uint64), dp, dp,
                                     dp, dp]
    L.wt_fbb_reference_centerlines.argtypes = [vp, i32, ip, dp, dp, C.c_uint64, C.POINTER(C.c_uint64), dp, dp, dp,
                                               dp]
    L.wt_fbb_embedding.argtypes = [i32]
    L.wt_fbb_embedding.restype = i32
    L.wt_fbb_normals.argtypes = [vp, C.c_uint64, i32, C.c_int64, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def nek(N: int) -> int:
    return int(load_library().wt_nek(N))


def fbb_embedding(npoint: int) -> int:
    """Circulant embedding size m of the device fBb generator for a tube of npoint points."""
    return int(load_library().wt_fbb_embedding(int(npoint)))


def sp_separations(N: int) -> np.ndarray:
    L = load_library()
    n = L.wt_sp_separations(N, None, 0)
    r = np.zeros(n, dtype=np.int32)
    L.wt_sp_separations(N, _ip(r), n)
    return r


class WovenGPU:
    """One device context: an N^3 grid in FP32 or FP64 on one GPU, or, with `sizes`, an
    ensemble of independent fields of grid sizes `sizes` processed together."""

    def __init__(self, N: int = 0, precision: int = 32, device: int = 0, sizes=None):
        self._L = load_library()
        h = C.c_void_p()
        if sizes is None:
            rc = self._L.wt_create(C.byref(h), N, precision, device)
            sizes = [N]
        else:
            arr = np.ascontiguousarray(sizes, dtype=np.int32)
            rc = self._L.wt_create_batch(C.byref(h), len(arr), _ip(arr), precision, device)
        if rc != 0:
            raise WovenError(rc, self._L.wt_last_error(None).decode())
        self._h = h
        self.sizes = [int(n) for n in sizes]
        self.N = self.sizes[0]
        self.precision = precision
        self.dtype = np.float64 if precision == 64 else np.float32
        self.nline = 0
        self._ntubes = {}
        self.instance = 0

    def select(self, inst: int):
        """Address instance `inst` with the per-instance calls."""
        self._check(self._L.wt_select_instance(self._h, inst))
        self.instance = inst
        self.N = self.sizes[inst]
        return self

    def scalars(self):
        e, u, en = C.c_double(), C.c_double(), C.c_double()
        self._check(self._L.wt_instance_scalars(self._h, C.byref(e), C.byref(u), C.byref(en)))
        return e.value, u.value, en.value

    def segment_stats(self) -> dict:
        s = _SegStats()
        self._check(self._L.wt_get_seg_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _SegStats._fields_ if k != "reserved"}

    def deposit_stats(self) -> dict:
        s = _DepStats()
        self._check(self._L.wt_get_deposit_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _DepStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._L.wt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise WovenError(rc, self._L.wt_last_error(self._h).decode())
        return rc

    # --- inputs --------------------------------------------------------------------------
    def set_params(self, p: RunParams):
        self._params = _Params(p.N, p.n_d, p.line_s, p.line_e, p.sigma_ratio, p.eta_t,
                               p.ntf_times, p.ntf_pre, p.Rtube_ratio, p.dzeta_sup)
        self._check(self._L.wt_set_params(self._h, C.byref(self._params)))

    def set_tubes(self, t: Tubes):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (t.sigma, t.gamma, t.cx, t.cy, t.cz)]
        npt = np.ascontiguousarray(t.npoint, dtype=np.int32)
        nT = np.ascontiguousarray(t.n_T, dtype=np.int32)
        filt = np.ascontiguousarray(t.if_filter, dtype=np.int32)
        self._check(self._L.wt_set_tubes(self._h, len(npt), _ip(npt), _dp(a[0]), _dp(a[1]), _ip(nT),
                                         _ip(filt), _dp(a[2]), _dp(a[3]), _dp(a[4])))
        self.nline = len(npt)
        self._ntubes[self.instance] = len(npt)

    # --- stages --------------------------------------------------------------------------
    def build_segments(self) -> dict:
        s = _SegStats()
        self._check(self._L.wt_build_segments(self._h, C.byref(s)))
        self.seg_stats = {k: getattr(s, k) for k, _ in _SegStats._fields_ if k != "reserved"}
        return self.seg_stats

    def bin(self):
        self._check(self._L.wt_bin(self._h))

    def deposit(self) -> dict:
        s = _DepStats()
        self._check(self._L.wt_deposit(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _DepStats._fields_}

    def biot_savart(self):
        e, u, en = C.c_double(), C.c_double(), C.c_double()
        self._check(self._L.wt_biot_savart(self._h, C.byref(e), C.byref(u), C.byref(en)))
        return e.value, u.value, en.value

    def generate(self):
        e, u, en = C.c_double(), C.c_double(), C.c_double()
        self._check(self._L.wt_generate(self._h, C.byref(e), C.byref(u), C.byref(en)))
        return e.value, u.value, en.value

    def stage_times(self) -> np.ndarray:
        """[segments, bin, deposit, biot_savart] of the last generate, then per-kernel ms."""
        ms = np.full(32, -1.0)
        n = self._check(self._L.wt_stage_times(self._h, _dp(ms), len(ms)))
        return ms[:n]

    # --- statistics ----------------------------------------------------------------------
    def spectrum(self, field: int = 1) -> np.ndarray:
        E = np.zeros(nek(self.N))
        n = self._check(self._L.wt_spectrum(self._h, field, _dp(E), len(E)))
        return E[:n]

    def structure_functions(self, mode: int, pmin: int, pmax: int, seps) -> np.ndarray:
        s = np.ascontiguousarray(seps, dtype=np.int32)
        S = np.zeros((pmax - pmin + 1) * len(s))
        self._check(self._L.wt_structure_functions(self._h, mode, pmin, pmax, len(s), _ip(s), _dp(S)))
        return S.reshape(pmax - pmin + 1, len(s))

    def batch_statistics(self, pmin: int, pmax: int, seps):
        """(spectra, S) of every instance: spectra[i] is instance i's integer-shell spectrum of the
        normalised velocity; S[i, p - pmin, m] its x-increment structure function of order p at
        separation seps[i][m] (seps: one row per instance, or one row for all)."""
        n = len(self.sizes)
        sp = np.ascontiguousarray(np.broadcast_to(np.asarray(seps, dtype=np.int32), (n, np.shape(seps)[-1])))
        nsep = sp.shape[1]
        neks = [nek(N) for N in self.sizes]
        E = np.zeros(sum(neks))
        S = np.zeros((n, pmax - pmin + 1, nsep))
        self._check(self._L.wt_batch_statistics(self._h, pmin, pmax, nsep, _ip(sp), _dp(E), _dp(S)))
        offs = np.concatenate([[0], np.cumsum(neks)])
        return [E[offs[i]:offs[i + 1]] for i in range(n)], S

    def autotune(self, cache_path: str | None = None) -> str:
        """Times and keeps the fastest FFT tile shapes for this context's grid sizes and, once
        tubes are set, the fastest deposition tiles; returns the per-size report."""
        buf = C.create_string_buffer(8192)
        self._check(self._L.wt_autotune(self._h, cache_path.encode() if cache_path else None, buf, len(buf)))
        return buf.value.decode()

    def set_deposit_tiles(self, bx: int, by: int, bz: int, colcap: int = 8, N: int = 0):
        """Brick shape and columns per pass of the deposition for grid size N (0: every instance)."""
        t = (C.c_int32 * 4)(bx, by, bz, colcap)
        self._check(self._L.wt_set_deposit_tiles(self._h, int(N), t))

    def deposit_tiles(self, N: int = 0) -> tuple:
        """(bx, by, bz, colcap) of grid size N (0: the selected instance's)."""
        t = (C.c_int32 * 4)()
        self._check(self._L.wt_get_deposit_tiles(self._h, int(N), t))
        return tuple(int(v) for v in t)

    def pdf(self, field: int = 1, nbins: int = 1000):
        """(centres, pdf, counts) of the three components of omega (0) or u (1), as cal_pdf."""
        val = np.zeros(nbins, dtype=np.float32)
        pdf = np.zeros(nbins, dtype=np.float32)
        cnt = np.zeros(nbins, dtype=np.int64)
        fp = C.POINTER(C.c_float)
        self._check(self._L.wt_pdf(self._h, field, nbins, val.ctypes.data_as(fp), pdf.ctypes.data_as(fp),
                                   cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return val, pdf, cnt

    def filter_spec(self, k_c: int):
        """Sharp spectral cut of the resident vorticity at k_c (filter_spec)."""
        self._check(self._L.wt_filter_spec(self._h, int(k_c)))

    # --- fBb centerline generator ----------------------------------------------------------
    def fbb_centerlines(self, npoint, hurst, winding, wiggle: float, seed: int, gauss=None, keys=None):
        """(cx, cy, cz) of closed fBb centerlines generated on this context's GPU: tube l has
        npoint[l] points, Hurst index hurst[l] and lattice winding winding[l] (3 ints). gauss: None
        (the device's Philox stream keyed (seed, keys[l] or l)) or, tube after tube, 4m normals and
        3 uniforms (m = fbb_embedding(npoint[l]))."""
        npt = np.ascontiguousarray(npoint, dtype=np.int32)
        H = np.ascontiguousarray(np.broadcast_to(np.asarray(hurst, dtype=np.float64), npt.shape))
        w = np.ascontiguousarray(np.asarray(winding, dtype=np.int32).reshape(len(npt), 3))
        P = int(npt.sum())
        out = [np.empty(P) for _ in range(3)]
        g = None if gauss is None else np.ascontiguousarray(gauss, dtype=np.float64)
        k = None if keys is None else np.ascontiguousarray(keys, dtype=np.uint64)
        self._check(self._L.wt_fbb_centerlines(self._h, len(npt), _ip(npt), _dp(H), _ip(w), float(wiggle),
                                               C.c_uint64(int(seed)),
                                               None if k is None else k.ctypes.data_as(C.POINTER(C.c_uint64)),
                                               None if g is None else _dp(g),
                                               _dp(out[0]), _dp(out[1]), _dp(out[2])))
        return out

    def fbb_reference_centerlines(self, npoint, hurst, dB, seed: int, gauss=None, keys=None):
        """(cx, cy, cz) of centerlines built as FBB_with_parameters.m builds them (no winding drift,
        step rms dB[l], tubes of fewer than 128 points sampled from a 128-point path, the wrap
        mod(x, 2 pi) - pi); random inputs as for fbb_centerlines, with m the embedding of
        max(npoint[l], 128) points."""
        npt = np.ascontiguousarray(npoint, dtype=np.int32)
        H = np.ascontiguousarray(np.broadcast_to(np.asarray(hurst, dtype=np.float64), npt.shape))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dB, dtype=np.float64), npt.shape))
        P = int(npt.sum())
        out = [np.empty(P) for _ in range(3)]
        g = None if gauss is None else np.ascontiguousarray(gauss, dtype=np.float64)
        k = None if keys is None else np.ascontiguousarray(keys, dtype=np.uint64)
        self._check(self._L.wt_fbb_reference_centerlines(self._h, len(npt), _ip(npt), _dp(H), _dp(d),
                                                         C.c_uint64(int(seed)),
                                                         None if k is None else k.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                         None if g is None else _dp(g),
                                                         _dp(out[0]), _dp(out[1]), _dp(out[2])))
        return out

    def fbb_normals(self, seed: int, tube: int, count: int) -> np.ndarray:
        out = np.empty(int(count))
        self._check(self._L.wt_fbb_normals(self._h, C.c_uint64(int(seed)), int(tube), int(count), _dp(out)))
        return out

    def set_graph(self, enable: bool = True):
        self._check(self._L.wt_set_graph(self._h, int(enable)))

    # --- fields --------------------------------------------------------------------------
    def get_field(self, which, fp64: bool = True) -> np.ndarray:
        w = WT_FIELD[which] if isinstance(which, str) else int(which)
        out = np.empty((self.N,) * 3, dtype=np.float64 if fp64 else np.float32)
        self._check(self._L.wt_get_field(self._h, w, out.ctypes.data_as(C.c_void_p), int(fp64)))
        return out

    def get_fields(self, name: str, fp64: bool = True) -> np.ndarray:
        base = 0 if name == "omega" else 3
        return np.stack([self.get_field(base + c, fp64) for c in range(3)])

    def set_field(self, which, arr: np.ndarray):
        w = WT_FIELD[which] if isinstance(which, str) else int(which)
        a = np.ascontiguousarray(arr)
        if a.shape != (self.N,) * 3:
            raise ValueError(f"field must have shape {(self.N,) * 3}")
        fp64 = a.dtype == np.float64
        if not fp64:
            a = a.astype(np.float32, copy=False)
        self._check(self._L.wt_set_field(self._h, w, a.ctypes.data_as(C.c_void_p), int(fp64)))

    def set_fields(self, name: str, arr: np.ndarray):
        base = 0 if name == "omega" else 3
        for c in range(3):
            self.set_field(base + c, arr[c])

    # --- intermediate arrays (parity inspection) -------------------------------------------
    def sections(self):
        S = self.segment_stats()["nsec_all"]
        ntf = np.zeros(S, dtype=np.int32)
        dtf = np.zeros(S)
        self._check(self._L.wt_get_sections(self._h, _ip(ntf), _dp(dtf)))
        return ntf, dtf

    def subpoints(self):
        n = self.segment_stats()["nsub_all"]
        ci = np.zeros((n, 3))
        s = np.zeros(n)
        self._check(self._L.wt_get_subpoints(self._h, _dp(ci), _dp(s)))
        return ci, s

    def tube_lengths(self) -> np.ndarray:
        """2 pi Len_Disc of every tube of the selected instance (tube.f90:501)."""
        out = np.zeros(self._ntubes.get(self.instance, 0))
        self._check(self._L.wt_get_tube_lengths(self._h, _dp(out)))
        return out

    def bins(self):
        n = self.segment_stats()["nsub_all"]
        lohi = np.zeros((n, 6), dtype=np.int32)
        R = np.zeros(n)
        self._check(self._L.wt_get_bins(self._h, _ip(lohi), _dp(R)))
        return lohi, R


def run_woven(p: RunParams, t: Tubes, precision: int = 32, device: int = 0):
    """Whole path on one GPU; returns the context with omega and u resident, and stats."""
    g = WovenGPU(p.N, precision, device)
    g.set_params(p)
    g.set_tubes(t)
    etot, up, ens = g.generate()
    return g, {"etot": etot, "up": up, "enstrophy": ens}
