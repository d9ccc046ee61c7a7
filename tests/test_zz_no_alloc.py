"""wt_structure_functions_noalloc (partial sums in a fixed device array, no device allocation)
against wt_structure_functions on the same resident velocity: two calls in a row, the second on
a smaller grid with fewer separations and orders, and an x-only and a Sp_calculate call; results
bitwise equal. Device buffers for the separations and results are the caller's, allocated here
through the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

from fastweavturb_amd import fbb

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceArray:
    """A caller-owned device buffer holding a copy of a numpy array."""

    def __init__(self, hip, arr):
        self.hip = hip
        self.nbytes = arr.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _hip():
    hip = C.CDLL("libamdhip64.so.7")      # the runtime libwovengpu.so is linked against
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _ctx(N, seed, precision):
    from fastweavturb_amd.api import WovenGPU
    p, t = fbb.config_small(N=N, seed=seed, ntubes=(1, 2), npts=(40, 30), radii=(0.5, 0.3), ntf_pre=100)
    g = WovenGPU(N, precision)
    g.set_params(p)
    g.set_tubes(t)
    g.generate()
    return g


def _noalloc(L, hip, g, mode, pmin, pmax, seps):
    seps = np.ascontiguousarray(seps, dtype=np.int32)
    S = np.zeros((pmax - pmin + 1) * len(seps))
    d_seps, d_S = DeviceArray(hip, seps), DeviceArray(hip, S)
    try:
        rc = L.wt_structure_functions_noalloc(g._h, mode, pmin, pmax, len(seps), d_seps.ptr, d_S.ptr)
        assert rc == 0, L.wt_last_error(g._h)
        return d_S.get(S).reshape(pmax - pmin + 1, len(seps))
    finally:
        d_seps.free()
        d_S.free()


@pytest.mark.parametrize("precision", [32, 64])
def test_noalloc_equals_allocating_entry_point(precision):
    from fastweavturb_amd.api import load_library
    L = load_library()
    L.wt_structure_functions_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                 C.c_void_p, C.c_void_p]
    hip = _hip()
    calls = [(64, 5, 1, 2, 6, np.arange(1, 41)),          # x-only, 40 separations x 5 orders
             (32, 6, 1, 3, 5, np.array([1, 3, 7, 16, 31, 2])),   # smaller grid, fewer of both
             (48, 7, 0, 1, 6, np.array([1, 1, 2, 3, 4, 6, 8, 11, 16, 23, 32, 45]))]   # Sp_calculate
    for N, seed, mode, pmin, pmax, seps in calls:
        print(f"noalloc: N={N} FP{precision} mode {mode} orders {pmin}..{pmax} nsep {len(seps)}", flush=True)
        g = _ctx(N, seed, precision)
        ref = g.structure_functions(mode, pmin, pmax, seps)
        got = _noalloc(L, hip, g, mode, pmin, pmax, seps)
        np.testing.assert_array_equal(got, ref)
        # again on the same context: the shared array keeps the last call's contents
        np.testing.assert_array_equal(_noalloc(L, hip, g, mode, pmin, pmax, seps), ref)
        g.close()


def test_noalloc_rejects_requests_beyond_its_scratch():
    from fastweavturb_amd.api import load_library
    L = load_library()
    L.wt_structure_functions_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                 C.c_void_p, C.c_void_p]
    hip = _hip()
    g = _ctx(32, 9, 32)
    seps = np.ones(3000, dtype=np.int32)
    S = np.zeros(5 * 3000)
    d_seps, d_S = DeviceArray(hip, seps), DeviceArray(hip, S)
    try:
        rc = L.wt_structure_functions_noalloc(g._h, 1, 2, 6, 3000, d_seps.ptr, d_S.ptr)
        assert rc == -5, rc                                   # WT_ERANGE, nothing launched
        assert b"scratch" in L.wt_last_error(g._h)
    finally:
        d_seps.free()
        d_S.free()
        g.close()
