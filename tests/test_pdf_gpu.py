"""wt_pdf (cal_pdf / export_pdf, subroutine.f90:1-62) against the PDF oracle: bin centres, counts
and densities bit-exact on the same field values, in FP32 and FP64 contexts; planted fields cover
values on bin edges, duplicates, a field offset far from zero (bin width near the rounding unit)
and a constant field."""
import numpy as np
import pytest

from fastweavturb_amd import fbb
from oracle import pdf_oracle as PO

pytestmark = pytest.mark.gpu


def _ctx(N, precision):
    from fastweavturb_amd.api import WovenGPU
    return WovenGPU(N, precision)


def _check(g, field, nbins, arr):
    N = g.N
    val, pdf, cnt = g.pdf(field, nbins)
    rv, rp, rc = PO.cal_pdf(arr.astype(np.float32), nbins, (N, N, 3 * N))
    np.testing.assert_array_equal(val, rv)
    np.testing.assert_array_equal(cnt, rc)
    np.testing.assert_array_equal(pdf, rp)
    return cnt


@pytest.mark.parametrize("precision", [32, 64])
def test_pdf_of_generated_fields(precision):
    from fastweavturb_amd.api import run_woven
    p, t = fbb.config_small(N=32)
    g, _ = run_woven(p, t, precision)
    for field, name in ((1, "u"), (0, "omega")):
        arr = g.get_fields(name, fp64=precision == 64)
        for nb in (1000, 1, 7, 8192):
            cnt = _check(g, field, nb, arr)
            # rounded edges can leave gaps between hi(l) and lo(l+1): values there count nowhere
            assert int(cnt.sum()) <= arr.size
    g.close()


@pytest.mark.parametrize("precision", [32, 64])
def test_pdf_planted_edge_cases(precision):
    N = 16
    g = _ctx(N, precision)
    rng = np.random.default_rng(3)
    dt = np.float64 if precision == 64 else np.float32
    # values exactly on the bin edges of a 50-bin table, plus duplicates
    base = rng.standard_normal((3, N, N, N)).astype(np.float32)
    _, _, lo, hi = PO.bin_edges(base, 50)[1:]
    flat = base.ravel().copy()
    flat[1:50] = lo[1:50]
    flat[100:149] = hi[:49]
    flat[200:400] = flat[5]
    fields = [flat.reshape(3, N, N, N),
              (1000.0 + 1e-3 * rng.standard_normal((3, N, N, N))),          # bin width ~ rounding unit
              (rng.standard_normal((3, N, N, N)) * 1e30),
              np.full((3, N, N, N), 0.25)]
    for k, arr in enumerate(fields):
        arr = arr.astype(dt)
        g.set_fields("u", arr)
        for nb in (50, 1000):
            cnt = _check(g, 1, nb, arr)
            if k == 3:
                assert cnt.sum() == 0
    g.close()


def test_pdf_full_size_properties():
    """256^3 generated velocity: bit-exact against the oracle, densities integrate to the counted
    fraction, and the counts do not depend on the number of bins beyond merging (1000 bins summed in
    pairs equal 500 bins, up to values that move across edges by one rounding)."""
    from fastweavturb_amd.api import run_woven
    p, t = fbb.config_small(N=256, seed=4, ntubes=(2, 4), npts=(60, 40), radii=(0.8, 0.4))
    g, _ = run_woven(p, t, 32)
    u = g.get_fields("u", fp64=False)
    cnt = _check(g, 1, 1000, u)
    val, pdf, _ = g.pdf(1, 1000)
    bs = np.float64(val[1]) - np.float64(val[0])
    assert abs(pdf.astype(np.float64).sum() * bs - cnt.sum() / u.size) < 1e-4
    c500 = g.pdf(1, 500)[2]
    assert np.abs(cnt.reshape(500, 2).sum(axis=1) - c500).sum() <= 1e-4 * u.size
    g.close()


def test_pdf_argument_errors():
    from fastweavturb_amd.api import WovenError
    g = _ctx(16, 32)
    with pytest.raises(WovenError):
        g.pdf(1, 100)                      # velocity not resident
    g.set_fields("u", np.zeros((3, 16, 16, 16), dtype=np.float32))
    for field, nb in ((2, 100), (1, 0), (1, 8193)):
        with pytest.raises(WovenError):
            g.pdf(field, nb)
    g.close()
