"""Ensemble batching: several independent woven fields of uneven grid sizes in one context.
Each instance must equal the single-instance path (bitwise in FP32, whose deposition is
fixed-point and whose per-instance arithmetic is unchanged by batching) and the oracle."""
import numpy as np
import pytest

from fastweavturb_amd import fbb

pytestmark = pytest.mark.gpu

SIZES = [24, 16, 32, 16, 12]


def configs():
    out = []
    for i, N in enumerate(SIZES):
        p, t = fbb.config_small(N=N, seed=100 + i, ntubes=(1, 1 + i % 2), npts=(30 + 3 * i, 24), radii=(0.5, 0.3))
        out.append((p, t))
    return out


def single(p, t, precision):
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(p.N, precision)
    g.set_params(p)
    g.set_tubes(t)
    g.generate()
    return g


@pytest.mark.parametrize("precision", [32, 64])
def test_batch_equals_single_instances(oracle, precision):
    from fastweavturb_amd.api import WovenGPU
    cfg = configs()
    b = WovenGPU(precision=precision, sizes=SIZES)
    for i, (p, t) in enumerate(cfg):
        b.select(i)
        b.set_params(p)
        b.set_tubes(t)
    b.generate()
    for i, (p, t) in enumerate(cfg):
        s = single(p, t, precision)
        b.select(i)
        for name in ("omega", "u"):
            a, r = b.get_fields(name, fp64=precision == 64), s.get_fields(name, fp64=precision == 64)
            if precision == 32:
                np.testing.assert_array_equal(a, r)
            else:
                assert np.linalg.norm(a - r) <= 1e-13 * np.linalg.norm(r)
        np.testing.assert_allclose(b.spectrum(1), s.spectrum(1), rtol=1e-12 if precision == 64 else 1e-6)
        e1, u1, n1 = b.scalars()
        e2, u2, n2 = s.scalars()
        assert u1 == pytest.approx(u2, rel=1e-12) and n1 == pytest.approx(n2, rel=1e-10)
    # one instance against the oracle
    p, t = cfg[2]
    P = oracle.Params(p.N, p.n_d, p.line_s, p.line_e, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre,
                      p.Rtube_ratio, p.dzeta_sup)
    om, u, st = oracle.run(P, t.npoint, t.cx, t.cy, t.cz, t.sigma, t.gamma, t.n_T)
    b.select(2)
    tol = 1e-12 if precision == 64 else 1e-5
    assert np.linalg.norm(b.get_fields("u") - u) <= tol * np.linalg.norm(u)
    seps = np.array([1, 2, 4, 8])
    np.testing.assert_allclose(b.structure_functions(1, 2, 6, seps), oracle.sf_x(u[0], 2, 6, seps),
                               rtol=1e-9 if precision == 64 else 1e-4)


def test_batch_requires_every_instance():
    from fastweavturb_amd.api import WovenError, WovenGPU
    p, t = fbb.config_small(N=16)
    b = WovenGPU(precision=32, sizes=[16, 16])
    b.set_params(p)
    b.set_tubes(t)
    with pytest.raises(WovenError) as e:
        b.generate()
    assert e.value.code == -2
