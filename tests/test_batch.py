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


def test_batch_statistics_match_per_instance_calls():
    from fastweavturb_amd.api import WovenGPU
    cfg = configs()
    b = WovenGPU(precision=64, sizes=SIZES)
    for i, (p, t) in enumerate(cfg):
        b.select(i)
        b.set_params(p)
        b.set_tubes(t)
    b.generate()
    seps = np.array([[1, 2, 3, 5], [2, 4, 6, 8], [1, 3, 5, 7], [4, 5, 6, 7], [1, 1, 2, 11]], dtype=np.int32)
    E, S = b.batch_statistics(2, 6, seps)
    for i in range(len(SIZES)):
        b.select(i)
        np.testing.assert_array_equal(E[i], b.spectrum(1))
        np.testing.assert_allclose(S[i], b.structure_functions(1, 2, 6, seps[i]), rtol=1e-12, atol=1e-300)


GRAPH_SIZES = [48, 32, 64, 48, 32, 64]


def _ensemble(seed):
    from fastweavturb_amd import surrogate as Sg
    rng = np.random.default_rng(seed)
    return [Sg.instance_config(Sg.draw_design(rng, N, nlev=3), N, 100 * seed + i, ntf_pre=100)
            for i, N in enumerate(GRAPH_SIZES)]


def test_graph_replay_equals_eager_runs():
    """The spectral stage replayed from one hipGraph, on new tubes each time and with eager work
    between replays, equals the eager path bitwise."""
    from fastweavturb_amd.api import WovenGPU
    sizes = [p.N for p, _ in _ensemble(1)]
    gg = WovenGPU(precision=32, sizes=sizes)
    gg.set_graph(True)
    ge = WovenGPU(precision=32, sizes=sizes)
    for rep, seed in enumerate((1, 2, 1, 3)):
        cfg = _ensemble(seed)
        assert [p.N for p, _ in cfg] == sizes
        for g in (gg, ge):
            for i, (p, t) in enumerate(cfg):
                g.select(i)
                g.set_params(p)
                g.set_tubes(t)
            g.generate()
        for i in range(len(cfg)):
            gg.select(i)
            ge.select(i)
            for name in ("omega", "u"):
                np.testing.assert_array_equal(gg.get_fields(name, fp64=False), ge.get_fields(name, fp64=False))
            # shell sums add in LDS with atomics (order varies at the last bit); Etot, Enstrophy
            # and up are reduced in a fixed order
            np.testing.assert_allclose(gg.spectrum(1), ge.spectrum(1), rtol=1e-14)
            assert gg.scalars() == ge.scalars()
        # eager work on the graph's context between replays
        gg.select(0)
        gg.structure_functions(1, 2, 4, [1, 2])


def test_whole_generation_graph_replay_equals_eager_runs():
    """wt_set_graph(2): geometry, binning, deposition and the spectral stage of the current tubes
    replayed from one hipGraph equal the eager path bitwise, replay after replay, with eager work
    on the context (including overwritten fields) between replays; new tubes are captured again."""
    from fastweavturb_amd.api import WovenGPU
    sizes = [p.N for p, _ in _ensemble(1)]
    gg = WovenGPU(precision=32, sizes=sizes)
    gg.set_graph(2)
    ge = WovenGPU(precision=32, sizes=sizes)
    for seed in (1, 2):
        cfg = _ensemble(seed)
        for g in (gg, ge):
            for i, (p, t) in enumerate(cfg):
                g.select(i)
                g.set_params(p)
                g.set_tubes(t)
        ge.generate()
        ref = []
        for i in range(len(cfg)):
            ge.select(i)
            ref.append((ge.get_fields("omega", fp64=False), ge.get_fields("u", fp64=False), ge.spectrum(1),
                        ge.scalars()))
        dep = ge.deposit_stats()
        for rep in range(4):                     # eager + capture, then replays
            gg.generate()
            st = gg.stage_times()
            assert (st[1] == -1.0) == (rep > 0), st[:4]
            for i in range(len(cfg)):
                gg.select(i)
                np.testing.assert_array_equal(gg.get_fields("omega", fp64=False), ref[i][0])
                np.testing.assert_array_equal(gg.get_fields("u", fp64=False), ref[i][1])
                np.testing.assert_allclose(gg.spectrum(1), ref[i][2], rtol=1e-14)
                assert gg.scalars() == ref[i][3]
            assert gg.deposit_stats()["hits"] == dep["hits"]
            # eager work between replays: statistics, and instance 0's fields overwritten
            gg.select(0)
            gg.structure_functions(1, 2, 4, [1, 2])
            N0 = sizes[0]
            gg.set_fields("omega", np.zeros((3, N0, N0, N0), dtype=np.float32))
