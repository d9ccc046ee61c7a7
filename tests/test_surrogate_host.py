"""Host logic of the inverse design (fastweavturb_amd/surrogate.py) on the CPU: design encoding,
the statistics' definitions, and the centerline bank's construction."""
import math

import numpy as np
import pytest

from fastweavturb_amd import surrogate as S


def test_design_vector_round_trip_and_mask():
    d = S.family_design(1.3, 1.4142135623730951, 0.7, nlev=4)
    v, m = d.vector()
    assert v.shape == (24,) and m.shape == (24,)
    assert m.reshape(6, 4)[:4].sum() == 15 and m.reshape(6, 4)[4:].sum() == 0   # a_0 masked
    d2 = S.Design.from_vector(v, 4)
    np.testing.assert_allclose(d2.radii(), 0.8 * 2.0 ** -np.arange(4), rtol=1e-15)
    np.testing.assert_allclose(d2.a, 1.3 * np.log2(d2.radii() / 0.8), rtol=1e-14)
    np.testing.assert_allclose(d2.b, 0.5, rtol=1e-14)
    assert np.all(d2.H == 0.7)


def test_level_law_matches_the_512_workload():
    d = S.family_design(4.0 / 3.0, 1.0, 0.75)
    lv = d.levels()
    assert [x.ntubes for x in lv] == [4 * 4 ** l for l in range(6)]
    np.testing.assert_allclose([x.radius for x in lv], 0.8 * 2.0 ** -np.arange(6))
    g = np.array([x.gamma for x in lv])
    np.testing.assert_allclose(g / g[0], (np.array([x.radius for x in lv]) / 0.8) ** (4.0 / 3.0), rtol=1e-12)


def test_log_shells_definition():
    for nek in (32, 64, 256):
        np.testing.assert_allclose(S.log_shells(np.full(nek, 3.5)), 3.5, rtol=1e-14)   # weights sum to 1
        E = np.arange(1, nek + 1, dtype=float) ** (-5.0 / 3.0)
        ls = S.log_shells(E)
        assert np.all(np.diff(ls) <= 1e-15)                  # a decreasing spectrum stays non-increasing
        # the first centre (k = 1) sees only its own shell; the last (k = 16) a narrow average
        # around shell 16 (Gaussian width of one centre spacing in ln k)
        assert ls[0] == pytest.approx(1.0, rel=1e-9)
        assert ls[31] == pytest.approx(16.0 ** (-5.0 / 3.0), rel=0.02)
    # a spectrum that does not resolve k = 16 is not a valid input for every centre
    assert S.shell_weights(16).shape == (32, 16)


def test_flatness_separations_are_box_units():
    for N in S.GRIDS + (512,):
        seps = S.flat_separations(N)
        np.testing.assert_array_equal(seps * 32, np.array(S.FLAT_M) * N)
    with pytest.raises(ValueError):
        S.flat_separations(100)


def test_features_of_gaussian_increments():
    # Gaussian increments: S4 = 3 S2^2, S6 = 15 S2^3
    s2 = np.linspace(0.1, 2.0, 8)
    Sm = np.stack([s2, np.zeros(8), 3 * s2 ** 2, np.zeros(8), 15 * s2 ** 3])
    f = S.features(np.full(32, 0.01), Sm)
    np.testing.assert_allclose(f[32:40], math.log10(3.0), rtol=1e-13)
    np.testing.assert_allclose(f[40:48], math.log10(15.0), rtol=1e-13)
    np.testing.assert_allclose(f[:32], -2.0, rtol=1e-13)


def test_density_snapping():
    lv = np.log2(S.DENSITIES)
    for i, v in enumerate(lv):
        assert S.snap_density([v + 0.2, v - 0.1]) == pytest.approx(v)
    assert S.snap_density([5.0]) == pytest.approx(1.0)


@pytest.fixture(scope="module")
def bank():
    return S.CenterlineBank(31, nlev=3)


def test_bank_is_deterministic_and_loads(bank, tmp_path):
    b2 = S.CenterlineBank(31, nlev=3)
    for k in bank.a:
        np.testing.assert_array_equal(np.asarray(bank.a[k]), np.asarray(b2.a[k]))
    p = str(tmp_path / "bank.npz")
    bank.save(p)
    b3 = S.CenterlineBank.load(p)
    d = S.family_design(1.3, 1.0, 0.63, nlev=3)
    t1, t3 = bank.tubes(d), b3.tubes(d)
    np.testing.assert_array_equal(t1.cx, t3.cx)
    np.testing.assert_array_equal(t1.gamma, t3.gamma)


@pytest.mark.parametrize("H", [0.5, 0.63, 0.9])
def test_bank_wiggle_and_wrap(bank, H):
    d = S.family_design(1.3, 1.0, H, nlev=3)
    t = bank.tubes(d)
    assert t.cx.min() >= -math.pi and t.cx.max() < math.pi
    off = np.concatenate([[0], np.cumsum(t.npoint)])
    k = 4 + 16                                  # first tube of level 2
    a, e = off[k], off[k + 1]
    n = e - a
    xs = np.stack([t.cx[a:e], t.cy[a:e], t.cz[a:e]])
    st = (np.diff(xs, axis=1) + math.pi) % (2 * math.pi) - math.pi
    drift = 2 * math.pi * np.asarray(bank.a["w2"]) / n
    dev = st - drift[:, None]
    assert np.sqrt((dev ** 2).sum() / (n - 1)) == pytest.approx(0.35 * np.linalg.norm(drift), rel=1e-5)


def test_bank_fractional_tube_is_continuous(bank):
    d1 = S.family_design(1.3, 1.0, 0.7, nlev=3)
    d2 = S.family_design(1.3, 1.0 + 1e-6, 0.7, nlev=3)
    t1, t2 = bank.tubes(d1), bank.tubes(d2)
    # 4, 16, 64 tubes become 5, 17, 65 with the new tubes at ~1e-6 of a tube's circulation
    assert t2.nline == t1.nline + 3
    new = np.concatenate([[4], [4 + 1 + 16], [4 + 1 + 16 + 1 + 64]])
    assert np.all(np.abs(t2.gamma[new]) < 1e-4 * np.abs(t1.gamma).max())
