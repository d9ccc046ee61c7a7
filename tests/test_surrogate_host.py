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
