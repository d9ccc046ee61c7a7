"""Pins of the PDF oracle (oracle/pdf_oracle.py): the vectorised first-match search equals the
reference's literal loop (subroutine.f90:23-33) on inputs with ties, duplicates, edge values and a
constant field; densities of a Gaussian sample follow the normal density."""
import math

import numpy as np

from oracle import pdf_oracle as PO


def _check(v, nb, shape):
    a = PO.cal_pdf(v, nb, shape)
    b = PO.cal_pdf_loop(v, nb, shape)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    return a


def test_vectorised_search_equals_the_reference_loop():
    rng = np.random.default_rng(0)
    for trial in range(6):
        v = rng.standard_normal(3 * 6 ** 3).astype(np.float32) * (10.0 ** rng.uniform(-3, 3))
        if trial % 2:
            v[::7] = v[3]                                  # duplicates
        _, _, cnt = _check(v, 37, (6, 6, 18))
        assert cnt.sum() <= v.size
    # a large offset: bin width at or below the rounding unit of the values (collapsed edges)
    for off, spread in ((1000.0, 1e-3), (1000.0, 3e-2), (-5.0e4, 1.0)):
        v = (off + spread * rng.standard_normal(900)).astype(np.float32)
        _check(v, 200, (1, 1, 900))
    # values placed exactly on bin edges
    v = rng.standard_normal(600).astype(np.float32)
    _, _, lo, hi = PO.bin_edges(v, 25)[1:]
    v2 = np.concatenate([v, lo[3:20], hi[3:20]])
    _check(v2, 25, (1, 1, v2.size))


def test_maximum_is_dropped_when_the_last_interval_excludes_it():
    v = np.array([0.0, 1.0, 2.0, 3.0, 4.0], dtype=np.float32)
    val, pdf, cnt = _check(v, 4, (1, 1, 5))
    np.testing.assert_array_equal(val, [0.5, 1.5, 2.5, 3.5])
    np.testing.assert_array_equal(cnt, [1, 1, 1, 1])      # 4.0 lies in no [val - 0.5, val + 0.5)
    np.testing.assert_allclose(pdf, 1.0 / 5.0)


def test_constant_field_counts_nothing():
    v = np.full(30, 2.5, dtype=np.float32)
    _, pdf, cnt = _check(v, 10, (1, 1, 30))
    assert cnt.sum() == 0 and np.all(np.isnan(pdf))        # bin_size 0: 0/0, as the reference


def test_gaussian_sample_density():
    rng = np.random.default_rng(1)
    v = rng.standard_normal(2_000_000).astype(np.float32)
    val, pdf, cnt = PO.cal_pdf(v, 200, (1, 1, v.size))
    mid = np.abs(val) < 2.0
    ref = np.exp(-0.5 * val[mid].astype(np.float64) ** 2) / math.sqrt(2 * math.pi)
    np.testing.assert_allclose(pdf[mid], ref, rtol=0.03)
    bs = float(val[1] - val[0])
    assert abs(float(pdf.sum()) * bs - 1.0) < 1e-4
