"""Autotuner (wt_autotune): the tuned FFT and deposition tiles give the same fields (the per-line
FFT arithmetic is unchanged; only the grouping of the spectral partial sums moves, so up differs in
the last bits; the deposition's fixed-point sums do not depend on the bricks), and the JSON cache
round-trips."""
import json

import numpy as np
import pytest

from fastweavturb_amd import fbb

pytestmark = pytest.mark.gpu

SIZES = [48, 32, 64, 32]


def _ctx(graph=False):
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(precision=32, sizes=SIZES)
    g.set_graph(graph)
    for i, N in enumerate(SIZES):
        p, t = fbb.config_small(N=N, seed=300 + i, ntubes=(1, 2), npts=(40, 30), radii=(0.5, 0.3))
        g.select(i)
        g.set_params(p)
        g.set_tubes(t)
    return g


def test_tuned_shapes_give_the_same_fields(tmp_path):
    ref = _ctx()
    ref.generate()
    g = _ctx(graph=True)
    path = str(tmp_path / "tiles.json")
    rep = g.autotune(path)
    assert all(f"N={N}" in rep for N in set(SIZES))
    for _ in range(3):                         # eager, capture, replay
        g.generate()
        for i in range(len(SIZES)):
            g.select(i)
            ref.select(i)
            for name in ("omega", "u"):
                a, b = g.get_fields(name), ref.get_fields(name)
                assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
            np.testing.assert_allclose(g.spectrum(1), ref.spectrum(1), rtol=1e-5)
    tab = json.load(open(path))
    keys = [k for k in tab if k.endswith(":32:64")]
    assert keys and set(tab[keys[0]]) == {"ppb", "tx", "wy", "ty", "wz", "tz"}
    # a second context reads the cache instead of measuring
    h = _ctx()
    rep2 = h.autotune(path)
    assert rep2.count("from cache") == 2 * len(set(SIZES))      # FFT and deposition tiles
    dep = [k for k in tab if ":32:dep:" in k]
    assert len(dep) == len(set(SIZES)) and all(set(tab[k]) == {"bx", "by", "bz", "colcap"} for k in dep)
