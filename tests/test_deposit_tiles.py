"""Deposition tiles (wt_set_deposit_tiles, tuned by wt_autotune): the brick shape and the columns
walked per pass only regroup the work, so every choice gives the same deposition (tube.f90:800-892):
the same hits, FP32 fields bitwise equal (fixed-point accumulation), FP64 fields equal up to the
order of additions, and pair counts equal to the host restatement of the binning for that shape,
including bricks that do not divide N and boxes of more than 256 bricks."""
import numpy as np
import pytest

from binning_rule import slab_pairs
from fastweavturb_amd import fbb

pytestmark = pytest.mark.gpu

TILES = [(16, 16, 16, 8), (32, 16, 8, 8), (8, 16, 32, 4), (8, 8, 8, 16), (4, 4, 4, 1), (64, 8, 8, 64),
         (32, 32, 4, 8)]


def _oracle_geo(O, p, t):
    P = O.Params(p.N, p.n_d, p.line_s, p.line_e, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre, p.Rtube_ratio,
                 p.dzeta_sup)
    return O.Geometry(P, t.npoint, t.cx, t.cy, t.cz)


@pytest.mark.parametrize("N", [48, 64])
def test_every_tile_gives_the_same_deposition(oracle, N):
    from fastweavturb_amd.api import WovenGPU
    p, t = fbb.config_small(N=N, seed=41, ntubes=(1, 3), npts=(60, 40), radii=(0.6, 0.25), ntf_pre=60)
    geo = _oracle_geo(oracle, p, t)
    b = geo.bins(t.sigma, t.n_T)
    kept = {tl[:3]: slab_pairs(geo, b, N, tl[:3])[0] for tl in TILES}
    _, hits = geo.deposit(t.sigma, t.gamma, t.n_T)
    geo.close()
    for prec in (32, 64):
        g = WovenGPU(N, prec)
        g.set_params(p)
        g.set_tubes(t)
        ref = None
        for tl in TILES:
            g.set_deposit_tiles(*tl)
            assert g.deposit_tiles() == tl and g.deposit_tiles(N) == tl
            g.build_segments()
            g.bin()
            d = g.deposit()
            om = g.get_fields("omega")
            print(f"N={N} FP{prec} tiles {tl}: pairs {d['pairs']} hits {d['hits']}", flush=True)
            assert d["hits"] == hits
            assert d["pairs"] == kept[tl[:3]], tl
            if ref is None:
                ref = om
            elif prec == 32:
                assert np.array_equal(om, ref), tl
            else:
                assert np.linalg.norm(om - ref) <= 1e-13 * np.linalg.norm(ref), tl
        g.close()


def test_invalid_tiles_are_rejected():
    from fastweavturb_amd.api import WovenError, WovenGPU
    g = WovenGPU(32, 32)
    for tl in [(3, 16, 16, 8), (32, 32, 32, 8), (128, 4, 4, 8), (16, 16, 16, 0), (16, 16, 16, 65), (2, 16, 16, 8)]:
        with pytest.raises(WovenError):
            g.set_deposit_tiles(*tl)
    with pytest.raises(WovenError):
        g.set_deposit_tiles(16, 16, 16, 8, N=64)      # no instance of that size
    assert g.deposit_tiles() == (16, 16, 16, 8)
    g.close()


def test_batched_tiles_per_grid_size():
    """A batch whose grid sizes use different tiles equals the batch on the defaults bitwise
    (FP32), instance by instance, also when the whole generation is replayed from a graph."""
    from fastweavturb_amd.api import WovenGPU
    sizes = [48, 32, 64]

    def ctx():
        g = WovenGPU(precision=32, sizes=sizes)
        for i, N in enumerate(sizes):
            p, t = fbb.config_small(N=N, seed=500 + i, ntubes=(1, 2), npts=(40, 30), radii=(0.5, 0.3))
            g.select(i)
            g.set_params(p)
            g.set_tubes(t)
        return g

    ref = ctx()
    ref.generate()
    g = ctx()
    g.set_deposit_tiles(8, 8, 8, 4, N=32)
    g.set_deposit_tiles(32, 16, 8, 16, N=64)
    g.set_graph(2)
    for _ in range(3):
        g.generate()
        for i, N in enumerate(sizes):
            g.select(i)
            ref.select(i)
            assert g.deposit_tiles() == {48: (16, 16, 16, 8), 32: (8, 8, 8, 4), 64: (32, 16, 8, 16)}[N]
            for name in ("omega", "u"):
                assert np.array_equal(g.get_fields(name), ref.get_fields(name)), (N, name)
    g.close()
    ref.close()
