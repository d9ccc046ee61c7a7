"""CPU-only checks of the host side: input formats, the fBb generator, the C ABI library
(loads and exports every declared symbol; host-only functions), and error paths."""
import os

import numpy as np
import pytest

from fastweavturb_amd import fbb
from fastweavturb_amd.woven_io import read_centerlines, read_inputs, read_parameters, write_inputs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_read_wt2_reference_inputs():
    p, t = read_inputs(os.path.join(GOLDEN, "wt2"))
    assert (p.N, p.n_d, p.line_s, p.line_e) == (256, 256, 1, 73)
    assert (p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre, p.Rtube_ratio, p.dzeta_sup) == (4.0, 0.0, 50, 1000, 3.0, 1e-4)
    assert t.nline == 73 and np.all(t.npoint == 148)
    assert t.npoint_all == 10804
    assert t.sigma[0] == pytest.approx(0.0604305) and t.gamma[1] == pytest.approx(0.39685)
    assert t.n_T[0] == 1 and t.n_T[1] == 2 and np.all(t.if_filter == 0)
    assert t.cx[0] == pytest.approx(0.389573) and t.cz[2] == pytest.approx(-2.50238)
    assert np.all(np.abs(t.cx) <= np.pi)


def test_inputs_round_trip(tmp_path):
    p, t = fbb.config_small(N=32)
    write_inputs(tmp_path, p, t)
    p2, t2 = read_inputs(tmp_path)
    assert p2 == p
    np.testing.assert_array_equal(t2.npoint, t.npoint)
    np.testing.assert_allclose(t2.cx, t.cx, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(t2.sigma, t.sigma, rtol=1e-9)
    np.testing.assert_array_equal(t2.n_T, t.n_T)


def test_truncated_and_malformed_inputs(tmp_path):
    with pytest.raises(ValueError):
        read_centerlines("")
    with pytest.raises(ValueError):
        read_centerlines("2\n3\n3\n0 0 0\n1 1 1\n")
    p, t = fbb.config_small(N=32)
    write_inputs(tmp_path, p, t)
    text = open(tmp_path / "parameters_input.dat").read().splitlines()
    with pytest.raises(ValueError):
        read_parameters("\n".join(text[:-1]), t.nline)
    with pytest.raises(ValueError):
        read_parameters("\n".join(text[:5]), t.nline)
    t.cx[3] = np.nan
    with pytest.raises(ValueError):
        t.validate()


def test_fgn_covariance():
    """Davies-Harte fGn: unit variance and lag-1 correlation (2^{2H} - 2)/2."""
    rng = np.random.default_rng(0)
    H = 0.75
    x = fbb.fgn_davies_harte(4096, H, rng, 64)
    assert x.var() == pytest.approx(1.0, rel=0.03)
    c1 = np.mean(x[:, 1:] * x[:, :-1])
    assert c1 == pytest.approx(0.5 * (2 ** (2 * H) - 2), abs=0.02)


def test_fbm_increment_scaling():
    """Second-order structure function of fBm increments grows as r^{2H}
    (the check of FBB_with_parameters.m:150-165)."""
    rng = np.random.default_rng(1)
    H = 0.75
    g = fbb.fgn_davies_harte(8192, H, rng, 16)
    B = np.cumsum(g, axis=1)
    rs = np.array([1, 4, 16, 64])
    S2 = np.array([np.mean((B[:, r:] - B[:, :-r]) ** 2) for r in rs])
    slope = np.polyfit(np.log(rs), np.log(S2), 1)[0]
    assert slope == pytest.approx(2 * H, abs=0.05)


def test_bridge_closes_and_winds():
    rng = np.random.default_rng(2)
    bb = fbb.fbb_path(500, 0.75, rng)
    assert np.allclose(bb[:, -1], 0.0)
    w = fbb.lattice_winding(np.array([0.9, -0.2, 0.55]))
    assert list(w) == [1, 0, 1]
    x = fbb.tube_centerline(400, 0.75, w, 0.35, rng)
    assert np.all(x >= -np.pi) and np.all(x < np.pi)
    # the unwrapped displacement over one period is 2*pi*w
    d = np.diff(np.concatenate([x, x[:, :1]], axis=1), axis=1)
    d = (d + np.pi) % (2 * np.pi) - np.pi
    np.testing.assert_allclose(d.sum(axis=1), 2 * np.pi * w, atol=1e-9)


def test_config_shapes():
    p, t = fbb.config_plain_128()
    assert p.N == 128 and t.nline == 8 + 32 + 128 and np.all(t.npoint == 1024)
    radii = np.unique(np.round(t.sigma * np.sqrt(2), 12))
    np.testing.assert_allclose(sorted(radii), [0.1, 0.2, 0.4])
    r = t.sigma * np.sqrt(2)
    ratio = t.gamma / r ** (4 / 3)
    assert np.all((ratio >= 0.8) & (ratio <= 1.2))


def test_library_exports_every_declared_symbol():
    from fastweavturb_amd import api
    names = api.header_functions()
    assert len(names) >= 20
    L = api.load_library()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_host_only_functions():
    from fastweavturb_amd import api
    assert api.nek(512) == 256 and api.nek(96) == 48
    assert list(api.sp_separations(512)) == [1, 1, 2, 3, 4, 6, 8, 11, 16, 23, 32, 45, 64, 91, 128, 181, 256, 362, 512]


def test_create_rejects_bad_arguments_without_a_device():
    from fastweavturb_amd.api import WovenError, WovenGPU
    with pytest.raises(WovenError) as e:
        WovenGPU(100, 32)          # 100 = 4 * 25: radix 5 is not supported
    assert e.value.code == -5
    with pytest.raises(WovenError) as e:
        WovenGPU(64, 16)
    assert e.value.code == -1


def test_product_never_imports_the_oracle():
    """The product (package sources) must not import, load or link the oracle."""
    import re
    root = os.path.join(os.path.dirname(os.path.dirname(__file__)), "fastweavturb_amd")
    pat = re.compile(r"(^\s*(from|import)\s+oracle)|liboracle|oracle_py|woven_oracle", re.M)
    for dirpath, _, files in os.walk(root):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".f90", "Makefile")):
                text = open(os.path.join(dirpath, f)).read()
                assert not pat.search(text), f


def test_centerline_from_given_inputs_equals_the_seeded_draw():
    """tube_centerline_from on the normals and uniforms tube_centerline draws gives its output
    bitwise (the form the device generator is checked against)."""
    n, H, w = 700, 0.7, np.array([1, -1, 0])
    x = fbb.tube_centerline(n, H, w, 0.35, np.random.default_rng(5))
    rng = np.random.default_rng(5)
    m = fbb.embedding_size(n - 1)
    zre, zim, u3 = rng.standard_normal((2, m)), rng.standard_normal((2, m)), rng.random(3)
    np.testing.assert_array_equal(fbb.tube_centerline_from(n, H, w, 0.35, zre, zim, u3), x)
