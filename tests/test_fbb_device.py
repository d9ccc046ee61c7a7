"""The fBb centerline generator on the GPU (wt_fbb_centerlines), FBB_with_parameters.m:73-120.

Pinned by: numpy's Philox bit generator (the device's random stream, word for word, and its
Box-Muller normals); fbb.py's construction on the same fixed random inputs (every embedding size
class: one-step FFTs up to m = 1024, four-step beyond, up to m = 32768); and properties of the
output that the construction fixes (closure and winding, the exact step rms, the r^2H scaling of
the bridge's increments, unit-variance fGn before scaling)."""
import numpy as np
import pytest

from fastweavturb_amd import fbb

pytestmark = pytest.mark.gpu


def ctx():
    from fastweavturb_amd.api import WovenGPU
    return WovenGPU(16, 64)


def minimg(d):
    return (d + np.pi) % (2 * np.pi) - np.pi


@pytest.mark.parametrize("seed,tube", [(1234, 0), (5, 77), (2 ** 40 + 3, 4095)])
def test_random_stream_is_numpys_philox(seed, tube):
    g = ctx()
    z = g.fbb_normals(seed, tube, 2000)
    g.close()
    raw = np.random.Philox(key=(tube << 64) | seed).random_raw(2000)
    u = (raw >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(1.0 - u[0::2]))
    ref = np.empty(2000)
    ref[0::2] = r * np.cos(2 * np.pi * u[1::2])
    ref[1::2] = r * np.sin(2 * np.pi * u[1::2])
    np.testing.assert_allclose(z, ref, rtol=1e-13, atol=1e-13)
    assert abs(z.mean()) < 0.1 and abs(z.std() - 1) < 0.05


def test_device_equals_host_construction_for_fixed_inputs():
    """Tubes of 8..16384 points in one call (grouped by embedding size on the device), with the
    normals and uniforms given: the device's positions equal fbb.tube_centerline_from's."""
    from fastweavturb_amd.api import fbb_embedding
    npts = [8, 100, 513, 600, 700, 1000, 3000, 16384, 600, 2500]
    H = [0.75, 0.55, 0.9, 0.6, 0.75, 0.8, 0.7, 0.75, 0.65, 0.5]
    W = [[1, 0, 0], [0, 1, 1], [1, -1, 1], [0, 0, -1], [1, 1, 0], [-1, 0, 1], [0, 1, 0], [1, 0, 1], [1, 1, 1],
         [0, -1, 0]]
    wiggle = 0.35
    rng = np.random.default_rng(11)
    gauss, host = [], []
    for n, h, w in zip(npts, H, W):
        m = fbb.embedding_size(n - 1)
        assert m == fbb_embedding(n)
        zre, zim, u3 = rng.standard_normal((2, m)), rng.standard_normal((2, m)), rng.random(3)
        gauss.append(np.concatenate([zre.ravel(), zim.ravel(), u3]))
        host.append(fbb.tube_centerline_from(n, h, np.array(w), wiggle, zre, zim, u3))
    g = ctx()
    cx, cy, cz = g.fbb_centerlines(npts, H, W, wiggle, seed=0, gauss=np.concatenate(gauss))
    g.close()
    off = np.concatenate([[0], np.cumsum(npts)])
    worst = 0.0
    for i, x in enumerate(host):
        dev = np.stack([cx[off[i]:off[i + 1]], cy[off[i]:off[i + 1]], cz[off[i]:off[i + 1]]])
        assert np.all(dev >= -np.pi) and np.all(dev < np.pi)
        worst = max(worst, float(np.abs(minimg(dev - x)).max()))
    print(f"fbb device vs host, fixed inputs: max |dx| {worst:.2e}", flush=True)
    assert worst <= 1e-10


def test_device_paths_close_wind_and_scale():
    """Philox-drawn paths: the unwrapped displacement over a period is 2 pi w, the bridge's step
    rms is wiggle * 2 pi |w| / n, and S2(r) of its increments grows as r^2H."""
    n, nt, wiggle = 8192, 24, 0.35
    for H in (0.55, 0.85):
        W = [[1, 0, 0]] * (nt // 2) + [[1, 1, -1]] * (nt // 2)
        g = ctx()
        cx, cy, cz = g.fbb_centerlines([n] * nt, [H] * nt, W, wiggle, seed=99)
        g.close()
        X = np.stack([cx, cy, cz]).reshape(3, nt, n).transpose(1, 0, 2)
        slopes = []
        for k in range(nt):
            w = np.array(W[k], dtype=np.float64)
            d = minimg(np.diff(np.concatenate([X[k], X[k][:, :1]], axis=1), axis=1))
            np.testing.assert_allclose(d.sum(axis=1), 2 * np.pi * w, atol=1e-9)
            steps = d[:, :-1] - (2 * np.pi * w / n)[:, None]          # bridge increments
            rms = np.sqrt(np.sum(steps ** 2) / (n - 1))
            assert rms == pytest.approx(wiggle * 2 * np.pi * np.linalg.norm(w) / n, rel=1e-9)
            B = np.cumsum(steps, axis=1)
            rs = np.array([1, 4, 16, 64])
            S2 = np.array([np.mean((B[:, r:] - B[:, :-r]) ** 2) for r in rs])
            slopes.append(np.polyfit(np.log(rs), np.log(S2), 1)[0])
        print(f"fbb device H={H}: S2 slope {np.mean(slopes):.3f}", flush=True)
        assert np.mean(slopes) == pytest.approx(2 * H, abs=0.06)


def test_device_config_matches_its_host_twin_statistically():
    """woven_config_device against woven_config on the 512^3 level layout's first three levels:
    the same point counts and circulations per tube class, fields of the same character."""
    from fastweavturb_amd.api import WovenGPU
    levels = [fbb.Level(0.8 * 2.0 ** -l, 4 * 4 ** l, (256 * 2 ** l, 256 * 2 ** (l + 1))) for l in range(3)]
    g = WovenGPU(64, 32)
    p, t = fbb.woven_config_device(g, 64, levels, 1234)
    assert t.nline == 84 and np.all(t.npoint >= 256)
    g.set_params(p)
    g.set_tubes(t)
    etot, up, ens = g.generate()
    assert np.isfinite(up) and up > 0
    assert g.spectrum(1).sum() == pytest.approx(1.5, rel=0.02)
    g.close()


def test_reference_mode_equals_host_construction():
    """wt_fbb_reference_centerlines (the reference's construction: level dB, no drift, tubes below
    128 points sampled from a 128-point path, mod(x, 2 pi) - pi) against
    fbb.reference_centerline_from: on given inputs at 8..3000 points, and on the device's own
    Philox stream for the family of the reference's shipped input pair (73 tubes)."""
    npts = [8, 20, 64, 100, 127, 128, 148, 1000, 3000]
    H = [0.834, 0.6, 0.75, 0.9, 0.55, 0.834, 0.834, 0.7, 0.65]
    dB = [2.4, 1.2, 0.6, 0.3, 0.5, 1.0, 0.6043, 0.05, 0.01]
    rng = np.random.default_rng(13)
    gauss, host = [], []
    for n, h, d in zip(npts, H, dB):
        m = fbb.embedding_size(max(n, 128) - 1)
        zre, zim, u3 = rng.standard_normal((2, m)), rng.standard_normal((2, m)), rng.random(3)
        gauss.append(np.concatenate([zre.ravel(), zim.ravel(), u3]))
        host.append(fbb.reference_centerline_from(n, h, d, zre, zim, u3))
    g = ctx()
    cx, cy, cz = g.fbb_reference_centerlines(npts, H, dB, seed=0, gauss=np.concatenate(gauss))
    fam = fbb.ReferenceFamily(rho_w_obj=0.019)
    _, td = fbb.reference_inputs(fam, seed=7, g=g)
    g.close()
    off = np.concatenate([[0], np.cumsum(npts)])
    worst = 0.0
    for i, x in enumerate(host):
        dev = np.stack([cx[off[i]:off[i + 1]], cy[off[i]:off[i + 1]], cz[off[i]:off[i + 1]]])
        assert np.all(dev >= -np.pi) and np.all(dev < np.pi)
        worst = max(worst, float(np.abs(minimg(dev - x)).max()))
    _, th = fbb.reference_inputs(fam, seed=7)
    fam_worst = max(float(np.abs(minimg(a - b)).max()) for a, b in ((td.cx, th.cx), (td.cy, th.cy), (td.cz, th.cz)))
    print(f"fbb reference mode, device vs host: given inputs {worst:.2e}, Philox family {fam_worst:.2e}", flush=True)
    assert worst <= 1e-10 and fam_worst <= 1e-10
    assert np.array_equal(td.npoint, th.npoint) and np.array_equal(td.sigma, th.sigma)
