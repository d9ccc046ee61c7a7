"""The reference's own centerline construction (FBB_with_parameters.m), restated in
fbb.reference_levels / reference_centerline_from / reference_inputs.

Pinned by the reference's shipped input pair (tests/golden/wt2, written by that script): the level
scalings reproduce every line of its parameters_input.dat and the tube and point counts of its
centerline_input.dat (for rho_w_obj = 0.019, the one input its counts fix: the script's header
value 0.024 gives 187 points per tube, the shipped files have 148), and its centerlines' step rms
equals the level's dB; by what the construction fixes (the bridge pins the last point to the
shift alone, the step rms is exactly dB); and by a loop-by-loop restatement of the script's
sub-sampling of short tubes (FBB:78-84)."""
import math
import os

import numpy as np
import pytest

from fastweavturb_amd import fbb
from fastweavturb_amd.woven_io import read_centerlines, write_inputs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "wt2")
SHIPPED = fbb.ReferenceFamily(rho_w_obj=0.019)


def test_level_scalings_reproduce_the_shipped_parameter_file(tmp_path):
    p, t = fbb.reference_inputs(SHIPPED, seed=3)
    write_inputs(str(tmp_path), p, t, reference_digits=True)
    ours = open(tmp_path / "parameters_input.dat").read().splitlines()
    ref = open(os.path.join(GOLDEN, "parameters_input.dat")).read().splitlines()
    assert ours == ref
    c_ours = open(tmp_path / "centerline_input.dat").read().splitlines()
    c_ref = open(os.path.join(GOLDEN, "centerline_input.dat")).read().splitlines()
    assert c_ours[:1 + t.nline] == c_ref[:1 + t.nline] and len(c_ours) == len(c_ref)
    # the script's own header value gives other point counts
    assert fbb.reference_levels(fbb.ReferenceFamily())["Np"] == [187, 187, 187]


def test_shipped_centerlines_have_the_level_step_rms():
    """Step rms of the shipped (wrapped, six-digit) centerlines of levels 2 and 3 against the
    level's dB = dB_1 r_s^(i-1); level 1's steps (dB = 2.42 per point over three axes) exceed the
    half box, so its wrapped differences do not give its steps back."""
    npoint, cx, cy, cz = read_centerlines(os.path.join(GOLDEN, "centerline_input.dat"))
    lv = fbb.reference_levels(SHIPPED)
    off = np.concatenate([[0], np.cumsum(npoint)])
    X = np.stack([cx, cy, cz])
    assert X.min() >= -np.pi and X.max() < np.pi
    level = np.repeat(np.arange(3), lv["m_line"])
    for l in range(len(npoint)):
        if level[l] == 0:
            continue
        d = np.diff(X[:, off[l]:off[l + 1]], axis=1)
        d = (d + np.pi) % (2 * np.pi) - np.pi
        rms = math.sqrt(float(np.sum(d ** 2)) / (npoint[l] - 1))
        assert rms == pytest.approx(lv["dB"][level[l]], rel=3e-5), l


def _subsample_literal(P128, Np):
    """FBB:78-84 as the script writes it, 1-based: FBM(i_t) = FBM_temp(ceil(i_t * r_t))."""
    r_t = 128 / Np
    out = np.empty((3, Np))
    for co in range(3):
        for i_t in range(1, Np + 1):
            out[co, i_t - 1] = P128[co, math.ceil(i_t * r_t) - 1]
    return out


@pytest.mark.parametrize("n", [8, 20, 64, 100, 127, 128, 148, 1000])
def test_reference_construction(n):
    H, dB = 0.834, 0.6043
    m = fbb.embedding_size(max(n, 128) - 1)
    zre, zim, u3 = fbb.philox_inputs(11, n, m)
    x = fbb.reference_centerline_from(n, H, dB, zre, zim, u3)
    assert x.shape == (3, n) and x.min() >= -np.pi and x.max() < np.pi
    # the bridge pins the last point to 0: it is the wrapped shift alone
    x0 = (u3 - 0.5) * 2.0 * np.pi
    np.testing.assert_array_equal(x[:, -1], (x0 - np.floor(x0 / (2 * np.pi)) * (2 * np.pi)) - np.pi)
    d = np.diff(x, axis=1)
    d = (d + np.pi) % (2 * np.pi) - np.pi
    assert math.sqrt(float(np.sum(d ** 2)) / (n - 1)) == pytest.approx(dB, rel=1e-12)
    if n < 128:
        g = fbb.fgn_from_normals(127, H, zre, zim, 3)
        P = np.concatenate([np.zeros((3, 1)), np.cumsum(g, axis=1)], axis=1)
        F = _subsample_literal(P, n)
        j = np.arange(1, n + 1) / n
        bb = F - 0.5 * ((j ** (2 * H) + 1) - (1 - j) ** (2 * H)) * F[:, -1:]
        s = np.diff(bb, axis=1)
        bb *= dB / math.sqrt(np.sum(s ** 2) / (n - 1))
        want = bb + x0[:, None]
        want = (want - np.floor(want / (2 * np.pi)) * (2 * np.pi)) - np.pi
        np.testing.assert_allclose(x, want, rtol=0, atol=1e-13)


def test_philox_inputs_are_numpys_stream():
    zre, zim, u3 = fbb.philox_inputs(5, 7, 16)
    raw = np.random.Philox(key=(7 << 64) | 5).random_raw(67)
    u = (raw >> np.uint64(11)).astype(np.float64) / 2.0 ** 53
    assert np.array_equal(u3, u[64:67])
    z0 = math.sqrt(-2 * math.log(1 - u[0])) * math.cos(2 * math.pi * u[1])
    assert zre[0, 0] == z0 and zre.shape == (2, 16) and zim.shape == (2, 16)


def test_command_line_writes_the_shipped_parameter_file(tmp_path):
    fbb.main(["--out", str(tmp_path), "--seed", "3", "--rho-w-obj", "0.019", "--host"])
    assert (open(tmp_path / "parameters_input.dat").read() ==
            open(os.path.join(GOLDEN, "parameters_input.dat")).read())
    from fastweavturb_amd.woven_io import read_inputs
    p, t = read_inputs(str(tmp_path))
    assert t.nline == 73 and t.npoint_all == 73 * 148 and p.N == 256 and p.sigma_ratio == 4.0
