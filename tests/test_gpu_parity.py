"""Parity of the HIP path (through the C ABI) against the FP64 oracle, stage by stage.

Bars: integers (ntf, bounding boxes, hit counts) and FP64 geometry bit-exact; FP64 fields
within 1e-12 relative L2; FP32 vorticity and velocity within 1e-5 relative L2; spectrum
shells within 1e-4 relative (FP32) / 1e-10 (FP64).
"""
import os

import numpy as np
import pytest

from binning_rule import slab_pairs
from fastweavturb_amd import fbb
from fastweavturb_amd.woven_io import Tubes, read_inputs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def oracle_params(O, p, t):
    return O.Params(p.N, p.n_d, p.line_s, p.line_e, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre,
                    p.Rtube_ratio, p.dzeta_sup)


def gpu_ctx(p, t, precision):
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(p.N, precision)
    g.set_params(p)
    g.set_tubes(t)
    return g


CASES = {
    "small32": lambda: fbb.config_small(N=32),
    "ratio4_eta": lambda: fbb.config_small(N=32, seed=11, sigma_ratio=4.0, ntubes=(2, 3), npts=(40, 30),
                                           radii=(0.25, 0.12)),
    "n24_radix3": lambda: fbb.config_small(N=24, seed=5, ntubes=(1, 2), npts=(50, 40), radii=(0.7, 0.4)),
}


def _case(name):
    p, t = CASES[name]()
    if name == "ratio4_eta":
        p.eta_t = 0.7
    return p, t


@pytest.mark.parametrize("name", list(CASES))
def test_segments_bitexact(oracle, name):
    p, t = _case(name)
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    g = gpu_ctx(p, t, 64)
    st = g.build_segments()
    assert st["nsec_all"] == len(geo.ntf)
    assert st["nseg_all"] == geo.nseg
    ntf, dtf = g.sections()
    np.testing.assert_array_equal(ntf, geo.ntf)
    np.testing.assert_array_equal(dtf.view(np.int64), geo.dtf.view(np.int64))
    ci, s = g.subpoints()
    assert ci.shape == geo.ci.shape
    np.testing.assert_array_equal(ci.view(np.int64), geo.ci.view(np.int64))
    np.testing.assert_array_equal(s.view(np.int64), geo.s.view(np.int64))
    assert st["ntf_max"] == geo.ntf.max()
    assert st["dtf_min"] == geo.dtf.min()


@pytest.mark.parametrize("ntf_pre", [1, 2, 7, 101, 1000])
def test_turning_angle_sample_counts(oracle, ntf_pre):
    """The FP32 turning angle takes the ntf_pre + 1 samples in symmetric pairs (odd and even
    ntf_pre pair them differently); ntf and dtf must stay the reference's, and most sections must
    be decided without the exact re-evaluation."""
    p, t = fbb.config_small(N=32, seed=23, ntubes=(2, 3), npts=(80, 60), radii=(0.5, 0.3), ntf_pre=ntf_pre)
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    g = gpu_ctx(p, t, 64)
    st = g.build_segments()
    ntf, dtf = g.sections()
    np.testing.assert_array_equal(ntf, geo.ntf)
    np.testing.assert_array_equal(dtf.view(np.int64), geo.dtf.view(np.int64))
    assert st["theta_exact_sections"] < 0.2 * st["nsec_all"], st


@pytest.mark.parametrize("name", list(CASES))
def test_bins_bitexact(oracle, name):
    p, t = _case(name)
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    ob = geo.bins(t.sigma, t.n_T)
    g = gpu_ctx(p, t, 64)
    g.build_segments()
    lohi, R = g.bins()
    q = geo.sp_base[geo.seg_sec] + geo.seg_j
    np.testing.assert_array_equal(lohi[q, :3], ob["lo"])
    np.testing.assert_array_equal(lohi[q, 3:], ob["hi"])
    if p.sigma_ratio == 1.0:
        np.testing.assert_array_equal(R[q].view(np.int64), ob["R"].view(np.int64))
    else:
        # Rtube = Rtube_ratio*sigma(s) uses sin (subroutine.f90:172): the device libm and the
        # host libm may differ by an ulp or two; the integer boxes above must still agree.
        np.testing.assert_allclose(R[q], ob["R"], rtol=5e-16, atol=0)
    mask = np.ones(len(R), bool)
    mask[q] = False
    assert np.all(R[mask] == -1.0)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [64, 32])
def test_deposit(oracle, name, precision):
    p, t = _case(name)
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    om_ref, hits = geo.deposit(t.sigma, t.gamma, t.n_T)
    g = gpu_ctx(p, t, precision)
    g.build_segments()
    g.bin()
    d = g.deposit()
    assert d["hits"] == hits, "the set of depositing (cell, sub-segment) pairs must be the reference's"
    om = g.get_fields("omega")
    err = rel_l2(om, om_ref)
    assert err <= (1e-12 if precision == 64 else 1e-5), err


@pytest.mark.parametrize("precision", [64, 32])
def test_deposit_wide_cores(oracle, precision):
    """Boxes of 7 bricks per axis (more than 256 (brick, image) pairs per sub-segment) are binned by
    a wave per sub-segment; the depositing pairs and the field must still be the reference's."""
    p, t = fbb.config_small(N=128, seed=17, ntubes=(1, 1), npts=(24, 20), radii=(0.9, 0.5), ntf_pre=100)
    t.sigma = np.array([0.8, 0.3])
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    om_ref, hits = geo.deposit(t.sigma, t.gamma, t.n_T)
    g = gpu_ctx(p, t, precision)
    g.build_segments()
    g.bin()
    box, R = g.bins()
    valid = R >= 0.0
    widest = int((box[valid, 3:] - box[valid, :3]).max())
    assert widest > 6 * 16, widest
    d = g.deposit()
    assert d["hits"] == hits
    assert d["pairs"] > 256 * 10
    err = rel_l2(g.get_fields("omega"), om_ref)
    assert err <= (1e-12 if precision == 64 else 1e-5), err


def _planar_tubes(n=48):
    """Circles in the x-y and y-z planes and a straight tube along x: many chords have exact zero
    components, so end planes parallel to a column axis and chords along an axis both occur."""
    th = 2 * np.pi * np.arange(n) / n
    cx = np.concatenate([3.0 + 1.2 * np.cos(th), np.full(n, 4.5), 2 * np.pi * np.arange(n) / n])
    cy = np.concatenate([2.5 + 1.2 * np.sin(th), 3.0 + 1.0 * np.cos(th), np.full(n, 5.0)])
    cz = np.concatenate([np.full(n, 1.75), 3.5 + 1.4 * np.sin(th), np.full(n, 0.6)])
    return Tubes(np.full(3, n, np.int32), cx, cy, cz, np.array([0.3, 0.25, 0.2]), np.array([1.0, -0.7, 0.5]),
                 np.ones(3, np.int32), np.zeros(3, np.int32))


@pytest.mark.parametrize("precision", [64, 32])
def test_deposit_planar_tubes(oracle, precision):
    p, _ = fbb.config_small(N=64)
    p.line_s, p.line_e = 1, 3
    t = _planar_tubes()
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    tci = geo.tci[np.linalg.norm(geo.tci, axis=1) > 0]
    ax = np.argmax(np.abs(tci), 1)
    r = np.arange(len(tci))
    B = tci[r, (ax + 1) % 3]
    C = tci[r, (ax + 2) % 3]
    assert ((B == 0) & (C != 0)).sum() > 50 and ((B == 0) & (C == 0)).sum() > 50
    om_ref, hits = geo.deposit(t.sigma, t.gamma, t.n_T)
    g = gpu_ctx(p, t, precision)
    g.build_segments()
    g.bin()
    d = g.deposit()
    assert d["hits"] == hits
    err = rel_l2(g.get_fields("omega"), om_ref)
    assert err <= (1e-12 if precision == 64 else 1e-5), err


def test_wide_box_list_growth(oracle):
    """More wide boxes (over 256 pairs each) than the list's first size: the list grows and the pair
    lists stay complete (pair count equal to the host restatement of the binning; hits equal the
    oracle's; FP32 and FP64 agree)."""
    p, t = fbb.config_small(N=128, seed=29, ntubes=(1,), npts=(400,), radii=(0.9,), ntf_pre=100)
    t.sigma = np.array([0.8])
    P = oracle_params(oracle, p, t)
    geo = oracle.Geometry(P, t.npoint, t.cx, t.cy, t.cz)
    kept, _ = slab_pairs(geo, geo.bins(t.sigma, t.n_T), 128)
    _, hits = geo.deposit(t.sigma, t.gamma, t.n_T)
    out = {}
    for precision in (64, 32):
        g = gpu_ctx(p, t, precision)
        g.build_segments()
        g.bin()
        _, R = g.bins()
        valid = R >= 0.0
        d = g.deposit()
        assert d["pairs"] == kept
        assert d["hits"] == hits
        out[precision] = (d["hits"], g.get_fields("omega"))
        g.close()
    assert int(valid.sum()) > 1024
    assert out[64][0] == out[32][0]
    assert rel_l2(out[32][1], out[64][1]) < 1e-5


@pytest.mark.parametrize("precision", [64, 32])
def test_biot_savart_from_oracle_vorticity(oracle, precision):
    p, t = _case("small32")
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    om, _ = geo.deposit(t.sigma, t.gamma, t.n_T)
    u_ref = oracle.biot_savart(om)
    om_ref = om.copy()
    st_ref = oracle.normalize(om_ref, u_ref)
    E_ref = oracle.spectrum(u_ref)
    Ew_ref = oracle.spectrum(om)
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(p.N, precision)
    g.set_fields("omega", om if precision == 64 else om.astype(np.float32))
    etot, up, ens = g.biot_savart()
    tol = 1e-12 if precision == 64 else 1e-5
    assert abs(etot - st_ref[0]) <= tol * st_ref[0] * 10
    assert abs(up - st_ref[1]) <= tol * st_ref[1] * 10
    assert abs(ens - st_ref[2]) <= tol * st_ref[2] * 10
    assert rel_l2(g.get_fields("u"), u_ref) <= tol
    assert rel_l2(g.get_fields("omega"), om_ref) <= tol
    E = g.spectrum(1)
    Ew = g.spectrum(0)
    stol = 1e-10 if precision == 64 else 1e-4
    big = E_ref > 1e-9 * E_ref.max()
    np.testing.assert_allclose(E[big], E_ref[big], rtol=stol)
    bigw = Ew_ref > 1e-9 * Ew_ref.max()
    np.testing.assert_allclose(Ew[bigw], Ew_ref[bigw], rtol=stol)


@pytest.mark.parametrize("N", [8, 12, 16, 48])
def test_biot_savart_random_field_sizes(oracle, N):
    """Random (non-solenoidal) vorticity including Nyquist content: the half-spectrum path
    with the Nyquist rule must equal the reference's full c2c + real part."""
    rng = np.random.default_rng(N)
    om = rng.standard_normal((3, N, N, N))
    u_ref = oracle.biot_savart(om)
    om_ref = om.copy()
    oracle.normalize(om_ref, u_ref)
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(N, 64)
    g.set_fields("omega", om)
    g.biot_savart()
    assert rel_l2(g.get_fields("u"), u_ref) <= 1e-12


def test_spectrum_and_sf_of_given_velocity(oracle):
    N = 16
    rng = np.random.default_rng(3)
    u = rng.standard_normal((3, N, N, N))
    from fastweavturb_amd.api import WovenGPU, sp_separations
    g = WovenGPU(N, 64)
    g.set_fields("u", u)
    E = g.spectrum(1)
    np.testing.assert_allclose(E, oracle.spectrum(u), rtol=1e-11)
    seps = sp_separations(N)
    np.testing.assert_array_equal(seps, oracle.sp_separations(N))
    S = g.structure_functions(0, 1, 10, seps)
    Sref, _ = oracle.sp_calculate(u, 10)
    np.testing.assert_allclose(S, Sref, rtol=1e-10, atol=1e-12 * np.abs(Sref).max())
    xs = np.arange(1, N // 2 + 1)
    Sx = g.structure_functions(1, 2, 6, xs)
    Sx_ref = oracle.sf_x(u[0], 2, 6, xs)
    # odd orders at r = N/2 vanish identically (the increments pair up as +d, -d): compare those
    # against the scale of the whole table
    np.testing.assert_allclose(Sx, Sx_ref, rtol=1e-11, atol=1e-13 * np.abs(Sx_ref).max())


@pytest.mark.parametrize("precision", [64, 32])
def test_pipeline_end_to_end(oracle, precision):
    p, t = _case("small32")
    om_ref, u_ref, st = oracle.run(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz, t.sigma,
                                   t.gamma, t.n_T)
    g = gpu_ctx(p, t, precision)
    etot, up, ens = g.generate()
    tol = 1e-12 if precision == 64 else 1e-5
    assert rel_l2(g.get_fields("omega"), om_ref) <= tol
    assert rel_l2(g.get_fields("u"), u_ref) <= tol
    assert abs(up - st[1]) <= 10 * tol * st[1]
    assert abs(ens - st[2]) <= 10 * tol * st[2]
    ms = g.stage_times()
    assert np.all(ms[:4] >= 0) and ms[:4].sum() > 0


def test_fp32_deposition_is_bitwise_reproducible():
    p, t = _case("ratio4_eta")
    a = gpu_ctx(p, t, 32)
    a.generate()
    b = gpu_ctx(p, t, 32)
    b.generate()
    for name in ("omega", "u"):
        np.testing.assert_array_equal(a.get_fields(name, fp64=False), b.get_fields(name, fp64=False))


def test_wt2_reference_input_geometry(oracle):
    """The reference's shipped WT2 case (73 tubes x 148 points, N=256): geometry bit-exact."""
    p, t = read_inputs(os.path.join(GOLDEN, "wt2"))
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    g = gpu_ctx(p, t, 64)
    st = g.build_segments()
    ntf, dtf = g.sections()
    np.testing.assert_array_equal(ntf, geo.ntf)
    np.testing.assert_array_equal(dtf.view(np.int64), geo.dtf.view(np.int64))
    ci, _ = g.subpoints()
    np.testing.assert_array_equal(ci.view(np.int64), geo.ci.view(np.int64))
    lohi, R = g.bins()
    ob = geo.bins(t.sigma, t.n_T)
    q = geo.sp_base[geo.seg_sec] + geo.seg_j
    np.testing.assert_array_equal(lohi[q, :3], ob["lo"])
    np.testing.assert_array_equal(lohi[q, 3:], ob["hi"])
    assert st["nseg_all"] == geo.nseg


def test_wt2_deposit_subset(oracle):
    """WT2 at N=64 (reduced grid), tubes 1..4 (one per level group): hit set and field parity."""
    p, t = read_inputs(os.path.join(GOLDEN, "wt2"))
    p.N = p.n_d = 64
    p.line_s, p.line_e = 1, 4
    geo = oracle.Geometry(oracle_params(oracle, p, t), t.npoint, t.cx, t.cy, t.cz)
    om_ref, hits = geo.deposit(t.sigma, t.gamma, t.n_T)
    for prec, tol in ((64, 1e-12), (32, 1e-5)):
        g = gpu_ctx(p, t, prec)
        g.build_segments()
        g.bin()
        d = g.deposit()
        assert d["hits"] == hits
        assert rel_l2(g.get_fields("omega"), om_ref) <= tol
