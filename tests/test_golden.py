"""Committed golden vectors (tests/golden/make_golden.py): the oracle must reproduce them
(CPU), and the GPU path must match them directly (GPU)."""
import os

import numpy as np
import pytest

from fastweavturb_amd.woven_io import RunParams, Tubes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load():
    z = np.load(os.path.join(GOLDEN, "woven_small16.npz"))
    pr = z["params"]
    p = RunParams(int(pr[0]), int(pr[1]), int(pr[2]), int(pr[3]), float(pr[4]), float(pr[5]), int(pr[6]),
                  int(pr[7]), float(pr[8]), float(pr[9]))
    t = Tubes(z["npoint"], z["cx"], z["cy"], z["cz"], z["sigma"], z["gamma"], z["n_T"],
              np.zeros(len(z["npoint"]), np.int32))
    return p, t, z


def test_fft_known_answers(oracle):
    z = np.load(os.path.join(GOLDEN, "fft_known.npz"))
    for n in (8, 12, 16, 24):
        np.testing.assert_allclose(oracle.fft1d(z[f"x{n}"], True), z[f"X{n}"], atol=1e-13)


def test_oracle_reproduces_golden(oracle):
    p, t, z = load()
    P = oracle.Params(p.N, p.n_d, p.line_s, p.line_e, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre,
                      p.Rtube_ratio, p.dzeta_sup)
    geo = oracle.Geometry(P, t.npoint, t.cx, t.cy, t.cz)
    np.testing.assert_array_equal(geo.ntf, z["ntf"])
    np.testing.assert_array_equal(geo.dtf, z["dtf"])
    np.testing.assert_array_equal(geo.ci, z["ci"])
    b = geo.bins(t.sigma, t.n_T)
    np.testing.assert_array_equal(b["lo"], z["lo"])
    np.testing.assert_array_equal(b["hi"], z["hi"])
    om_raw, hits = geo.deposit(t.sigma, t.gamma, t.n_T)
    assert hits == int(z["hits"][0])
    np.testing.assert_allclose(om_raw, z["omega_raw"], rtol=0, atol=1e-13 * np.abs(z["omega_raw"]).max())
    om, u, st = oracle.run(P, t.npoint, t.cx, t.cy, t.cz, t.sigma, t.gamma, t.n_T)
    np.testing.assert_allclose(u, z["u"], atol=1e-12 * np.abs(z["u"]).max())
    np.testing.assert_allclose(st[:4], z["stats"], rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_gpu_matches_golden(precision):
    from fastweavturb_amd.api import WovenGPU
    p, t, z = load()
    g = WovenGPU(p.N, precision)
    g.set_params(p)
    g.set_tubes(t)
    st = g.build_segments()
    ntf, dtf = g.sections()
    np.testing.assert_array_equal(ntf, z["ntf"])
    if precision == 64:
        np.testing.assert_array_equal(dtf, z["dtf"])
    lohi, R = g.bins()
    q = z["seg_q"]
    np.testing.assert_array_equal(lohi[q, :3], z["lo"])
    np.testing.assert_array_equal(lohi[q, 3:], z["hi"])
    g.bin()
    d = g.deposit()
    assert d["hits"] == int(z["hits"][0])
    tol = 1e-12 if precision == 64 else 1e-5
    om_raw = g.get_fields("omega")
    assert np.linalg.norm(om_raw - z["omega_raw"]) <= tol * np.linalg.norm(z["omega_raw"])
    g.biot_savart()
    u = g.get_fields("u")
    assert np.linalg.norm(u - z["u"]) <= tol * np.linalg.norm(z["u"])
    E = g.spectrum(1)
    big = z["E_u"] > 1e-9 * z["E_u"].max()
    np.testing.assert_allclose(E[big], z["E_u"][big], rtol=1e-10 if precision == 64 else 1e-4)
    S = g.structure_functions(0, 1, 6, z["r"])
    np.testing.assert_allclose(S, z["Sp"], rtol=1e-9 if precision == 64 else 1e-4, atol=1e-12)
