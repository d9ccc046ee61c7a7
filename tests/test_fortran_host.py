"""The Fortran drop-in host (ISO_C_BINDING over the C ABI) reads the reference's two input
files and writes the reference's output files; its numbers must match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from fastweavturb_amd import fbb
from fastweavturb_amd.woven_io import write_inputs

pytestmark = pytest.mark.gpu

HOST = os.path.join(os.path.dirname(os.path.dirname(__file__)), "fastweavturb_amd", "fortran", "weave_host")


def run_host(workdir, precision):
    r = subprocess.run([HOST, str(precision)], cwd=workdir, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("precision", [64, 32])
def test_fortran_host_outputs_match_oracle(oracle, tmp_path, precision):
    p, t = fbb.config_small(N=16, seed=21, ntubes=(1, 2), npts=(30, 24), radii=(0.9, 0.5), ntf_pre=200)
    write_inputs(tmp_path, p, t)
    run_host(tmp_path, precision)
    # the oracle reads the same numbers back from the written files
    from fastweavturb_amd.woven_io import read_inputs
    p2, t2 = read_inputs(tmp_path)
    P = oracle.Params(p2.N, p2.n_d, p2.line_s, p2.line_e, p2.sigma_ratio, p2.eta_t, p2.ntf_times, p2.ntf_pre,
                      p2.Rtube_ratio, p2.dzeta_sup)
    om, u, st = oracle.run(P, t2.npoint, t2.cx, t2.cy, t2.cz, t2.sigma, t2.gamma, t2.n_T)
    E = np.loadtxt(tmp_path / "output/stat/spectrum.dat")
    Eref = oracle.spectrum(u)
    assert E.shape == (oracle.nek(16), 2)
    big = Eref > 1e-6 * Eref.max()
    np.testing.assert_allclose(E[big, 1], Eref[big], rtol=2e-5)
    Sp = np.loadtxt(tmp_path / "output/stat/Sp.dat")
    Sref, r = oracle.sp_calculate(u, 10)
    np.testing.assert_allclose(Sp[:, 0], r * 2 * np.pi / 16, rtol=1e-5)
    np.testing.assert_allclose(Sp[:, 1:].T, Sref, rtol=5e-5, atol=1e-9)
    from oracle import pdf_oracle as PO
    for fname, fld in (("pdf_um.dat", u), ("pdf_wm.dat", om)):
        pd = np.loadtxt(tmp_path / "output/stat" / fname)
        rv, rp, rc = PO.cal_pdf(fld.astype(np.float32), 1000, (16, 16, 48))
        assert pd.shape == (1000, 2)
        span = float(rv[-1] - rv[0])
        np.testing.assert_allclose(pd[:, 0], rv, rtol=0, atol=1e-4 * span)
        bs = span / 999
        assert abs(pd[:, 1].sum() * bs - rc.sum() / fld.size) < 1e-3
    wh = np.fromfile(tmp_path / "output/field_wh.dat", dtype=np.float32).reshape(2, 16, 16, 16)
    wmag = np.sqrt((om ** 2).sum(axis=0))
    hel = (om * u).sum(axis=0)
    # field_wh samples i*r_d with r_d = N/n_d = 1: Fortran index i -> numpy index i-1, i = 1..N
    np.testing.assert_allclose(wh[0], wmag, rtol=1e-5, atol=1e-5 * wmag.max())
    np.testing.assert_allclose(wh[1], hel, rtol=1e-4, atol=1e-5 * np.abs(hel).max())
    td = open(tmp_path / "output/test_data.dat").read().split()
    assert td.count("ntf_max=") == t.nline
    pout = open(tmp_path / "output/parameters_output.dat").read().split()
    assert pout[1] == "16"
    # spectrum of the vorticity before normalisation (tube.f90:972-978): omega_raw = up * omega
    Ew = np.loadtxt(tmp_path / "output/stat/spectrum_vor_temp.dat")
    Ewref = oracle.spectrum(om) * st[1] ** 2
    big = Ewref > 1e-6 * Ewref.max()
    np.testing.assert_allclose(Ew[big, 1], Ewref[big], rtol=2e-5)
    check_info(tmp_path, oracle, P, t2, st)


def check_info(workdir, oracle, P, t, st):
    """info.dat in the reference's layout (tube.f90:519-525, 941-952, 1033-1034, 1193-1207)."""
    lines = open(workdir / "output/info.dat").read().splitlines()
    sep = "=============================="
    keys = [ln.strip() if ln.strip() == sep else (ln.split("=")[0].strip() if "=" in ln else ln.strip())
            for ln in lines]
    order = ["Initialization time is", "==============================", "[i] [Section count] [Spline time] [Construction time]"]
    assert lines[0].strip().startswith(order[0]) and keys[1] == order[1] and keys[2] == order[2]
    ntube = P.line_e - P.line_s + 1
    rows = [ln.split() for ln in lines[3:3 + ntube]]
    assert [int(r[0]) for r in rows] == list(range(P.line_s, P.line_e + 1))
    assert [int(r[1]) for r in rows] == [int(n) for n in t.npoint[P.line_s - 1:P.line_e]]
    tail = keys[3 + ntube:]
    assert tail == ["==============================", "Total construction time", "velocity calculation begin",
                    "velocity calculation done", "Etot_temp", "up_temp", "Post processing time",
                    "==============================", "N", "Total length of curve", "Gamma_1", "twist eta", "Etot",
                    "Enstrophy"]
    val = {k: float(ln.split("=")[1]) for k, ln in zip(keys, lines) if "=" in ln and k != sep}
    geo = oracle.Geometry(P, t.npoint, t.cx, t.cy, t.cz)
    assert val["N"] == P.N
    assert val["Total length of curve"] == pytest.approx(float(geo.length.sum()), rel=1e-12)
    assert val["Etot_temp"] == pytest.approx(st[0], rel=1e-4)
    assert val["up_temp"] == pytest.approx(st[1], rel=1e-4)
    assert val["Gamma_1"] == pytest.approx(t.gamma[0] / st[1], rel=1e-4)
    assert val["Etot"] == pytest.approx(1.5, rel=1e-5)
    assert val["Enstrophy"] == pytest.approx(st[2], rel=1e-4)


def test_fortran_host_on_reference_wt2_inputs(tmp_path):
    """The reference's shipped input pair drives the host unchanged (grid reduced to 64^3 by
    editing only the N line, to keep the run short)."""
    src = os.path.join(os.path.dirname(__file__), "golden", "wt2")
    lines = open(os.path.join(src, "parameters_input.dat")).read().splitlines()
    lines[1] = "      64"
    lines[3] = "      64"
    (tmp_path / "parameters_input.dat").write_text("\n".join(lines) + "\n")
    (tmp_path / "centerline_input.dat").write_text(open(os.path.join(src, "centerline_input.dat")).read())
    out = run_host(tmp_path, 32)
    assert "up =" in out
    E = np.loadtxt(tmp_path / "output/stat/spectrum.dat")
    assert E.shape == (32, 2) and np.all(np.isfinite(E)) and E[:, 1].sum() > 0
    td = open(tmp_path / "output/test_data.dat").read().split()
    assert td.count("ntf_average=") == 73
