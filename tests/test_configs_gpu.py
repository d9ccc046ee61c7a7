"""The benchmark workloads at their stated sizes, on the GPU.

* 128^3 FP64 plain method: fields, hits, ntf, spectrum and structure functions against the oracle.
* 512^3 FP32 batched workload (5460 tubes): the certified FP32 turning angle against the reference
  sum on every section; FP32 fields, hits and spectrum against the FP64 path on the same seed; one
  tube per level against the oracle on a 64-plane slab.
* The spectral stage at production sizes (128, 256, 512, 768, 1024) with default and tuned tiles
  against torch.fft (rocFFT) in float64.
* 1024^3 at vortex densities 0.5, 1 and 2; the 64-instance uneven ensemble.
Errors are printed so that the run's log records them.
"""
import json
import os
import shutil

import numpy as np
import pytest

from fastweavturb_amd import fbb

pytestmark = pytest.mark.gpu

TILES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fastweavturb_amd", "data",
                     "tiles.json")


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def oparams(O, p):
    return O.Params(p.N, p.n_d, p.line_s, p.line_e, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre,
                    p.Rtube_ratio, p.dzeta_sup)


def ctx(p, t, precision):
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(p.N, precision)
    g.set_params(p)
    g.set_tubes(t)
    return g


# ---- 128^3 FP64 --------------------------------------------------------------------------------
def test_plain_128_fp64_matches_oracle(oracle):
    p, t = fbb.config_plain_128(1234)
    P = oparams(oracle, p)
    print("128^3 FP64: oracle run", flush=True)
    om_ref, u_ref, st = oracle.run(P, t.npoint, t.cx, t.cy, t.cz, t.sigma, t.gamma, t.n_T)
    geo = oracle.Geometry(P, t.npoint, t.cx, t.cy, t.cz)
    g = ctx(p, t, 64)
    g.generate()
    ntf, dtf = g.sections()
    np.testing.assert_array_equal(ntf, geo.ntf)
    np.testing.assert_array_equal(dtf, geo.dtf)
    assert g.deposit_stats()["hits"] == int(st[4])
    e_om, e_u = rel_l2(g.get_fields("omega"), om_ref), rel_l2(g.get_fields("u"), u_ref)
    E, E_ref = g.spectrum(1), oracle.spectrum(u_ref)
    e_E = float(np.max(np.abs(E / E_ref - 1)))
    seps = np.arange(1, 65, dtype=np.int32)
    S, S_ref = g.structure_functions(1, 2, 6, seps), oracle.sf_x(u_ref[0], 2, 6, seps)
    e_S = float(np.max(np.abs(S - S_ref) / np.abs(S_ref).max(axis=1, keepdims=True)))
    print(f"128^3 FP64 vs oracle: omega {e_om:.2e} u {e_u:.2e} shells {e_E:.2e} SF {e_S:.2e} "
          f"hits {int(st[4])}", flush=True)
    assert e_om <= 1e-12 and e_u <= 1e-12
    assert e_E <= 1e-10 and e_S <= 1e-10
    g.close()


# ---- 512^3 FP32 workload ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2():
    return fbb.config_batched_512(1234)


def test_512_fast_turning_angle_equals_reference_sum(c2):
    """k_theta_fast decides ntf alone for all but a few sections; on every one of the 55M sections
    ntf and dtf equal those of the reference's FP64 sum (WT_THETA_EXACT=1)."""
    p, t = c2
    g = ctx(p, t, 32)
    st = g.build_segments()
    ntf, dtf = g.sections()
    g.close()
    os.environ["WT_THETA_EXACT"] = "1"
    try:
        h = ctx(p, t, 32)
    finally:
        del os.environ["WT_THETA_EXACT"]
    st2 = h.build_segments()
    ntf2, dtf2 = h.sections()
    h.close()
    assert st2["theta_exact_sections"] == st2["nsec_all"] == len(ntf)
    print(f"512^3 turning angle: {len(ntf)} sections, {st['theta_exact_sections']} re-evaluated by the "
          f"fast path, ntf equal: {np.array_equal(ntf, ntf2)}", flush=True)
    np.testing.assert_array_equal(ntf, ntf2)
    np.testing.assert_array_equal(dtf, dtf2)


def test_512_fp32_against_fp64_same_seed(c2):
    """The headline workload: FP32 fields within 1e-5 relative L2 of the FP64 path, the same
    depositing pairs, spectrum shells within 1e-4."""
    p, t = c2
    out = {}
    for prec in (32, 64):
        print(f"512^3 FP{prec} generation", flush=True)
        g = ctx(p, t, prec)
        g.generate()
        out[prec] = dict(om=g.get_fields("omega"), u=g.get_fields("u"), hits=g.deposit_stats()["hits"],
                         E=g.spectrum(1), Ew=g.spectrum(0), scal=g.scalars())
        g.close()
    a, b = out[32], out[64]
    e_om, e_u = rel_l2(a["om"], b["om"]), rel_l2(a["u"], b["u"])
    e_E = float(np.max(np.abs(a["E"] / b["E"] - 1)))
    e_Ew = float(np.max(np.abs(a["Ew"] / b["Ew"] - 1)))
    print(f"512^3 FP32 vs FP64: omega {e_om:.2e} u {e_u:.2e} shells(u) {e_E:.2e} shells(omega) {e_Ew:.2e} "
          f"hits {a['hits']} / {b['hits']} up {a['scal'][1]:.9e} / {b['scal'][1]:.9e}", flush=True)
    assert a["hits"] == b["hits"]
    assert e_om <= 1e-5 and e_u <= 1e-5
    assert e_E <= 1e-4 and e_Ew <= 1e-4


def test_512_one_tube_per_level_against_oracle_slab(oracle):
    """One tube of each of the six levels at 512^3 (the level-0 tube's boxes are 277 cells wide),
    deposited by the GPU in FP64 and FP32 and by the oracle on the 64 planes z = 225..288."""
    levels = [fbb.Level(0.8 * 2.0 ** -l, 1, (256 * 2 ** l, 256 * 2 ** (l + 1))) for l in range(6)]
    p, t = fbb.woven_config(512, levels, 1234)
    P = oparams(oracle, p)
    geo = oracle.Geometry(P, t.npoint, t.cx, t.cy, t.cz)
    k0, k1 = 225, 288
    ref, ref_hits = geo.deposit_planes(t.sigma, t.gamma, t.n_T, k0, k1)
    assert ref_hits > 0
    for prec, tol in ((64, 1e-12), (32, 1e-5)):
        g = ctx(p, t, prec)
        g.build_segments()
        ntf, _ = g.sections()
        np.testing.assert_array_equal(ntf, geo.ntf)
        g.bin()
        g.deposit()
        om = g.get_fields("omega")[:, k0 - 1:k1]
        e = rel_l2(om, ref)
        print(f"512^3 six tubes FP{prec} vs oracle on planes {k0}..{k1}: omega {e:.2e}", flush=True)
        assert e <= tol
        g.close()


# ---- spectral stage at production sizes -----------------------------------------------------------
def biot_savart_reference(om):
    """u = Re(IFFT(i q x w_hat / k^2)) with the Nyquist components of q zeroed (the reference's
    dreal of a full c2c inverse), normalised to u' = 1 (tube.f90:985-1043), by torch.fft in float64
    on the device of `om` (a (3, N, N, N) tensor)."""
    import torch
    dev = om.device
    N = om.shape[-1]
    k = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=dev)
    k[N // 2] = -N // 2
    kx = torch.arange(N // 2 + 1, dtype=torch.float64, device=dev)
    KZ, KY, KX = k[:, None, None], k[None, :, None], kx[None, None, :]
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    ny = lambda K: torch.where(K.abs() == N // 2, torch.zeros_like(K), K)
    qx, qy, qz = ny(KX), ny(KY), ny(KZ)
    w = [torch.fft.rfftn(om[c].to(torch.float64)) for c in range(3)]
    u = torch.empty((3, N, N, N), dtype=torch.float64, device=dev)
    for c in range(3):
        a, b, qa, qb = ((w[2], w[1], qy, qz), (w[0], w[2], qz, qx), (w[1], w[0], qx, qy))[c]
        uh = 1j * (qa * a - qb * b) / k2
        uh[0, 0, 0] = 0
        u[c] = torch.fft.irfftn(uh, s=(N, N, N))
        del uh
    del w
    up = torch.sqrt(2.0 * (0.5 * (u ** 2).sum(dim=0).mean()) / 3.0)
    return u / up


def _reference_child(N, seed, out_dir, fp64):
    """Runs in a child process (torch brings its own HIP runtime, which cannot share a process
    with the library's once either has opened the GPU): draws the random vorticity on the GPU,
    computes the reference velocity and writes both as .npy files."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    om = torch.randn((3, N, N, N), generator=gen, device="cuda", dtype=torch.float32)
    np.save(os.path.join(out_dir, "omega.npy"), om.cpu().numpy())
    u = biot_savart_reference(om)
    np.save(os.path.join(out_dir, "u_ref.npy"), u.cpu().numpy().astype(np.float64 if fp64 else np.float32))


def _scratch_dir(tmp_path, nbytes):
    shm = "/dev/shm"
    try:
        if shutil.disk_usage(shm).free > 2 * nbytes:
            d = os.path.join(shm, f"wt_fft_{os.getpid()}")
            os.makedirs(d, exist_ok=True)
            return d
    except OSError:
        pass
    return str(tmp_path)


@pytest.mark.parametrize("N,precision,tuned", [
    (128, 64, False), (128, 32, True), (256, 64, False), (256, 32, True),
    (512, 32, False), (512, 32, True), (768, 32, False), (1024, 32, True)])
def test_spectral_stage_production_sizes(N, precision, tuned, tmp_path):
    """Random vorticity (i.i.d. normal values: content at every wavenumber, Nyquist included)
    through wt_biot_savart against torch.fft (rocFFT) in float64; 768 and 1024 take four passes.
    The reference's u is stored in float32 for FP32 runs (6e-8 relative, far below the bar)."""
    import subprocess
    import sys
    from fastweavturb_amd.api import WovenGPU
    print(f"spectral stage N={N} FP{precision} tuned={tuned}", flush=True)
    d = _scratch_dir(tmp_path, 3 * N ** 3 * (4 + (8 if precision == 64 else 4)))
    try:
        code = ("import sys; sys.path.insert(0, %r); sys.path.insert(0, %r); import test_configs_gpu as T; "
                "T._reference_child(%d, %d, %r, %s)" % (os.path.dirname(os.path.abspath(__file__)),
                                                       os.path.dirname(TILES).rsplit(os.sep, 2)[0],
                                                       N, 40 + N, d, precision == 64))
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-3000:]
        om = np.load(os.path.join(d, "omega.npy"), mmap_mode="r")
        ref = np.load(os.path.join(d, "u_ref.npy"), mmap_mode="r")
        g = WovenGPU(N, precision)
        if tuned:
            cache = str(tmp_path / "tiles.json")
            shutil.copy(TILES, cache)
            rep = g.autotune(cache)
            assert f"gfx950:{precision}:{N}" in json.load(open(cache)), rep
        for c in range(3):
            g.set_field(c, np.asarray(om[c], dtype=np.float64 if precision == 64 else np.float32))
        g.biot_savart()
        num = den = 0.0
        for c in range(3):
            u = g.get_field(3 + c, fp64=precision == 64).astype(np.float64)
            rc = np.asarray(ref[c], dtype=np.float64)
            num += float(np.sum((u - rc) ** 2))
            den += float(np.sum(rc ** 2))
            del u, rc
        g.close()
        del om, ref
    finally:
        if d.startswith("/dev/shm/"):
            shutil.rmtree(d, ignore_errors=True)
    e = (num / den) ** 0.5
    print(f"spectral stage N={N} FP{precision} tuned={tuned}: u {e:.2e}", flush=True)
    assert e <= (1e-12 if precision == 64 else 1e-5)


# ---- 1024^3 densities ------------------------------------------------------------------------------
@pytest.mark.parametrize("density", [0.5, 1.0, 2.0])
def test_1024_density_sweep(density):
    from fastweavturb_amd.api import WovenGPU
    print(f"1024^3 density {density}: building tubes", flush=True)
    p, t = fbb.config_c3(density=density)
    g = WovenGPU(1024, 32)
    g.set_params(p)
    g.set_tubes(t)
    etot, up, ens = g.generate()
    dep = g.deposit_stats()
    E = g.spectrum(1)
    seps = np.arange(1, 513, dtype=np.int32)
    S = g.structure_functions(1, 2, 6, seps)
    print(f"1024^3 density {density}: {t.nline} tubes, hits {dep['hits']}, up {up:.6e}, "
          f"shell sum {E.sum():.6f}, S2(1) {S[0, 0]:.4e}, F4(1) {S[2, 0] / S[0, 0] ** 2:.3f}", flush=True)
    assert dep["overflow"] == 0 and dep["hits"] > 0
    assert 1.4 < E.sum() <= 1.5 + 1e-6
    assert np.all(np.isfinite(S)) and np.all(S[0] > 0) and np.all(S[2] > 0) and np.all(S[4] > 0)
    g.close()


# ---- the 64-instance ensemble ----------------------------------------------------------------------
def test_ensemble_64_uneven_instances(oracle):
    """64 instances drawn over {64, 96, 128, 192, 256}^3 in one context: instances of every grid
    size equal their single-instance runs bitwise (FP32), and a 64^3 instance matches the oracle."""
    from fastweavturb_amd import surrogate as Sg
    from fastweavturb_amd.api import WovenGPU
    jobs = Sg.ensemble_jobs(64, seed=5000)
    cfg = [Sg.instance_config(d, N, s) for d, N, s in jobs]
    sizes = [N for _, N, _ in jobs]
    assert set(sizes) == set(Sg.GRIDS)
    b = WovenGPU(precision=32, sizes=sizes)
    for i, (p, t) in enumerate(cfg):
        b.select(i)
        b.set_params(p)
        b.set_tubes(t)
    b.generate()
    pick = [sizes.index(N) for N in Sg.GRIDS]
    for i in pick:
        p, t = cfg[i]
        s = ctx(p, t, 32)
        s.generate()
        b.select(i)
        for name in ("omega", "u"):
            np.testing.assert_array_equal(b.get_fields(name, fp64=False), s.get_fields(name, fp64=False))
        s.close()
        print(f"ensemble instance {i} (N={sizes[i]}) equals its single run", flush=True)
    i = pick[0]
    p, t = cfg[i]
    om, u, st = oracle.run(oparams(oracle, p), t.npoint, t.cx, t.cy, t.cz, t.sigma, t.gamma, t.n_T)
    b.select(i)
    e_om, e_u = rel_l2(b.get_fields("omega"), om), rel_l2(b.get_fields("u"), u)
    print(f"ensemble instance {i} (N=64) vs oracle: omega {e_om:.2e} u {e_u:.2e}", flush=True)
    assert e_om <= 1e-5 and e_u <= 1e-5
    b.close()
