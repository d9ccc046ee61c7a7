#!/usr/bin/env python3
"""Woven-field throughput on one MI355X: Mpts/s of a 512^3 FP32 woven field.

One step = one whole woven-field generation from the tube table and centerline points already
resident in device memory to normalised vorticity and velocity resident in device memory:
geometry (cze, splines, turning angle, sub-points), binning, deposition, forward FFT,
Biot-Savart, inverse FFT, normalisation (tube.f90:392-1043). Workload: 6 levels, r = 0.8*2^-l,
4*4^l tubes per level (5460 tubes), 256*2^l..256*2^(l+1) points per tube (55M points), seeded
fBb centerlines with H = 0.75, synthetic. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts/s woven field at 512^3 FP32; time-to-target-spectrum (±5%) vs sweep"
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
SLOT_NAMES = ["theta", "subpoints", "bin_count", "deposit", "bin_fill", "fft_x", "fft_y",
              "fft_z_biot_savart", "fft_y_inv", "fft_x_inv", "spectral_reduce"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str):
    from fastweavturb_amd import fbb
    if name == "c2":
        return "batched_512_fp32", 32, fbb.config_batched_512(seed=1234)
    if name == "c1":
        return "plain_128_fp64", 64, fbb.config_plain_128(seed=1234)
    if name == "small":
        return "small_64_fp32", 32, fbb.woven_config(64, [fbb.Level(0.5, 4, (256, 512)), fbb.Level(0.25, 16, (512, 1024))], 1234)
    raise SystemExit(f"unknown config {name}")


def algorithmic_cost(slot: str, N: int, seg: dict, dep: dict, p):
    """(amount, unit, bound, peak) of one launch of a kernel: bytes for memory-bound kernels,
    FLOPs for arithmetic-bound ones: per-unit costs times the units one launch processes."""
    H = N // 2 + 1
    half = 3 * N * N * H * 8          # three FP32 half spectra, bytes
    real = 3 * N ** 3 * 4             # three FP32 fields, bytes
    if slot in ("fft_y", "fft_z_biot_savart", "fft_y_inv"):
        return 2 * half, "GB/s", "hbm", PEAK_HBM_GBS
    if slot == "fft_x":
        return real + half, "GB/s", "hbm", PEAK_HBM_GBS
    if slot == "fft_x_inv":
        return half + real + 2 * real, "GB/s", "hbm", PEAK_HBM_GBS
    if slot == "theta":
        # certified FP32 re-evaluation of the reference's ntf_pre+1 samples per section, in pairs
        # t = 1/2 +- s: per pair s^2, the even and odd parts of the 4 shifted polynomials (25 FMA),
        # 4 products s*O and 8 sums (63 FLOP), the pair's sum, its FP64 accumulation and the smaller
        # denominator (3 FLOP); per sample |w|^2 (1 mul + 2 FMA), sqrt, reciprocal, 1 mul (8 FLOP):
        # 41 FLOP per sample
        return seg["nsec_all"] * (p.ntf_pre + 1) * 41.0, "TFLOP/s", "valu", PEAK_FP32_TFLOPS
    if slot == "deposit":
        # FP64 on the vector ALU (its peak equals the published FP64 peak): 16 FLOP per grid column
        # walked (two plane positions and the interval), 33 per candidate cell (the reference's
        # three tests), 150 per hit (zeta, spline point and tangent, offset, Gaussian core,
        # fixed-point conversion)
        return dep["columns"] * 16 + dep["candidates"] * 33 + dep["hits"] * 150, "TFLOP/s", "valu", PEAK_FP64_TFLOPS
    return None


def host_cpu() -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def tube_subsegments(p, t, ntf) -> np.ndarray:
    """Sub-segments of every tube: the sum of ntf over its real sections (its npoint sections;
    the periodic ghost section deposits nothing)."""
    sec_base = np.concatenate([[0], np.cumsum(t.npoint.astype(np.int64) + 1)])
    tot = np.add.reduceat(ntf.astype(np.int64), sec_base[:-1])
    return tot - ntf[sec_base[1:] - 1]


def cpu_baseline(p, t, nseg_tube, budget_s: float):
    """The reference's CPU path for the same workload, on every host core: geometry and deposition
    by the C restatement of tube.f90 (threaded: geometry one tube per thread, deposition split into
    z-slabs of planes the way the reference's MPI ranks split it, tube.f90:803-804), the 3-D FFTs of
    the Biot-Savart stage by scipy.fft (pocketfft, FP64 c2c like the reference's FFTW plans) on
    every core. Geometry and deposition run on a sample of each level's tubes (and of the grid
    planes when a level's boxes are large) sized to the time budget and are scaled to the level by
    its sub-segment (deposition) and point (geometry) counts; the FFTs run at the full size. Core
    seconds are measured (process CPU time) and converted to wall time assuming the workload's
    thousands of tubes and N z-planes keep every core busy."""
    import scipy.fft as sfft
    from oracle import oracle_py as O
    from fastweavturb_amd.woven_io import Tubes
    cores = O.threads()
    N = p.N
    dx = 2 * np.pi / N
    rng = np.random.default_rng(0)
    P = O.Params(N, N, 1, 1, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre, p.Rtube_ratio, p.dzeta_sup)
    sig = np.round(t.sigma, 12)
    levels = sorted(set(sig.tolist()), reverse=True)
    pt_base = np.concatenate([[0], np.cumsum(t.npoint.astype(np.int64))])
    # predicted cost per tube: box cells walked (6 ns per cell and core) and turning-angle samples
    # (25 ns per sample and core)
    box = (2 * p.Rtube_ratio * t.sigma / dx + 3) ** 3
    dep_s = nseg_tube * box * 6e-9 / cores
    cost_s = dep_s + t.npoint * (p.ntf_pre + 1) * 25e-9 / cores
    dep_budget = 0.7 * budget_s
    full = cost_s.sum() <= dep_budget
    geo_core = dep_core = 0.0
    parts = []
    for lv in levels:
        idx = np.flatnonzero(sig == lv)
        order = idx if full else rng.permutation(idx)
        lvl_cost = float(cost_s[idx].sum())
        share = dep_budget / len(levels)
        if full:
            take, frac = order, 1.0
        else:
            take, acc = [], 0.0
            for i in order:
                take.append(int(i))
                acc += cost_s[i]
                if acc >= share:
                    break
            take = np.array(take)
            g_s, d_s = float((cost_s - dep_s)[take].sum()), float(dep_s[take].sum())
            frac = min(1.0, max((share - g_s) / max(d_s, 1e-12), 0.02))
        # grid planes: every 64/nplane-th block of 8 planes when only a fraction is needed
        nblk = N // 8
        step = max(1, int(np.ceil(1.0 / max(frac, 1e-6) - 1e-9)))
        blocks = list(range(0, nblk, step))
        planes = N if step == 1 else 8 * len(blocks)
        sel = np.concatenate([np.arange(pt_base[i], pt_base[i + 1]) for i in take])
        sub = Tubes(t.npoint[take].copy(), t.cx[sel].copy(), t.cy[sel].copy(), t.cz[sel].copy(), t.sigma[take].copy(),
                    t.gamma[take].copy(), t.n_T[take].copy(), t.if_filter[take].copy())
        P.line_e = len(take)
        c0 = time.process_time()
        geo = O.Geometry(P, sub.npoint, sub.cx, sub.cy, sub.cz)
        cg = time.process_time() - c0
        c0 = time.process_time()
        if step == 1:
            geo.deposit_planes(sub.sigma, sub.gamma, sub.n_T, 1, N)
        else:
            for bk in blocks:
                geo.deposit_planes(sub.sigma, sub.gamma, sub.n_T, 8 * bk + 1, 8 * bk + 8)
        cd = time.process_time() - c0
        geo.close()
        seg_s = int(nseg_tube[take].sum())
        geo_scale = float(t.npoint[idx].sum()) / float(sub.npoint.sum())
        dep_scale = float(nseg_tube[idx].sum()) / seg_s * (N / planes)
        geo_core += cg * geo_scale
        dep_core += cd * dep_scale
        parts.append(f"r={lv * np.sqrt(2):.4g}: {len(take)}/{len(idx)} tubes, {planes}/{N} planes")
        log(f"[cpu] level r={lv * np.sqrt(2):.4g}: {len(take)}/{len(idx)} tubes, {planes}/{N} planes, "
            f"geometry {cg:.1f} core-s (x{geo_scale:.3g}), deposition {cd:.1f} core-s (x{dep_scale:.3g}), "
            f"predicted level {lvl_cost:.1f} s")
    # spectral stage at full size: 3 forward and 3 inverse c2c transforms of N^3 FP64 fields
    x = np.empty((N, N, N), dtype=np.complex128)
    x.real = rng.standard_normal((N, N, N))
    x.imag = 0.0
    w0 = time.perf_counter()
    for _ in range(3):
        y = sfft.fftn(x, workers=cores)
        x = sfft.ifftn(y, workers=cores)
    t_fft = time.perf_counter() - w0
    del x, y
    t_geo, t_dep = geo_core / cores, dep_core / cores
    total = t_geo + t_dep + t_fft
    sample = ("port on %d cores: tube.f90 geometry and deposition restated in C (threaded per tube / per "
              "z-slab), %s; Biot-Savart FFTs by scipy.fft at the full %d^3 (6 c2c, measured); core seconds "
              "scaled by sub-segment and point counts, divided by the cores. %s" % (
                  cores, "every tube and plane measured" if full else "a seeded sample of each level",
                  N, "; ".join(parts)))
    return {"value": N ** 3 / total / 1e6, "unit": "Mpts/s", "cores": cores, "kind": "port",
            "cpu": host_cpu(), "sample": sample, "seconds_estimated": total,
            "seconds_parts": {"geometry": t_geo, "deposition": t_dep, "fft": t_fft},
            "whole_workload_measured": bool(full),
            "note": "the reference Fortran (tube.f90) needs FFTW3 and MPI, absent here; its C restatement is timed"}


def time_to_target(n_targets: int, n_sweeps: int, n_synth: int = 4, sweep_cap: int = 100):
    """Time-to-target-spectrum on 512^3 FP32 woven fields (every candidate costs one full 512^3
    generation from the GPU-drawn centerline bank plus its statistics).
    * Reachable targets: hidden designs of the plain method's family (seeded); the surrogate
      (trained here on the GPU from the shipped generator-produced pairs) predicts a design and
      refines it until all 32 log shells are within +-5%; the plain sweep walks its (gamma, density,
      hurst) grid until one candidate is. Both are timed.
    * Synthetic targets: seeded E(k) = C k^-5/3 exp(-beta k eta) spectra and the reference's model
      spectrum (flatness inputs: the training pairs' median); the surrogate first, the sweep as the
      fallback (capped at sweep_cap runs); the fraction each finds and the best shell error."""
    from fastweavturb_amd import surrogate as S
    bank = S.CenterlineBank(S.PAIRS_RECIPE["bank_seed"])
    X, Y, M = S.load_pairs()
    sur = S.Surrogate(seed=1)
    t0 = time.perf_counter()
    hist = sur.fit(X, Y, M, epochs=300)
    t_train = time.perf_counter() - t0
    log(f"[ttt] surrogate trained on {len(X)} pairs in {t_train:.1f}s (loss {hist[0]:.3f} -> {hist[-1]:.4f})")
    ver = S.Verifier(bank, 512)
    rows, synth = [], []
    try:
        for k, d in enumerate(S.target_designs(n_targets)):
            tgt = ver.run(d)
            r = S.surrogate_design(sur, tgt, ver)
            row = {"target": S.sweep_family(d), "surrogate_s": r["seconds"], "surrogate_runs": r["runs"],
                   "surrogate_found": r["found"], "surrogate_error": r["errors"][-1]}
            if k < n_sweeps:
                sw = S.sweep_design(tgt, ver)
                row.update({"sweep_s": sw["seconds"], "sweep_runs": sw["runs"], "sweep_found": sw["found"],
                            "sweep_grid": sw["grid_size"]})
            log(f"[ttt] target {k}: {json.dumps(row)}")
            rows.append(row)
        flat = np.median(X[:, S.NSHELL:], axis=0)
        for name, tgt in S.synthetic_targets(n_synth, flat):
            r = S.surrogate_design(sur, tgt, ver)
            row = {"target": name, "surrogate_found": r["found"], "surrogate_runs": r["runs"],
                   "surrogate_s": r["seconds"], "surrogate_best_error": float(min(r["errors"]))}
            if not r["found"] and sweep_cap > 0:
                sw = S.sweep_design(tgt, ver, max_runs=sweep_cap)
                row.update({"sweep_found": sw["found"], "sweep_runs": sw["runs"], "sweep_s": sw["seconds"],
                            "sweep_best_error": float(sw["best"][0]), "sweep_best": sw["best"][1]})
            log(f"[ttt] synthetic: {json.dumps(row)}")
            synth.append(row)
    finally:
        ver.close()
    s_ok = [r for r in rows if r["surrogate_found"]]
    both = [r for r in rows if "sweep_s" in r and r["surrogate_found"] and r["sweep_found"]]
    out = {"N": 512, "tolerance": 0.05, "shells": S.NSHELL, "targets": len(rows),
           "surrogate_found": len(s_ok), "surrogate_mean_s": float(np.mean([r["surrogate_s"] for r in s_ok])) if s_ok else None,
           "surrogate_mean_runs": float(np.mean([r["surrogate_runs"] for r in s_ok])) if s_ok else None,
           "sweep_targets": sum("sweep_s" in r for r in rows),
           "sweep_mean_s": float(np.mean([r["sweep_s"] for r in both])) if both else None,
           "sweep_mean_runs": float(np.mean([r["sweep_runs"] for r in both])) if both else None,
           "speedup": (float(np.mean([r["sweep_s"] for r in both]) / np.mean([r["surrogate_s"] for r in both]))
                       if both else None),
           # the network's training on the GPU, shared by the targets of this run
           "speedup_incl_training": (float(np.mean([r["sweep_s"] for r in both]) /
                                           (np.mean([r["surrogate_s"] for r in both]) + t_train / (len(rows) + len(synth))))
                                     if both else None),
           "training": {"pairs": int(len(X)), "train_s": t_train, "epochs": 300},
           "per_target": rows,
           "synthetic": {"targets": len(synth), "surrogate_found": sum(r["surrogate_found"] for r in synth),
                         "sweep_fallback_found": sum(bool(r.get("sweep_found")) for r in synth),
                         "sweep_cap": sweep_cap, "per_target": synth}}
    return out


def rocfft_compare(N: int, reps: int = 5) -> dict:
    """Times rocFFT (through torch.fft, which ROCm builds on hipFFT/rocFFT) on the transforms of
    the Biot-Savart stage: R2C of the three vorticity components and C2R of the three velocity
    components at N^3 FP32, as one batched call each, on torch's stream with torch events."""
    import torch
    x = torch.randn(3, N, N, N, device="cuda:0", dtype=torch.float32)
    X = torch.fft.rfftn(x, dim=(1, 2, 3))
    y = torch.fft.irfftn(X, s=(N, N, N), dim=(1, 2, 3))
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd = inv = 0.0
    for _ in range(reps):
        e[0].record()
        X = torch.fft.rfftn(x, dim=(1, 2, 3))
        e[1].record()
        y = torch.fft.irfftn(X, s=(N, N, N), dim=(1, 2, 3))
        e[2].record()
        torch.cuda.synchronize()
        fwd += e[0].elapsed_time(e[1])
        inv += e[1].elapsed_time(e[2])
    del x, X, y
    torch.cuda.empty_cache()
    return {"rocfft_r2c_3x_ms": fwd / reps, "rocfft_c2r_3x_ms": inv / reps,
            "rocfft_transforms_ms": (fwd + inv) / reps, "via": "torch.fft (hipFFT/rocFFT)", "N": N}


def _force_tiles(g, tiles):
    if tiles:
        g.set_deposit_tiles(*[int(v) for v in tiles.split(",")])
        log(f"[bench] deposition tiles set to {tiles}")


def ensemble_bench(steps: int, warmup: int, n: int = 64, graph: bool = True, tile_cache: str | None = None,
                   tiles: str | None = None) -> dict:
    """The ensemble workload: n uneven instances (grid sizes from {64, 96, 128, 192, 256}, 3-6
    levels, Hurst index, density and circulation exponent drawn per instance, tube seed 5000 + i)
    generated together in one context, the spectral stage replayed from one hipGraph, then the
    statistics of every instance in one call (32 log shells of E(k) and the flatness F4, F6 of
    x-increments at 8 separations)."""
    from fastweavturb_amd import surrogate as S
    from fastweavturb_amd.api import WovenGPU
    t0 = time.time()
    jobs = S.ensemble_jobs(n, seed=5000)
    cfgs = [S.instance_config(d, N, seed) for d, N, seed in jobs]
    sizes = [p.N for p, _ in cfgs]
    log(f"[c4] {n} instances built in {time.time() - t0:.1f}s; sizes {dict(zip(*np.unique(sizes, return_counts=True)))}")
    g = WovenGPU(precision=32, sizes=sizes)
    g.set_graph(graph)
    for i, (p, t) in enumerate(cfgs):
        g.select(i)
        g.set_params(p)
        g.set_tubes(t)
    log("[c4] autotune: " + g.autotune(tile_cache).strip().replace("\n", "; "))
    _force_tiles(g, tiles)
    for _ in range(max(1, warmup)):
        g.generate()
    t0 = time.perf_counter()
    stage = np.zeros(4)
    for _ in range(steps):
        g.generate()
        stage += g.stage_times()[:4]
    dt = (time.perf_counter() - t0) / steps
    # the same generations replayed from one hipGraph of the whole ensemble (wt_set_graph mode 2)
    g.set_graph(2)
    g.generate()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.generate()
    dt_graph = (time.perf_counter() - t0) / steps
    g.set_graph(graph)
    log(f"[c4] eager {dt * 1e3:.1f} ms, whole-ensemble graph {dt_graph * 1e3:.1f} ms per generation")
    seps = np.stack([S.flat_separations(N) for N in sizes])
    t1 = time.perf_counter()
    E, Sf = g.batch_statistics(2, 6, seps)
    t_stats = time.perf_counter() - t1
    feats = np.stack([S.features(E[i], Sf[i]) for i in range(n)])
    seg = g.segment_stats()
    dep = g.deposit_stats()
    g.close()
    pts = float(sum(N ** 3 for N in sizes))
    return {"metric": "ensemble woven fields per second (64 uneven instances)", "value": n / dt,
            "unit": "instances/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": dt * 1e3,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "ensemble_64_uneven", "instances": n, "grid_points": pts,
                       "sizes": {str(k): int(v) for k, v in zip(*np.unique(sizes, return_counts=True))},
                       "tubes": int(sum(t.nline for _, t in cfgs)), "points": int(sum(t.npoint_all for _, t in cfgs)),
                       "graph": graph},
            "mpts_per_s": pts / dt / 1e6, "stages_ms": dict(zip(["segments", "bin", "deposit", "biot_savart"],
                                                                 (stage / steps).tolist())),
            "graph_ms_per_step": dt_graph * 1e3, "graph_instances_per_s": n / dt_graph,
            "statistics_ms": t_stats * 1e3, "sub_segments": int(seg["nseg_all"]), "hits": int(dep["hits"]),
            "features_first": feats[0].round(4).tolist()}


def c3_bench(densities=(0.5, 1.0, 2.0), tile_cache: str | None = None, tiles: str | None = None) -> dict:
    """The 1024^3 workload: 8 levels (r = 2^-l), density swept over `densities`, fields resident
    on the GPU, structure functions of orders 2-6 of u_x at every separation 1..512."""
    from fastweavturb_amd import fbb
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(1024, 32)
    seps = np.arange(1, 513, dtype=np.int32)
    runs = []
    for d in densities:
        t0 = time.time()
        p, t = fbb.config_c3(d)
        log(f"[c3] density {d}: {t.nline} tubes, {t.npoint_all} points (built in {time.time() - t0:.1f}s)")
        g.set_params(p)
        g.set_tubes(t)
        if d == densities[0]:              # FFT and deposition tiles (same levels at every density)
            log("[c3] autotune: " + g.autotune(tile_cache).strip().replace("\n", "; "))
            _force_tiles(g, tiles)
        g.generate()                       # warm-up: grows the pair list and tube arenas once
        t0 = time.perf_counter()
        etot, up, ens = g.generate()
        t_gen = time.perf_counter() - t0
        E = g.spectrum(1)
        t0 = time.perf_counter()
        S = g.structure_functions(1, 2, 6, seps)
        t_sf = time.perf_counter() - t0
        free, total = _device_memory()
        dep = g.deposit_stats()
        row = {"density": d, "tubes": t.nline, "points": int(t.npoint_all), "generate_s": t_gen,
               "mpts_per_s": 1024 ** 3 / t_gen / 1e6, "stages_ms": g.stage_times()[:4].tolist(),
               "sf_s": t_sf, "hits": int(dep["hits"]), "pairs": int(dep["pairs"]), "up": up,
               "energy_in_shells": float(E.sum()), "F4_at_1": float(S[2, 0] / S[0, 0] ** 2),
               "F4_at_512": float(S[2, -1] / S[0, -1] ** 2), "device_used_gb": (total - free) / 1e9}
        log(f"[c3] {json.dumps(row)}")
        runs.append(row)
    g.close()
    return {"metric": "Mpts/s woven field at 1024^3 FP32 (8 levels)", "workload": "c3_1024_fp32", "runs": runs}


def _device_memory():
    """(free, total) bytes of device 0, read through the HIP runtime the library uses."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0:
            return free.value, total.value
    except OSError:
        pass
    return 0, 0


def copy_probe(nbytes: int = 1 << 30, reps: int = 3):
    """A plain elementwise copy of known size: each launch reads and writes exactly nbytes."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe": {"bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                                     "launches": reps, "kernel": "elementwise mul (y = x * 1)"}}), flush=True)


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
        e = tab["workloads"][workload][kernel]
        return float(e["fetch_kib"] * 1024 * tab["calibration"]["fetch_factor"] +
                     e["write_kib"] * 1024 * tab["calibration"]["write_factor"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c2", "c1", "small", "c4", "c3"])
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of host time the CPU baseline's sample may take")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ttt-targets", type=int, default=4,
                    help="targets of the time-to-target-spectrum measurement (0 skips it)")
    ap.add_argument("--ttt-sweeps", type=int, default=-1,
                    help="targets also searched by the plain sweep (-1: every target)")
    ap.add_argument("--ttt-synthetic", type=int, default=4,
                    help="seeded k^-5/3 synthetic target spectra (the reference's model spectrum is added)")
    ap.add_argument("--ttt-sweep-cap", type=int, default=100,
                    help="runs of the sweep fallback for a synthetic target the surrogate misses")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default FFT tile shapes")
    ap.add_argument("--no-graph", action="store_true", help="skip timing the whole-generation graph replay")
    ap.add_argument("--deposit-tiles", default=None,
                    help="bx,by,bz,colcap for every grid size after the autotune (default: the tuned ones)")
    ap.add_argument("--c3-densities", default="0.5,1,2", help="tube densities of the 1024^3 workload")
    ap.add_argument("--tile-cache", default=os.path.join(ROOT, "fastweavturb_amd", "data", "tiles.json"),
                    help="autotune cache (JSON): shapes recorded there are reused, new ones appended")
    ap.add_argument("--no-fft-compare", action="store_true",
                    help="skip timing rocFFT's transforms of the same Biot-Savart stage")
    ap.add_argument("--rocfft-only", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB -> 1 GiB device copy kernel (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the woven-field path runs on one GPU only (asked for --gpus {args.gpus}); "
              "the z-slab decomposition over several GPUs is not built.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.rocfft_only:
        print(json.dumps(rocfft_compare(args.rocfft_only)), flush=True)
        return
    if args.config == "c3":
        print(json.dumps(c3_bench(tuple(float(v) for v in args.c3_densities.split(",")), tile_cache=args.tile_cache,
                                  tiles=args.deposit_tiles)), flush=True)
        return
    if args.config == "c4":
        print(json.dumps(ensemble_bench(args.steps, args.warmup, tile_cache=args.tile_cache, tiles=args.deposit_tiles)),
              flush=True)
        return
    fftc = None
    if not args.no_fft_compare:
        # torch carries its own HIP runtime: it runs in a child process started before this
        # process touches the GPU
        import subprocess
        nfft = {"c2": 512, "c1": 128, "small": 64}[args.config]
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--rocfft-only", str(nfft)],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            log(f"[bench] rocFFT comparison failed: {r.stderr[-2000:]}")
        else:
            fftc = json.loads(r.stdout.strip().splitlines()[-1])

    from fastweavturb_amd.api import WovenGPU
    t0 = time.time()
    wname, precision, (p, t) = make_workload(args.config)
    log(f"[bench] workload {wname}: {t.nline} tubes, {t.npoint_all} points, N={p.N} "
        f"(generated in {time.time() - t0:.1f}s)")
    g = WovenGPU(p.N, precision)
    g.set_params(p)
    g.set_tubes(t)
    # the cold first generation: the tube table's upload, the first allocations of the geometry and
    # the work lists, default deposition tiles
    tc = time.perf_counter()
    g.generate()
    cold_ms = (time.perf_counter() - tc) * 1e3
    log(f"[bench] cold first generation {cold_ms:.1f} ms")
    if not args.no_autotune:
        log("[bench] autotune: " + g.autotune(args.tile_cache).strip().replace("\n", "; "))
    _force_tiles(g, args.deposit_tiles)
    for i in range(args.warmup):
        g.generate()
        log(f"[bench] warmup {i} done, stages ms {np.round(g.stage_times()[:4], 2).tolist()}")
    # every library call is stream-synchronous on return, so the clock brackets finished work
    t0 = time.perf_counter()
    stage_ms = np.zeros(4)
    step_ms = []
    for i in range(args.steps):
        ts = time.perf_counter()
        g.generate()
        step_ms.append((time.perf_counter() - ts) * 1e3)
        stage_ms += g.stage_times()[:4]
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = p.N ** 3 / (ms_per_step * 1e-3) / 1e6
    times = g.stage_times()
    kernels = {SLOT_NAMES[i]: float(times[4 + i]) for i in range(len(SLOT_NAMES)) if times[4 + i] > 0}
    seg = g.segment_stats()
    dep = g.deposit_stats()
    ntf, _ = g.sections()
    nseg_tube = tube_subsegments(p, t, ntf)
    del ntf
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step; stages {np.round(stage_ms / args.steps, 2).tolist()}")
    log(f"[bench] kernels ms {json.dumps({k: round(v, 3) for k, v in kernels.items()})}")
    log(f"[bench] geometry {json.dumps(seg)}; deposit {json.dumps(dep)}")
    dom = max(kernels, key=kernels.get)
    roof = None
    cost = algorithmic_cost(dom, p.N, seg, dep, p)
    if cost is not None:
        amount, unit, bound, peak = cost
        sec = kernels[dom] * 1e-3
        achieved = amount / sec / (1e9 if unit == "GB/s" else 1e12)
        resource = {"deposit": "FP64 vector ALU (peak = published FP64 peak)",
                    "theta": "FP32 vector ALU (packed FP32 disabled)"}.get(dom, "HBM")
        roof = {"kernel": dom, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                "frac": achieved / peak, "traffic": measured_traffic(wname, dom), "ms": kernels[dom],
                "resource": resource}
    # the path's own bound: every byte it must move once at 8 TB/s (omega written, the 6 FFT passes
    # of the half spectra, centerline points read, sub-segment records written and read once)
    b_w = p.N ** 3 * (12 + 144) + 24 * int(t.npoint_all) + 128 * int(seg["nseg_all"])
    t_roof = b_w / (PEAK_HBM_GBS * 1e9)
    headline = {"bytes": b_w, "t_roof_ms": t_roof * 1e3, "t_step_ms": ms_per_step,
                "frac": t_roof / (ms_per_step * 1e-3),
                "model": "B_W = N^3 (12 + 144) + 24 P + 128 S_sub bytes at 8 TB/s against the whole step"}
    if fftc is not None:
        spectral = [k for k in ("fft_x", "fft_y", "fft_z_biot_savart", "fft_y_inv", "fft_x_inv", "spectral_reduce")
                    if k in kernels]
        fftc["ours_biot_savart_stage_ms"] = float(stage_ms[3] / args.steps)
        fftc["ours_kernels_ms"] = float(sum(kernels[k] for k in spectral))
        log(f"[bench] fft compare {json.dumps(fftc)}")
    # the same generations replayed from one hipGraph (wt_set_graph mode 2): no host work or
    # launches between the stages; reported beside the eager figure, which carries the stage times
    graph_ms = None
    if not args.no_graph:
        g.set_graph(2)
        g.generate()
        t0 = time.perf_counter()
        for i in range(args.steps):
            g.generate()
        graph_ms = (time.perf_counter() - t0) / args.steps * 1e3
        g.set_graph(0)
        log(f"[bench] whole-generation graph: {graph_ms:.2f} ms/step")
    ttt = None
    g.close()
    if args.ttt_targets > 0 and args.config == "c2":
        ttt = time_to_target(args.ttt_targets, args.ttt_targets if args.ttt_sweeps < 0 else args.ttt_sweeps,
                             args.ttt_synthetic, args.ttt_sweep_cap)
    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(p, t, nseg_tube, args.cpu_budget)
        log(f"[bench] cpu baseline {cpu['value']:.4g} Mpts/s (sampled in {time.time() - t0:.1f}s)")
    out = {
        "metric": METRIC, "value": value, "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if precision == 32 else "f64", "data": "synthetic",
        "config": {"workload": wname, "N": p.N, "tubes": t.nline, "points": int(t.npoint_all),
                   "sub_segments": int(seg["nseg_all"]), "levels": 6 if args.config == "c2" else None},
        "stages_ms": dict(zip(["segments", "bin", "deposit", "biot_savart"], (stage_ms / args.steps).tolist())),
        "step_ms": step_ms, "median_ms": float(np.median(step_ms)), "cold_ms": cold_ms,
        "kernels_ms": kernels, "graph_ms_per_step": graph_ms, "roofline": roof, "headline_roofline": headline, "cpu_baseline": cpu,
        "time_to_target": ttt,
        "fft_vs_rocfft": fftc,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
