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
        # 4 products s*O and 8 sums (63 FLOP); per sample |w|^2 (1 mul + 2 FMA), sqrt, reciprocal,
        # 1 mul, the FP64 running sum and 3 bound updates (12 FLOP): 43.5 FLOP per sample
        return seg["nsec_all"] * (p.ntf_pre + 1) * 43.5, "TFLOP/s", "mfma", PEAK_FP32_TFLOPS
    if slot == "deposit":
        # FP64 on the vector ALU (its peak equals the published FP64 peak): 16 FLOP per grid column
        # walked (two plane positions and the interval), 33 per candidate cell (the reference's
        # three tests), 150 per hit (zeta, spline point and tangent, offset, Gaussian core,
        # fixed-point conversion)
        return dep["columns"] * 16 + dep["candidates"] * 33 + dep["hits"] * 150, "TFLOP/s", "mfma", PEAK_FP64_TFLOPS
    return None


def cpu_baseline(p, t, budget_s: float):
    """Time the FP64 oracle (single-threaded C restatement of the reference) on a bounded
    sample of the same workload and extrapolate to the whole field."""
    from oracle import oracle_py as O
    from fastweavturb_amd.woven_io import Tubes
    rng = np.random.default_rng(0)
    N = p.N
    # level boundaries by sigma
    sig = t.sigma
    levels = sorted(set(np.round(sig, 12)), reverse=True)
    pt_base = np.concatenate([[0], np.cumsum(t.npoint)])
    t_geo = 0.0
    t_dep = 0.0
    notes = []
    per_level_budget = budget_s * 0.8 / max(1, len(levels))
    for lv in levels:
        idx = np.flatnonzero(np.round(sig, 12) == lv)
        tube = int(rng.choice(idx))
        a, b = pt_base[tube], pt_base[tube + 1]
        sub = Tubes(t.npoint[tube:tube + 1].copy(), t.cx[a:b].copy(), t.cy[a:b].copy(), t.cz[a:b].copy(),
                    t.sigma[tube:tube + 1].copy(), t.gamma[tube:tube + 1].copy(), t.n_T[tube:tube + 1].copy(),
                    t.if_filter[tube:tube + 1].copy())
        P = O.Params(N, N, 1, 1, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre, p.Rtube_ratio, p.dzeta_sup)
        t0 = time.perf_counter()
        geo = O.Geometry(P, sub.npoint, sub.cx, sub.cy, sub.cz)
        tg = time.perf_counter() - t0
        level_points = int(t.npoint[idx].sum())
        t_geo += tg * level_points / int(sub.npoint[0])
        # deposit a growing chunk of sub-segments until the per-level budget is used
        nseg = geo.nseg
        chunk = 1
        spent, done, secs = 0.0, 0, 0.0
        while done < nseg and spent < per_level_budget:
            hi = min(nseg, done + chunk)
            dt, _ = O.deposit_range_timed(P, geo, sub.sigma, sub.gamma, sub.n_T, done, hi)
            secs += dt
            spent += dt
            done = hi
            chunk = min(chunk * 2, 1 << 20)
        seg_per_point = nseg / int(sub.npoint[0])
        t_dep += secs * (seg_per_point * level_points) / done
        notes.append(f"r={lv * np.sqrt(2):.4g}: tube {tube}, {done}/{nseg} sub-segments")
        geo.close()
    # spectral part: 6 3-D c2c transforms of N^3 = 18 N^2 lines of length N, timed on a sample
    line = np.random.default_rng(1).standard_normal(N) + 0j
    nl = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < min(2.0, budget_s * 0.1):
        O.fft1d(line, True)
        nl += 1
    t_line = (time.perf_counter() - t0) / nl
    t_fft = 18 * N * N * t_line
    total = t_geo + t_dep + t_fft
    sample = ("FP64 oracle, 1 thread: per level one random tube's geometry (turning angle, "
              "splines, sub-points) and a leading run of its sub-segments deposited, scaled by "
              "points/sub-segments to the whole level; 3-D FFTs from timed 1-D lines; "
              + "; ".join(notes))
    return {"value": N ** 3 / total / 1e6, "unit": "Mpts/s", "cores": 1, "kind": "port",
            "sample": sample, "seconds_estimated": total,
            "seconds_parts": {"geometry": t_geo, "deposition": t_dep, "fft": t_fft}}


def time_to_target(n_targets: int, n_sweeps: int):
    """Time-to-target-spectrum on 512^3 FP32 woven fields: hidden designs of the plain method's
    family (seeded) are generated as targets; the surrogate (trained here on the GPU from the
    shipped generator-produced pairs) predicts a design and refines it until all 32 log shells are
    within +-5%; the plain sweep walks its (gamma, density, hurst) grid until one candidate is.
    Each candidate costs one full 512^3 generation plus its statistics."""
    from fastweavturb_amd import surrogate as S
    t0 = time.perf_counter()
    bank = S.CenterlineBank(S.PAIRS_RECIPE["bank_seed"])
    t_bank = time.perf_counter() - t0
    X, Y, M = S.load_pairs()
    sur = S.Surrogate(seed=1)
    t0 = time.perf_counter()
    hist = sur.fit(X, Y, M, epochs=300)
    t_train = time.perf_counter() - t0
    log(f"[ttt] bank {t_bank:.1f}s; surrogate trained on {len(X)} pairs in {t_train:.1f}s "
        f"(loss {hist[0]:.3f} -> {hist[-1]:.4f})")
    ver = S.Verifier(bank, 512)
    rows = []
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
           "training": {"pairs": int(len(X)), "train_s": t_train, "epochs": 300},
           "per_target": rows}
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


def ensemble_bench(steps: int, warmup: int, n: int = 64, graph: bool = True, tile_cache: str | None = None) -> dict:
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
    log("[c4] autotune: " + g.autotune(tile_cache).strip().replace("\n", "; "))
    for i, (p, t) in enumerate(cfgs):
        g.select(i)
        g.set_params(p)
        g.set_tubes(t)
    for _ in range(max(1, warmup)):
        g.generate()
    t0 = time.perf_counter()
    stage = np.zeros(4)
    for _ in range(steps):
        g.generate()
        stage += g.stage_times()[:4]
    dt = (time.perf_counter() - t0) / steps
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
            "statistics_ms": t_stats * 1e3, "sub_segments": int(seg["nseg_all"]), "hits": int(dep["hits"]),
            "features_first": feats[0].round(4).tolist()}


def c3_bench(densities=(0.5, 1.0, 2.0), tile_cache: str | None = None) -> dict:
    """The 1024^3 workload: 8 levels (r = 2^-l), density swept over `densities`, fields resident
    on the GPU, structure functions of orders 2-6 of u_x at every separation 1..512."""
    from fastweavturb_amd import fbb
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(1024, 32)
    log("[c3] autotune: " + g.autotune(tile_cache).strip().replace("\n", "; "))
    seps = np.arange(1, 513, dtype=np.int32)
    runs = []
    for d in densities:
        t0 = time.time()
        p, t = fbb.config_c3(d)
        log(f"[c3] density {d}: {t.nline} tubes, {t.npoint_all} points (built in {time.time() - t0:.1f}s)")
        g.set_params(p)
        g.set_tubes(t)
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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c2", "c1", "small", "c4", "c3"])
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ttt-targets", type=int, default=4,
                    help="targets of the time-to-target-spectrum measurement (0 skips it)")
    ap.add_argument("--ttt-sweeps", type=int, default=1, help="targets also searched by the plain sweep")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default FFT tile shapes")
    ap.add_argument("--tile-cache", default=os.path.join(ROOT, "fastweavturb_amd", "data", "fft_tiles.json"),
                    help="autotune cache (JSON): shapes recorded there are reused, new ones appended")
    ap.add_argument("--fft-compare", action="store_true",
                    help="also time rocFFT's transforms of the same Biot-Savart stage")
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
        print(json.dumps(c3_bench(tile_cache=args.tile_cache)), flush=True)
        return
    if args.config == "c4":
        print(json.dumps(ensemble_bench(args.steps, args.warmup, tile_cache=args.tile_cache)), flush=True)
        return
    fftc = None
    if args.fft_compare:
        # torch carries its own HIP runtime: it runs in a child process started before this
        # process touches the GPU
        import subprocess
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--rocfft-only", "512"],
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
    if not args.no_autotune:
        log("[bench] autotune: " + g.autotune(args.tile_cache).strip().replace("\n", "; "))
    for i in range(args.warmup):
        g.generate()
        log(f"[bench] warmup {i} done, stages ms {np.round(g.stage_times()[:4], 2).tolist()}")
    # every library call is stream-synchronous on return, so the clock brackets finished work
    t0 = time.perf_counter()
    stage_ms = np.zeros(4)
    for i in range(args.steps):
        g.generate()
        stage_ms += g.stage_times()[:4]
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = p.N ** 3 / (ms_per_step * 1e-3) / 1e6
    times = g.stage_times()
    kernels = {SLOT_NAMES[i]: float(times[4 + i]) for i in range(len(SLOT_NAMES)) if times[4 + i] > 0}
    seg = g.segment_stats()
    dep = g.deposit_stats()
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
    if fftc is not None:
        spectral = [k for k in ("fft_x", "fft_y", "fft_z_biot_savart", "fft_y_inv", "fft_x_inv", "spectral_reduce")
                    if k in kernels]
        fftc["ours_biot_savart_stage_ms"] = float(stage_ms[3] / args.steps)
        fftc["ours_kernels_ms"] = float(sum(kernels[k] for k in spectral))
        log(f"[bench] fft compare {json.dumps(fftc)}")
    ttt = None
    if args.ttt_targets > 0 and args.config == "c2":
        g.close()
        ttt = time_to_target(args.ttt_targets, args.ttt_sweeps)
    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(p, t, args.cpu_budget)
        log(f"[bench] cpu baseline {cpu['value']:.4g} Mpts/s (sampled in {time.time() - t0:.1f}s)")
    out = {
        "metric": METRIC, "value": value, "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if precision == 32 else "f64", "data": "synthetic",
        "config": {"workload": wname, "N": p.N, "tubes": t.nline, "points": int(t.npoint_all),
                   "sub_segments": int(seg["nseg_all"]), "levels": 6 if args.config == "c2" else None},
        "stages_ms": dict(zip(["segments", "bin", "deposit", "biot_savart"], (stage_ms / args.steps).tolist())),
        "kernels_ms": kernels, "roofline": roof, "cpu_baseline": cpu, "time_to_target": ttt,
        "fft_vs_rocfft": fftc,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
