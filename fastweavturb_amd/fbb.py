"""Seeded fractional-Brownian-bridge (fBb) centerline generator for the woven-field workloads.

The reference builds its centerlines offline in MATLAB (FBB_with_parameters.m) with the
Wavelet Toolbox's wfbm, which is not reproducible here. This module restates the same
construction with an exact fBm sampler:
  - fractional Gaussian noise by circulant embedding (Davies & Harte 1987), cumulated to fBm;
  - the bridge B(t) - f(t) B(1), f(t) = (t^2H + 1 - (1-t)^2H)/2 (FBB_with_parameters.m:92-102);
  - step-rms normalisation of the 3-D increments to dB (FBB_with_parameters.m:104-112);
  - a uniform random shift and a wrap into [-pi, pi) (FBB_with_parameters.m:117-120).
A tube also carries a winding drift 2*pi*w*t along an integer lattice vector w nearest to its
level's mean axis direction, so that a closed periodic centerline has a mean direction.
Everything is seeded per tube (seed + global tube index), so subsets are reproducible.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .woven_io import RunParams, Tubes


_EIG_CACHE: dict = {}


def _smooth_at_least(n: int) -> int:
    """Smallest 2^a 3^b >= n (fast FFT length)."""
    best = 1 << max(0, (n - 1).bit_length())
    p3 = 1
    while p3 <= best:
        v = p3
        while v < n:
            v *= 2
        best = min(best, v)
        p3 *= 3
    return best


def _circulant_sqrt_eigs(n: int, H: float) -> np.ndarray:
    """sqrt(lambda / 2n) of the circulant embedding of the fGn covariance (cached per n, H)."""
    key = (n, float(H))
    v = _EIG_CACHE.get(key)
    if v is None:
        k = np.arange(0, n + 1, dtype=np.float64)
        gam = 0.5 * (np.abs(k + 1) ** (2 * H) - 2 * np.abs(k) ** (2 * H) + np.abs(k - 1) ** (2 * H))
        row = np.concatenate([gam, gam[-2:0:-1]])          # length 2n circulant first row
        lam = np.maximum(np.fft.fft(row).real, 0.0)        # exact method: lam >= 0 up to round-off
        v = np.sqrt(lam / len(row))
        if len(_EIG_CACHE) > 64:
            _EIG_CACHE.clear()
        _EIG_CACHE[key] = v
    return v


def embedding_size(n: int) -> int:
    """Circulant embedding length m = 2 n' (n' = smallest 2^a 3^b >= n) of an n-sample fGn draw."""
    return 2 * _smooth_at_least(n)


def fgn_from_normals(n: int, H: float, zre: np.ndarray, zim: np.ndarray, count: int) -> np.ndarray:
    """(count, n) fGn samples from given standard normals zre, zim of shape (pairs, m): the real
    and imaginary parts of FFT(sqrt(lambda/m) (zre + i zim)), truncated to n."""
    s = _circulant_sqrt_eigs(_smooth_at_least(n), H)
    z = zre + 1j * zim
    w = np.fft.fft(s[None, :] * z, axis=1)[:, :n]
    out = np.concatenate([w.real, w.imag], axis=0)
    return out[:count]


def fgn_davies_harte(n: int, H: float, rng: np.random.Generator, count: int = 1) -> np.ndarray:
    """(count, n) samples of unit-variance fractional Gaussian noise with Hurst index H.

    Generated at a smooth length n' >= n and truncated (a prefix of fGn is fGn). The real and
    imaginary parts of one complex circulant draw are two independent samples."""
    if n <= 1:
        return rng.standard_normal((count, max(n, 0)))
    m = embedding_size(n)
    pairs = (count + 1) // 2
    zre = rng.standard_normal((pairs, m))
    zim = rng.standard_normal((pairs, m))
    return fgn_from_normals(n, H, zre, zim, count)


def bridge_from_fgn(npts: int, H: float, g: np.ndarray) -> np.ndarray:
    """(3, npts) closed fractional Brownian bridge from (3, npts - 1) fGn samples
    (FBB_with_parameters.m:92-102)."""
    j = np.arange(1, npts + 1, dtype=np.float64) / npts
    f = 0.5 * (j ** (2 * H) + 1.0 - (1.0 - j) ** (2 * H))
    B = np.concatenate([np.zeros((3, 1)), np.cumsum(g, axis=1)], axis=1)
    return B - f[None, :] * B[:, -1:]


def fbb_path(npts: int, H: float, rng: np.random.Generator) -> np.ndarray:
    """(3, npts) closed fractional Brownian bridge with unnormalised amplitude."""
    return bridge_from_fgn(npts, H, fgn_davies_harte(npts - 1, H, rng, 3))


def lattice_winding(direction: np.ndarray) -> np.ndarray:
    """Integer lattice vector in {-1,0,1}^3 nearest in direction to a unit vector."""
    d = np.asarray(direction, dtype=np.float64)
    w = np.rint(d / np.max(np.abs(d))).astype(np.int64)
    if not np.any(w):
        w[np.argmax(np.abs(d))] = 1 if d[np.argmax(np.abs(d))] >= 0 else -1
    return w


def tube_centerline(npts: int, H: float, winding: np.ndarray, wiggle: float,
                    rng: np.random.Generator) -> np.ndarray:
    """(3, npts) centerline wrapped into [-pi, pi): shift + drift + normalised fBb."""
    bb = fbb_path(npts, H, rng)
    return _place(bb, npts, winding, wiggle, rng.random(3))


def tube_centerline_from(npts: int, H: float, winding: np.ndarray, wiggle: float, zre: np.ndarray,
                         zim: np.ndarray, u3: np.ndarray) -> np.ndarray:
    """tube_centerline from given random inputs: normals zre, zim of shape (2, m),
    m = embedding_size(npts - 1), and 3 uniforms u3 for the shift."""
    bb = bridge_from_fgn(npts, H, fgn_from_normals(npts - 1, H, zre, zim, 3))
    return _place(bb, npts, winding, wiggle, u3)


def _place(bb: np.ndarray, npts: int, winding: np.ndarray, wiggle: float, u3: np.ndarray) -> np.ndarray:
    """Step-rms normalisation, shift, winding drift and wrap (FBB_with_parameters.m:104-120)."""
    steps = np.diff(bb, axis=1)
    rms = float(np.sum(steps ** 2) / (npts - 1))
    drift_step = 2.0 * np.pi * float(np.linalg.norm(winding)) / npts
    dB = wiggle * drift_step
    bb *= dB / math.sqrt(rms) if rms > 0 else 0.0
    t = np.arange(npts, dtype=np.float64) / npts
    x0 = (np.asarray(u3, dtype=np.float64) - 0.5) * 2.0 * np.pi
    x = x0[:, None] + 2.0 * np.pi * winding[:, None] * t[None, :] + bb
    return np.mod(x + np.pi, 2.0 * np.pi) - np.pi


@dataclasses.dataclass
class Level:
    radius: float          # core radius r; sigma = r / sqrt(2)
    ntubes: int
    npoints: tuple         # (min, max) points per tube; max > min draws log-uniformly
    hurst: float = 0.75


def woven_config(N: int, levels: list[Level], seed: int, gamma_exp: float = 4.0 / 3.0,
                 gamma_jitter: float = 0.2, wiggle: float = 0.35, sigma_ratio: float = 1.0,
                 Rtube_ratio: float = 3.0, ntf_times: int = 50, ntf_pre: int = 1000,
                 dzeta_sup: float = 1e-4) -> tuple[RunParams, Tubes]:
    """Build the tube table and centerlines of one woven field.

    Circulation Gamma = r^gamma_exp * U(1 - jitter, 1 + jitter); sigma = r / sqrt(2) (so the
    leading Gaussian term exp(-rho^2 / 2 sigma^2) is exp(-rho^2 / r^2)); n_T = 1; no filter.
    """
    npoint, xs, sig, gam = [], [], [], []
    tube = 0
    for li, lev in enumerate(levels):
        lrng = np.random.default_rng(seed + 1_000_000 + li)
        axis = lrng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        w = lattice_winding(axis)
        for _ in range(lev.ntubes):
            rng = np.random.default_rng(seed + tube)
            lo, hi = lev.npoints
            n = lo if hi <= lo else int(math.floor(lo * (hi / lo) ** rng.random()))
            n = max(n, 8)
            xs.append(tube_centerline(n, lev.hurst, w, wiggle, rng))
            npoint.append(n)
            sig.append(lev.radius / math.sqrt(2.0))
            gam.append(lev.radius ** gamma_exp * rng.uniform(1 - gamma_jitter, 1 + gamma_jitter))
            tube += 1
    X = np.concatenate(xs, axis=1)
    t = Tubes(np.array(npoint, dtype=np.int32), X[0].copy(), X[1].copy(), X[2].copy(),
              np.array(sig), np.array(gam), np.ones(tube, dtype=np.int32),
              np.zeros(tube, dtype=np.int32))
    p = RunParams(N=N, n_d=N, line_s=1, line_e=tube, sigma_ratio=sigma_ratio, eta_t=0.0,
                  ntf_times=ntf_times, ntf_pre=ntf_pre, Rtube_ratio=Rtube_ratio, dzeta_sup=dzeta_sup)
    t.validate()
    return p, t


def config_plain_128(seed: int = 1234) -> tuple[RunParams, Tubes]:
    """128^3 plain method: 3 levels, r = 0.4/0.2/0.1, 8/32/128 tubes of 1024 points, H = 0.75."""
    levels = [Level(0.4, 8, (1024, 1024)), Level(0.2, 32, (1024, 1024)), Level(0.1, 128, (1024, 1024))]
    return woven_config(128, levels, seed)


def config_batched_512(seed: int = 1234, density: float = 1.0) -> tuple[RunParams, Tubes]:
    """512^3: 6 levels, r = 0.8 * 2^-l, round(density * 4 * 4^l) tubes, 256*2^l..256*2^(l+1) points."""
    levels = [Level(0.8 * 2.0 ** -l, max(1, int(round(density * 4 * 4 ** l))),
                    (256 * 2 ** l, 256 * 2 ** (l + 1))) for l in range(6)]
    return woven_config(512, levels, seed)


def config_small(N: int = 32, seed: int = 7, ntubes=(1, 2), npts=(64, 48), radii=(0.5, 0.3),
                 sigma_ratio: float = 1.0, ntf_pre: int = 100, **kw) -> tuple[RunParams, Tubes]:
    """Tiny multi-level case for parity tests (seconds on the CPU oracle)."""
    levels = [Level(r, n, (p, p)) for r, n, p in zip(radii, ntubes, npts)]
    return woven_config(N, levels, seed, sigma_ratio=sigma_ratio, ntf_pre=ntf_pre, **kw)


@dataclasses.dataclass
class LevelSpec:
    """One scale level of an ensemble or inverse-design instance: every tube of the level has
    the same point count, so the level is drawn as one batch."""
    radius: float          # core radius r; sigma = r / sqrt(2)
    ntubes: int
    npoints: int
    hurst: float
    gamma: float           # circulation of the level before the per-tube jitter


def woven_levels(N: int, levels: list, seed: int, gamma_jitter: float = 0.2, wiggle: float = 0.35,
                 Rtube_ratio: float = 3.0, ntf_times: int = 50, ntf_pre: int = 1000,
                 dzeta_sup: float = 1e-4, chunk: int = 1024) -> tuple[RunParams, Tubes]:
    """Tube table of one instance from LevelSpecs, with the construction of `woven_config`
    (fBb bridge, step-rms normalisation, winding drift, shift and wrap) but one random stream per
    level (seeded by (seed, level)) and the fGn of a level's tubes drawn in batches of `chunk`.
    Subsets of tubes are therefore not reproducible on their own, unlike `woven_config`."""
    npoint, xs, sig, gam = [], [], [], []
    for li, lev in enumerate(levels):
        rng = np.random.default_rng([int(seed), li])
        axis = rng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        w = lattice_winding(axis).astype(np.float64)
        n = max(8, int(lev.npoints))
        j = np.arange(1, n + 1, dtype=np.float64) / n
        f = 0.5 * (j ** (2 * lev.hurst) + 1.0 - (1.0 - j) ** (2 * lev.hurst))
        t = np.arange(n, dtype=np.float64) / n
        dB = wiggle * 2.0 * np.pi * float(np.linalg.norm(w)) / n
        for c0 in range(0, lev.ntubes, chunk):
            cnt = min(chunk, lev.ntubes - c0)
            g = fgn_davies_harte(n - 1, lev.hurst, rng, 3 * cnt).reshape(cnt, 3, n - 1)
            B = np.concatenate([np.zeros((cnt, 3, 1)), np.cumsum(g, axis=2)], axis=2)
            B -= f[None, None, :] * B[:, :, -1:]
            rms = np.sum(np.diff(B, axis=2) ** 2, axis=(1, 2)) / (n - 1)
            B *= (dB / np.sqrt(np.maximum(rms, 1e-300)))[:, None, None]
            x0 = (rng.random((cnt, 3)) - 0.5) * 2.0 * np.pi
            X = x0[:, :, None] + 2.0 * np.pi * w[None, :, None] * t[None, None, :] + B
            X = np.mod(X + np.pi, 2.0 * np.pi) - np.pi
            xs.append(X.transpose(1, 0, 2).reshape(3, cnt * n))
            npoint.extend([n] * cnt)
            sig.extend([lev.radius / math.sqrt(2.0)] * cnt)
            gam.append(lev.gamma * rng.uniform(1 - gamma_jitter, 1 + gamma_jitter, cnt))
    X = np.concatenate(xs, axis=1)
    nt = len(npoint)
    tb = Tubes(np.array(npoint, dtype=np.int32), X[0].copy(), X[1].copy(), X[2].copy(),
               np.array(sig), np.concatenate(gam), np.ones(nt, dtype=np.int32), np.zeros(nt, dtype=np.int32))
    p = RunParams(N=N, n_d=N, line_s=1, line_e=nt, sigma_ratio=1.0, eta_t=0.0, ntf_times=ntf_times,
                  ntf_pre=ntf_pre, Rtube_ratio=Rtube_ratio, dzeta_sup=dzeta_sup)
    tb.validate()
    return p, tb


def config_c3(density: float = 1.0, seed: int = 1234, N: int = 1024, nlev: int = 8, max_points: int = 1024):
    """1024^3: nlev levels, r = 2^-l, round(density * 4 * 4^l) tubes, min(256 * 2^l, max_points) points
    per tube (the cap keeps the centerlines at ~90M points), H = 0.75, Gamma ~ r^(4/3)."""
    levels = [LevelSpec(2.0 ** -l, max(1, int(round(density * 4 * 4 ** l))), int(min(256 * 2 ** l, max_points)),
                        0.75, (2.0 ** -l) ** (4.0 / 3.0)) for l in range(nlev)]
    return woven_levels(N, levels, seed)



# ---- the same constructions with the centerlines drawn on the GPU ------------------------------
def woven_config_device(g, N: int, levels: list[Level], seed: int, gamma_exp: float = 4.0 / 3.0,
                        gamma_jitter: float = 0.2, wiggle: float = 0.35, sigma_ratio: float = 1.0,
                        Rtube_ratio: float = 3.0, ntf_times: int = 50, ntf_pre: int = 1000,
                        dzeta_sup: float = 1e-4) -> tuple[RunParams, Tubes]:
    """woven_config with the centerlines generated on the GPU of WovenGPU context g
    (wt_fbb_centerlines, Philox stream keyed (seed, tube)). Per tube the host draws, from
    default_rng(seed + tube), the point count and then the circulation jitter; the level axes are
    woven_config's."""
    npoint, H, W, sig, gam = [], [], [], [], []
    tube = 0
    for li, lev in enumerate(levels):
        lrng = np.random.default_rng(seed + 1_000_000 + li)
        axis = lrng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        w = lattice_winding(axis)
        for _ in range(lev.ntubes):
            rng = np.random.default_rng(seed + tube)
            lo, hi = lev.npoints
            n = lo if hi <= lo else int(math.floor(lo * (hi / lo) ** rng.random()))
            npoint.append(max(n, 8))
            H.append(lev.hurst)
            W.append(w)
            sig.append(lev.radius / math.sqrt(2.0))
            gam.append(lev.radius ** gamma_exp * rng.uniform(1 - gamma_jitter, 1 + gamma_jitter))
            tube += 1
    cx, cy, cz = g.fbb_centerlines(npoint, H, np.array(W), wiggle, seed)
    t = Tubes(np.array(npoint, dtype=np.int32), cx, cy, cz, np.array(sig), np.array(gam),
              np.ones(tube, dtype=np.int32), np.zeros(tube, dtype=np.int32))
    p = RunParams(N=N, n_d=N, line_s=1, line_e=tube, sigma_ratio=sigma_ratio, eta_t=0.0,
                  ntf_times=ntf_times, ntf_pre=ntf_pre, Rtube_ratio=Rtube_ratio, dzeta_sup=dzeta_sup)
    t.validate()
    return p, t


def woven_levels_device(g, N: int, levels: list, seed: int, gamma_jitter: float = 0.2, wiggle: float = 0.35,
                        Rtube_ratio: float = 3.0, ntf_times: int = 50, ntf_pre: int = 1000,
                        dzeta_sup: float = 1e-4) -> tuple[RunParams, Tubes]:
    """woven_levels (LevelSpecs: one point count and Hurst index per level) with the centerlines
    generated on the GPU of context g; per level default_rng((seed, level)) draws the axis and the
    circulation jitters."""
    npoint, H, W, sig, gam = [], [], [], [], []
    for li, lev in enumerate(levels):
        rng = np.random.default_rng([int(seed), li])
        axis = rng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        w = lattice_winding(axis)
        n = max(8, int(lev.npoints))
        npoint.extend([n] * lev.ntubes)
        H.extend([lev.hurst] * lev.ntubes)
        W.extend([w] * lev.ntubes)
        sig.extend([lev.radius / math.sqrt(2.0)] * lev.ntubes)
        gam.append(lev.gamma * rng.uniform(1 - gamma_jitter, 1 + gamma_jitter, lev.ntubes))
    nt = len(npoint)
    cx, cy, cz = g.fbb_centerlines(npoint, H, np.array(W), wiggle, seed)
    tb = Tubes(np.array(npoint, dtype=np.int32), cx, cy, cz, np.array(sig), np.concatenate(gam),
               np.ones(nt, dtype=np.int32), np.zeros(nt, dtype=np.int32))
    p = RunParams(N=N, n_d=N, line_s=1, line_e=nt, sigma_ratio=1.0, eta_t=0.0, ntf_times=ntf_times,
                  ntf_pre=ntf_pre, Rtube_ratio=Rtube_ratio, dzeta_sup=dzeta_sup)
    tb.validate()
    return p, tb


# ---- the reference's own construction (FBB_with_parameters.m) ------------------------------------
@dataclasses.dataclass
class ReferenceFamily:
    """The header of FBB_with_parameters.m (lines 5-22), with its values as defaults: grid N,
    Kolmogorov scale eta, levels Nv, core-to-eta ratio c_e_s, target vortex volume fraction
    rho_w_obj, tube-count ratio r_nL, sigma_ratio, step-to-core ratio c_B, scale ratio r_s, Hurst
    index H, spectral exponent r_E, first circulation Gamma_1, first n_T."""
    N: int = 256
    eta: float = 0.00888684
    Nv: int = 3
    c_e_s: float = 1.7
    rho_w_obj: float = 0.024
    r_nL: float = 4.0
    sigma_ratio: float = 4.0
    c_B: float = 40.0
    r_s: float = 0.5
    H: float = 0.834
    r_E: float = 5.0 / 3.0
    Gamma_1: float = 1.0
    n_T_1: int = 1


def _mround(x: float) -> int:
    """MATLAB's round: half away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def reference_levels(f: ReferenceFamily) -> dict:
    """Per level i = 1..Nv of FBB_with_parameters.m:25-48 and 188-193: tube count m_line, points Np,
    step rms dB, core sigma, circulation Gamma and n_T, in the script's FP64 arithmetic."""
    r_line = math.floor(f.r_nL / f.r_s + 0.5)
    r_Np = math.floor(f.r_nL / f.r_s / r_line)
    r_rho = f.r_s * f.r_s * f.r_nL
    if r_rho != 1:
        rho_Nv = f.rho_w_obj / (1 - r_rho ** f.Nv) * (1 - r_rho) * r_rho ** (f.Nv - 1)
    else:
        rho_Nv = f.rho_w_obj / f.Nv
    sigma_Nv = f.c_e_s * f.eta
    sigma_1 = sigma_Nv / (f.r_s ** (f.Nv - 1))
    dB_Nv = f.c_B * sigma_Nv
    dB_1 = dB_Nv / f.r_s ** (f.Nv - 1)
    L_Nv_temp = rho_Nv * (2 * math.pi) ** 3 / sigma_Nv ** 2
    L_1_temp = L_Nv_temp / f.r_nL ** (f.Nv - 1)
    Np_1 = _mround(L_1_temp / (1.2 * dB_1))
    r_Gamma = f.r_s ** ((f.r_E - 1) / 2) * f.r_nL ** (-1 / 2)
    r_n_T = _mround(1 / f.r_s)
    i = range(f.Nv)
    return {"m_line": [int(r_line ** k) for k in i], "Np": [int(Np_1 * r_Np ** k) for k in i],
            "dB": [float(dB_1 * f.r_s ** k) for k in i], "sigma": [float(sigma_1 * f.r_s ** k) for k in i],
            "Gamma": [float(f.Gamma_1 * r_Gamma ** k) for k in i], "n_T": [int(f.n_T_1 * r_n_T ** k) for k in i]}


def reference_centerline_from(npts: int, H: float, dB: float, zre: np.ndarray, zim: np.ndarray,
                              u3: np.ndarray) -> np.ndarray:
    """(3, npts) centerline as FBB_with_parameters.m:73-120 builds it, from given random inputs
    (normals zre, zim of shape (2, m), m = embedding_size(max(npts, 128) - 1), and 3 uniforms u3):
    the fBm path P (P_0 = 0) of max(npts, 128) points, for npts < 128 its points at
    ceil(j 128 / npts) - 1, j = 1..npts (FBB:78-84); the bridge P_j - f_j P_npts (FBB:92-102); the
    step rms scaled to dB (FBB:104-112); the shift (u - 1/2) 2 pi and mod(x, 2 pi) - pi (FBB:117-120)."""
    ngen = 128 if npts < 128 else npts
    g = fgn_from_normals(ngen - 1, H, zre, zim, 3)
    P = np.concatenate([np.zeros((3, 1)), np.cumsum(g, axis=1)], axis=1)
    if ngen != npts:
        P = P[:, np.ceil(np.arange(1, npts + 1, dtype=np.float64) * (ngen / npts)).astype(np.int64) - 1]
    j = np.arange(1, npts + 1, dtype=np.float64) / npts
    f = 0.5 * ((j ** (2 * H) + 1.0) - (1.0 - j) ** (2 * H))
    bb = P - f[None, :] * P[:, -1:]
    steps = np.diff(bb, axis=1)
    rms = float(np.sum(steps ** 2) / (npts - 1))
    bb = bb * (dB / math.sqrt(rms) if rms > 0 else 0.0)
    x = bb + ((np.asarray(u3, dtype=np.float64) - 0.5) * 2.0 * np.pi)[:, None]
    return (x - np.floor(x / (2.0 * np.pi)) * (2.0 * np.pi)) - np.pi


def philox_inputs(seed: int, key: int, m: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """A tube's random inputs from the stream the GPU generator uses: numpy's Philox4x64-10 keyed
    (seed, key), uniforms (word >> 11) 2^-53, normals by Box-Muller on consecutive uniform pairs;
    4m normals (zre: rows 0, 1; zim: rows 0, 1) and then 3 uniforms."""
    raw = np.random.Philox(key=(int(key) << 64) | int(seed)).random_raw(4 * m + 3)
    u = (raw >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(1.0 - u[0:4 * m:2]))
    z = np.empty(4 * m)
    z[0::2] = r * np.cos(2 * np.pi * u[1:4 * m:2])
    z[1::2] = r * np.sin(2 * np.pi * u[1:4 * m:2])
    return z[:2 * m].reshape(2, m), z[2 * m:].reshape(2, m), u[4 * m:4 * m + 3]


def reference_inputs(f: ReferenceFamily, seed: int, g=None) -> tuple[RunParams, Tubes]:
    """The two input files' contents as FBB_with_parameters.m writes them for family f: tubes level
    by level, tube l (0-based, in file order) with the random stream keyed (seed, l); centerlines on
    the GPU of WovenGPU context g (wt_fbb_reference_centerlines) or, with g None, on the host from
    the same stream. Parameters as the script writes them (FBB:198-243): N_downsample = N,
    line_s = 1, line_e = all tubes, eta_t = 0, ntf_times 50, ntf_pre 1000, Rtube_ratio 3,
    dzeta_sup 1e-4, if_filter 0."""
    lv = reference_levels(f)
    npoint, dB, sig, gam, nT = [], [], [], [], []
    for i in range(f.Nv):
        c = lv["m_line"][i]
        npoint += [lv["Np"][i]] * c
        dB += [lv["dB"][i]] * c
        sig += [lv["sigma"][i]] * c
        gam += [lv["Gamma"][i]] * c
        nT += [lv["n_T"][i]] * c
    nt = len(npoint)
    if g is not None:
        cx, cy, cz = g.fbb_reference_centerlines(npoint, [f.H] * nt, dB, seed)
    else:
        xs = []
        for l in range(nt):
            m = embedding_size(max(npoint[l], 128) - 1)
            zre, zim, u3 = philox_inputs(seed, l, m)
            xs.append(reference_centerline_from(npoint[l], f.H, dB[l], zre, zim, u3))
        cx, cy, cz = np.concatenate(xs, axis=1)
    t = Tubes(np.array(npoint, dtype=np.int32), np.ascontiguousarray(cx), np.ascontiguousarray(cy),
              np.ascontiguousarray(cz), np.array(sig), np.array(gam), np.array(nT, dtype=np.int32),
              np.zeros(nt, dtype=np.int32))
    p = RunParams(N=f.N, n_d=f.N, line_s=1, line_e=nt, sigma_ratio=f.sigma_ratio, eta_t=0.0, ntf_times=50,
                  ntf_pre=1000, Rtube_ratio=3.0, dzeta_sup=1e-4)
    t.validate()
    return p, t


def main(argv=None) -> int:
    """Writes centerline_input.dat and parameters_input.dat as FBB_with_parameters.m does, for the
    family given by its header values (defaults: the script's)."""
    import argparse
    from .woven_io import write_inputs
    d = ReferenceFamily()
    ap = argparse.ArgumentParser(description="the reference's fBb input files (FBB_with_parameters.m)")
    ap.add_argument("--out", required=True, help="directory for the two files")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--host", action="store_true", help="draw the centerlines on the host instead of the GPU")
    for f in dataclasses.fields(ReferenceFamily):
        ap.add_argument("--" + f.name.replace("_", "-"), type=type(getattr(d, f.name)), default=getattr(d, f.name))
    a = ap.parse_args(argv)
    fam = ReferenceFamily(**{f.name: getattr(a, f.name) for f in dataclasses.fields(ReferenceFamily)})
    g = None
    if not a.host:
        from .api import WovenGPU
        g = WovenGPU(16, 64)
    try:
        p, t = reference_inputs(fam, a.seed, g)
    finally:
        if g is not None:
            g.close()
    write_inputs(a.out, p, t, reference_digits=True)
    print(f"{t.nline} tubes, {t.npoint_all} points written to {a.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
