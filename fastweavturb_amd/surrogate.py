"""Inverse design of woven fields: a learned surrogate that maps target statistics to per-level
tube parameters, and the plain parameter sweep it is measured against.

The reference sets its level parameters by closed-form scalings (FBB_with_parameters.m:25-41,
185-193: radius ratio r_s, density, r_Gamma from r_E, Hurst index H) and has no inverse design.
Here:

* Design (24 numbers, 6 levels x 4): per level l
    a_l = log2(Gamma_l / Gamma_0)   (a_0 = 0: the normalisation u' = 1 makes a common
                                     circulation factor unobservable)
    b_l = log2(density_l)           tubes_l = max(1, round(2^b_l * 4 * 4^l))
    q_l = radius ratio              r_l = r_{l-1} / q_l, r_{-1} = 1.6
    H_l = Hurst index of the level's fBb centerlines
  points per tube n_l = clamp(round(51.2 / r_l), 32, 1024); sigma = r / sqrt(2).
* Statistics (48 numbers) of the normalised velocity: 32 log shells of E(k) (Gaussian-weighted
  in ln k around 32 centres log-spaced over 1 <= k <= 16) and the flatness F4 = S4/S2^2 and
  F6 = S6/S2^3 of x-increments of u_x at the separations m * 2pi/32, m = 1,2,3,4,6,8,12,16.
  Both are defined in box units, so instances of every grid size N in {64, 96, 128, 192, 256}
  (and the 512^3 verification grid) describe the same scales.
* Pairs (statistics -> design) come from ensembles of uneven instances generated by the batched
  library path (one context per ensemble; geometry, binning and deposition as single launches).
* The network (48 -> 3 x 256 GELU -> 24, BF16 MFMA kernels of the library) is trained on
  standardised features and designs, with the slots of absent levels masked out of the loss.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import fbb
from .woven_io import RunParams, Tubes

NLEV = 6
NPAR = 4
NSHELL = 32
K_LO, K_HI = 1.0, 16.0
FLAT_M = (1, 2, 3, 4, 6, 8, 12, 16)
NFEAT = NSHELL + 2 * len(FLAT_M)
GRIDS = (64, 96, 128, 192, 256)
R_TOP = 1.6
SURROGATE_SIZES = [NFEAT, 256, 256, 256, NLEV * NPAR]

# parameter ranges of the designs (the box the sweep and the draws live in)
RANGE = {"b": (-1.0, 1.0), "q": (1.6, 2.4), "H": (0.5, 0.9)}


@dataclasses.dataclass
class Design:
    a: np.ndarray          # [L] log2 Gamma_l / Gamma_0, a[0] = 0
    b: np.ndarray          # [L] log2 density
    q: np.ndarray          # [L] radius ratio
    H: np.ndarray          # [L] Hurst index

    @property
    def nlev(self) -> int:
        return len(self.q)

    def radii(self) -> np.ndarray:
        return R_TOP / np.cumprod(self.q)

    def levels(self) -> list:
        r = self.radii()
        out = []
        for l in range(self.nlev):
            nt = max(1, int(round(2.0 ** self.b[l] * 4 * 4 ** l)))
            npts = int(min(1024, max(32, round(51.2 / r[l]))))
            gam = r[0] ** (4.0 / 3.0) * 2.0 ** self.a[l]
            out.append(fbb.LevelSpec(float(r[l]), nt, npts, float(self.H[l]), float(gam)))
        return out

    def vector(self) -> tuple[np.ndarray, np.ndarray]:
        """(24-vector, mask): per level (a, b, q, H); absent levels are zero with mask 0."""
        v = np.zeros((NLEV, NPAR))
        m = np.zeros((NLEV, NPAR))
        L = self.nlev
        v[:L] = np.stack([self.a, self.b, self.q, self.H], axis=1)
        m[:L] = 1.0
        m[0, 0] = 0.0      # a_0 is 0 by definition
        return v.ravel(), m.ravel()

    @staticmethod
    def from_vector(v, nlev: int = NLEV) -> "Design":
        v = np.asarray(v, dtype=np.float64).reshape(NLEV, NPAR)[:nlev]
        a = v[:, 0].copy()
        a[0] = 0.0
        return Design(a, np.clip(v[:, 1], *RANGE["b"]), np.clip(v[:, 2], *RANGE["q"]), np.clip(v[:, 3], *RANGE["H"]))


def max_levels(N: int, q) -> int:
    """Levels whose core radius is at least one grid spacing 2pi/N (at least 3)."""
    r = R_TOP / np.cumprod(np.asarray(q, dtype=np.float64))
    return max(3, int(np.sum(r >= 2.0 * np.pi / N)))


def draw_design(rng: np.random.Generator, N: int, nlev: int | None = None) -> Design:
    """Ensemble draw: instance-wide Hurst index U[0.5, 0.9], density U[0.5, 2] and circulation
    exponent U[1.1, 1.5] (Gamma_l ~ r_l^gamma), each level jittered around them; radius ratios
    U[1.6, 2.4]; 3..(levels resolved by N) levels when nlev is not given."""
    q = rng.uniform(*RANGE["q"], NLEV)
    if nlev is None:
        nlev = int(rng.integers(3, min(NLEV, max_levels(N, q)) + 1))
    q = q[:nlev]
    H0, d0, g0 = rng.uniform(0.5, 0.9), rng.uniform(0.5, 2.0), rng.uniform(1.1, 1.5)
    r = R_TOP / np.cumprod(q)
    a = g0 * np.log2(r / r[0]) + rng.uniform(-0.25, 0.25, nlev)
    a[0] = 0.0
    b = np.clip(math.log2(d0) + rng.uniform(-0.3, 0.3, nlev), *RANGE["b"])
    H = np.clip(H0 + rng.uniform(-0.08, 0.08, nlev), *RANGE["H"])
    return Design(a, b, q, H)


def instance_config(design: Design, N: int, seed: int, ntf_pre: int = 1000):
    return fbb.woven_levels(N, design.levels(), seed, ntf_pre=ntf_pre)


# ---- statistics -----------------------------------------------------------------------------
def shell_weights(nek: int) -> np.ndarray:
    """(32, nek) Gaussian weights in ln k of the integer shells around the log-spaced centres."""
    j = np.arange(NSHELL)
    lc = math.log(K_LO) + (math.log(K_HI) - math.log(K_LO)) * j / (NSHELL - 1)
    s = (math.log(K_HI) - math.log(K_LO)) / (NSHELL - 1)
    lk = np.log(np.arange(1, nek + 1, dtype=np.float64))
    w = np.exp(-0.5 * ((lk[None, :] - lc[:, None]) / s) ** 2)
    return w / w.sum(axis=1, keepdims=True)


def log_shells(E) -> np.ndarray:
    """32 log shells of an integer-shell spectrum E[i-1], i = 1..nek."""
    E = np.asarray(E, dtype=np.float64)
    return shell_weights(len(E)) @ E


def flat_separations(N: int) -> np.ndarray:
    if N % 32:
        raise ValueError("flatness separations need N divisible by 32")
    return np.array([m * N // 32 for m in FLAT_M], dtype=np.int32)


def features(E, S) -> np.ndarray:
    """48 statistics from the integer-shell spectrum and S[p-2][m] (p = 2..6) of x-increments."""
    S = np.asarray(S, dtype=np.float64)
    F4 = S[2] / S[0] ** 2
    F6 = S[4] / S[0] ** 3
    return np.concatenate([np.log10(np.maximum(log_shells(E), 1e-300)), np.log10(F4), np.log10(F6)])


def measure(g, inst: int | None = None) -> np.ndarray:
    """Statistics of instance `inst` of a generated WovenGPU context."""
    if inst is not None:
        g.select(inst)
    E = g.spectrum(1)
    S = g.structure_functions(1, 2, 6, flat_separations(g.N))
    return features(E, S)


# ---- fixed-seed centerline bank ---------------------------------------------------------------
HB = (0.5, 0.6, 0.7, 0.8, 0.9)


def level_points(l: int) -> int:
    """Points per tube of level l (64 * 2^l, at most 1024); independent of the radius so that a
    design's geometry changes continuously with its radius ratios."""
    return int(min(1024, 64 * 2 ** l))


def level_max_tubes(l: int) -> int:
    return 2 * 4 ** (l + 1)


class CenterlineBank:
    """Fixed-seed centerlines of the inverse design, drawn on the GPU by the fBb generator
    (wt_fbb_centerlines). Per level l: tube i has the Philox key l * 2^32 + i under the bank's
    seed, so its random numbers are the same for every design; level_points(l) points; the level's
    winding vector and each tube's circulation jitter U(0.8, 1.2) come from default_rng((seed, l, 0)).
    A design takes the first tubes of each level, with the design's Hurst index of that level
    (the same normals for every H, so the field changes continuously with H) and the wiggle 0.35 of
    the drift step (as `fbb.woven_config`). A non-integer tube count c gives ceil(c) tubes, the last
    with its circulation scaled by the fraction, so the field changes continuously with every
    design parameter, and equal designs give identical fields."""

    def __init__(self, seed: int, nlev: int = NLEV, wiggle: float = 0.35):
        self.seed, self.nlev, self.wiggle = int(seed), nlev, wiggle
        self.w, self.u = [], []
        for l in range(nlev):
            r0 = np.random.default_rng([self.seed, l, 0])
            axis = r0.standard_normal(3)
            axis /= np.linalg.norm(axis)
            self.w.append(fbb.lattice_winding(axis).astype(np.int32))
            self.u.append(r0.uniform(0.8, 1.2, level_max_tubes(l)))

    def tubes(self, design: Design, g):
        """Tube table of a design (sigma = r / sqrt(2), Gamma_l = r_0^(4/3) 2^a_l times the tube's
        jitter, n_T = 1, no filter), centerlines generated on the GPU of WovenGPU context g."""
        if design.nlev > self.nlev:
            raise ValueError("design has more levels than the bank")
        r = design.radii()
        npoint, H, W, keys, sig, gam = [], [], [], [], [], []
        for l in range(design.nlev):
            creal = min(float(level_max_tubes(l)), max(1.0, 2.0 ** design.b[l] * 4 * 4 ** l))
            cnt = int(math.ceil(creal - 1e-9))
            npoint.extend([level_points(l)] * cnt)
            H.extend([float(np.clip(design.H[l], HB[0], HB[-1]))] * cnt)
            W.extend([self.w[l]] * cnt)
            keys.extend(((l << 32) + np.arange(cnt)).tolist())
            sig.extend([r[l] / math.sqrt(2.0)] * cnt)
            gl = r[0] ** (4.0 / 3.0) * 2.0 ** design.a[l] * np.array(self.u[l][:cnt], dtype=np.float64)
            gl[-1] *= creal - (cnt - 1)
            gam.append(gl)
        cx, cy, cz = g.fbb_centerlines(npoint, H, np.array(W), self.wiggle, self.seed, keys=np.array(keys, np.uint64))
        nt = len(npoint)
        return Tubes(np.array(npoint, dtype=np.int32), cx, cy, cz, np.array(sig), np.concatenate(gam),
                     np.ones(nt, dtype=np.int32), np.zeros(nt, dtype=np.int32))

    def config(self, design: Design, N: int, g, ntf_pre: int = 1000):
        t = self.tubes(design, g)
        p = RunParams(N=N, n_d=N, line_s=1, line_e=t.nline, sigma_ratio=1.0, eta_t=0.0, ntf_times=50,
                      ntf_pre=ntf_pre, Rtube_ratio=3.0, dzeta_sup=1e-4)
        return p, t


def family_design(gamma: float, density: float, hurst: float, nlev: int = NLEV) -> Design:
    """The plain method's level law (the 512^3 workload's levels): r_l = 0.8 * 2^-l, one density,
    one Hurst index, Gamma_l ~ r_l^gamma."""
    q = np.full(nlev, 2.0)
    r = R_TOP / np.cumprod(q)
    a = gamma * np.log2(r / r[0])
    return Design(a, np.full(nlev, math.log2(density)), q, np.full(nlev, hurst))


# Densities of the inverse design. Below ~k = 4 a log shell holds only tens of Fourier modes, so
# the spectrum there is the interference pattern of the exact tube set: one tube more or less
# moves a shell by tens of percent, and no smooth model or coarse sweep can land on a continuous
# tube count. Density is therefore one of these values (the same tube set for every design that
# shares it), and the smooth parameters (circulations, radii, Hurst indices) are searched on it.
DENSITIES = tuple(float(v) for v in 2.0 ** np.arange(-1.0, 1.0001, 0.5))


def snap_density(b) -> float:
    """log2 density nearest to the mean of the given log2 densities."""
    lv = np.log2(DENSITIES)
    return float(lv[np.argmin(np.abs(lv - float(np.mean(b))))])


def draw_inverse_design(rng: np.random.Generator, jitter: bool, nlev: int = NLEV) -> Design:
    """Prior of the inverse-design pairs: the plain method's family with gamma U[1.1, 1.5],
    density from DENSITIES, H U[0.5, 0.9]; with `jitter`, each level's circulation, radius ratio
    and Hurst index perturbed too."""
    d = family_design(rng.uniform(1.1, 1.5), DENSITIES[int(rng.integers(len(DENSITIES)))], rng.uniform(0.5, 0.9),
                      nlev)
    if jitter:
        L = d.nlev
        d.a = d.a + np.concatenate([[0.0], rng.uniform(-0.3, 0.3, L - 1)])
        d.q = np.clip(d.q + rng.uniform(-0.15, 0.15, L), *RANGE["q"])
        d.H = np.clip(d.H + rng.uniform(-0.08, 0.08, L), *RANGE["H"])
    return d


# ---- pair generation ---------------------------------------------------------------------------
def generate_pairs(jobs, ensemble: int = 64, log=None, bank: CenterlineBank | None = None):
    """(statistics, design, mask, N) for each job (design, N, tube seed), from ensembles of
    `ensemble` instances, each ensemble one batched context on the GPU whose centerlines are drawn
    on the GPU too: from the bank, or (bank None) fresh from the job's seed by
    `fbb.woven_levels_device`."""
    from .api import WovenGPU
    n = len(jobs)
    X = np.zeros((n, NFEAT), dtype=np.float32)
    Y = np.zeros((n, NLEV * NPAR), dtype=np.float32)
    M = np.zeros((n, NLEV * NPAR), dtype=np.float32)
    Ns = np.zeros(n, dtype=np.int32)
    t0 = time.time()
    for lo in range(0, n, ensemble):
        chunk = jobs[lo:lo + ensemble]
        g = WovenGPU(precision=32, sizes=[N for _, N, _ in chunk])
        for k, (design, N, seed) in enumerate(chunk):
            p, t = bank.config(design, N, g) if bank is not None else \
                fbb.woven_levels_device(g, N, design.levels(), seed)
            g.select(k)
            g.set_params(p)
            g.set_tubes(t)
        g.generate()
        for k in range(len(chunk)):
            i = lo + k
            X[i] = measure(g, k)
            Y[i], M[i] = chunk[k][0].vector()
            Ns[i] = chunk[k][1]
        g.close()
        if log:
            log(f"[pairs] {lo + len(chunk)}/{n} in {time.time() - t0:.1f}s")
    return X, Y, M, Ns


def ensemble_jobs(n: int, seed: int = 5000, grids=GRIDS):
    """Uneven ensemble instances: grid size from `grids`, a design from `draw_design`, tube seed
    seed + i."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = int(rng.choice(grids))
        out.append((draw_design(rng, N), N, seed + i))
    return out


# ---- the network ---------------------------------------------------------------------------------
class Surrogate:
    """Standardisation + the library's MLP. Designs are predicted as 24-vectors."""

    def __init__(self, seed: int = 0):
        from .mlp import MLP
        self.net = MLP(SURROGATE_SIZES, seed=seed)
        self.xm = self.xs = self.ym = self.ys = None

    def fit(self, X, Y, M, epochs: int = 200, batch: int = 256, lr: float = 2e-3, lr_final: float = 2e-5,
            seed: int = 0) -> np.ndarray:
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        M = np.asarray(M, dtype=np.float64)
        self.xm, self.xs = X.mean(axis=0), X.std(axis=0) + 1e-6
        cnt = np.maximum(M.sum(axis=0), 1.0)
        self.ym = (Y * M).sum(axis=0) / cnt
        self.ys = np.sqrt(((Y - self.ym) ** 2 * M).sum(axis=0) / cnt) + 1e-6
        Xn = (X - self.xm) / self.xs
        Yn = (Y - self.ym) / self.ys * M
        self.net.set_data(Xn, Yn, M)
        return self.net.train(epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed)

    def predict(self, X) -> np.ndarray:
        Xn = (np.atleast_2d(np.asarray(X, dtype=np.float64)) - self.xm) / self.xs
        return self.net.forward(Xn).astype(np.float64) * self.ys + self.ym

    def state(self) -> dict:
        return {"params": self.net.get_params(), "xm": self.xm, "xs": self.xs, "ym": self.ym, "ys": self.ys}

    def load(self, st: dict):
        self.net.set_params(st["params"])
        self.xm, self.xs, self.ym, self.ys = st["xm"], st["xs"], st["ym"], st["ys"]


def masked_rmse(P, Y, M) -> np.ndarray:
    """Per-slot RMS error over the pairs where the slot is present."""
    cnt = np.maximum(M.sum(axis=0), 1.0)
    return np.sqrt(((P - Y) ** 2 * M).sum(axis=0) / cnt)


# ---- inverse design ------------------------------------------------------------------------------
def shell_error(f, target) -> float:
    """Largest relative deviation of the 32 log shells of statistics f from the target's."""
    return float(np.max(np.abs(10.0 ** (np.asarray(f)[:NSHELL] - np.asarray(target)[:NSHELL]) - 1.0)))


class Verifier:
    """Runs designs as woven fields from a centerline bank on one GPU context of grid N and
    measures their statistics; counts runs and wall time (centerlines, generation, statistics)."""

    def __init__(self, bank: CenterlineBank, N: int = 512):
        from .api import WovenGPU
        self.bank, self.N = bank, N
        self.nlev = bank.nlev
        self.g = WovenGPU(N, 32)
        self.runs = 0
        self.seconds = 0.0

    def run(self, design: Design) -> np.ndarray:
        t0 = time.perf_counter()
        p, t = self.bank.config(design, self.N, self.g)
        self.g.set_params(p)
        self.g.set_tubes(t)
        self.g.generate()
        f = measure(self.g)
        self.seconds += time.perf_counter() - t0
        self.runs += 1
        return f

    def close(self):
        self.g.close()


def surrogate_design(sur: Surrogate, target, verifier: Verifier, tol: float = 0.05, max_runs: int = 8,
                     damping: float = 0.7, log=None) -> dict:
    """Predict a design from the target statistics, verify it, and while some shell misses the
    tolerance correct the design with the network as an approximate inverse:
    theta <- theta + S(target) - S(measured)."""
    t0 = time.perf_counter()
    s_target = sur.predict(target)[0]
    theta = s_target.copy()
    bslots = np.arange(NLEV) * NPAR + 1
    theta[bslots] = snap_density(theta[bslots])       # the density is chosen once, then kept
    errs = []
    for k in range(max_runs):
        f = verifier.run(Design.from_vector(theta, verifier.nlev))
        errs.append(shell_error(f, target))
        if log:
            log(f"[surrogate] run {k + 1}: max shell error {errs[-1] * 100:.2f}%")
        if errs[-1] <= tol:
            return {"found": True, "runs": k + 1, "seconds": time.perf_counter() - t0, "errors": errs,
                    "design": theta}
        step = damping * (s_target - sur.predict(f)[0])
        step[bslots] = 0.0
        theta = theta + step
    return {"found": False, "runs": max_runs, "seconds": time.perf_counter() - t0, "errors": errs, "design": theta}


SWEEP_GRID = {"gamma": np.round(np.arange(1.10, 1.5001, 0.02), 2),
              "density": np.array(DENSITIES),
              "hurst": np.round(np.arange(0.5, 0.9001, 0.1), 2)}


def sweep_design(target, verifier: Verifier, tol: float = 0.05, max_runs: int | None = None, log=None,
                 grid: dict | None = None) -> dict:
    """The plain parameter sweep: the level law of the plain method over a grid of (gamma,
    density, hurst) in lexicographic order, each candidate verified by one run, until every shell
    is within the tolerance or the grid (or max_runs) is exhausted."""
    grid = grid or SWEEP_GRID
    t0 = time.perf_counter()
    best = (np.inf, None)
    k = 0
    for gm in grid["gamma"]:
        for dn in grid["density"]:
            for hh in grid["hurst"]:
                if max_runs is not None and k >= max_runs:
                    return {"found": False, "runs": k, "seconds": time.perf_counter() - t0, "best": best,
                            "grid_size": int(np.prod([len(v) for v in grid.values()]))}
                f = verifier.run(family_design(float(gm), float(dn), float(hh), verifier.nlev))
                k += 1
                e = shell_error(f, target)
                if e < best[0]:
                    best = (e, (float(gm), float(dn), float(hh)))
                if e <= tol:
                    return {"found": True, "runs": k, "seconds": time.perf_counter() - t0, "best": best,
                            "grid_size": int(np.prod([len(v) for v in grid.values()]))}
    return {"found": False, "runs": k, "seconds": time.perf_counter() - t0, "best": best,
            "grid_size": int(np.prod([len(v) for v in grid.values()]))}


def family_jobs(n: int, seed: int, N: int, jitter_frac: float = 0.3, nlev: int = NLEV):
    """Inverse-design pairs: designs from the plain method's family (a fraction with per-level
    jitter), all on the bank's tube seed, at grid N."""
    rng = np.random.default_rng(seed)
    return [(draw_inverse_design(rng, rng.random() < jitter_frac, nlev), N, 0) for _ in range(n)]


def ensemble_design_jobs(n: int, seed: int, grids=GRIDS):
    """Inverse-design pairs from the ensemble workload's draws (`draw_design`: grid sizes from
    `grids`, 3 to 6 levels as the grid resolves them, Hurst index U[0.5, 0.9], density U[0.5, 2] and
    circulation exponent U[1.1, 1.5] per instance, each level jittered), each level's density
    snapped to the bank's discrete tube sets."""
    rng = np.random.default_rng(seed)
    lv = np.log2(DENSITIES)
    out = []
    for _ in range(n):
        N = int(rng.choice(grids))
        d = draw_design(rng, N)
        d.b = np.array([lv[np.argmin(np.abs(lv - b))] for b in d.b])
        out.append((d, N, 0))
    return out


def target_designs(n: int, seed: int = 9000, nlev: int = NLEV):
    """Hidden designs of the reachable targets: the plain method's family with gamma, density and
    hurst drawn inside the sweep's range."""
    rng = np.random.default_rng(seed)
    return [family_design(rng.uniform(1.12, 1.48), DENSITIES[int(rng.integers(len(DENSITIES)))],
                          rng.uniform(0.52, 0.88), nlev) for _ in range(n)]


def spectrum_features(E, flat) -> np.ndarray:
    """48 statistics of a target given as an integer-shell spectrum E[i-1], i = 1..nek, scaled to
    the normalised fields' energy (sum 3/2), and 16 log10 flatness values."""
    E = np.asarray(E, dtype=np.float64)
    E = E * (1.5 / E.sum())
    return np.concatenate([np.log10(np.maximum(log_shells(E), 1e-300)), np.asarray(flat, dtype=np.float64)])


def synthetic_targets(n: int, flat, seed: int = 9000, N: int = 512) -> list:
    """Seeded synthetic target spectra E(k) = C k^(-5/3) exp(-beta k eta), beta ~ U(1.5, 3),
    eta ~ U(1/64, 1/16) (target s uses default_rng(seed + s)), plus the reference's model spectrum
    (output/stat/spectrum_model_classical.eqn:3-21): E = 0.2 k^(-5/3) (k/4) / sqrt((k/4)^2 + 1)
    exp(-2.3 k / 32). Each as (name, 48 statistics) with the flatness `flat` (16 log10 values)."""
    k = np.arange(1, N // 2 + 1, dtype=np.float64)
    out = []
    for s in range(n):
        rng = np.random.default_rng(seed + s)
        beta, eta = rng.uniform(1.5, 3.0), rng.uniform(1.0 / 64, 1.0 / 16)
        out.append((f"k-5/3 beta={beta:.3f} eta={eta:.4f}", spectrum_features(k ** (-5.0 / 3.0) * np.exp(-beta * k * eta),
                                                                             flat)))
    kl = 0.25 * k
    model = 0.2 * (k + 1e-8) ** (-5.0 / 3.0) * (kl / np.sqrt(kl * kl + 1.0)) ** (5.0 / 3.0 - 2.0 / 3.0) * \
        np.exp(-2.3 * k * 0.03125)
    out.append(("reference model spectrum", spectrum_features(model, flat)))
    return out


def sweep_family(design: Design, grid: dict | None = None) -> dict:
    """Where a family design sits on the sweep grid: the index of its density and the nearest grid
    points in gamma and hurst (for reporting)."""
    grid = grid or SWEEP_GRID
    r = design.radii()
    g = float(design.a[1] / np.log2(r[1] / r[0])) if design.nlev > 1 else float("nan")
    return {"gamma": g, "density": float(2.0 ** design.b[0]), "hurst": float(design.H[0])}


# ---- training set of the shipped surrogate --------------------------------------------------------
PAIRS_FILE = "data/inverse_pairs.npz"
# half from the ensemble workload's draws over its grid sizes, half from the plain method's family
# at 128^3 (30% of those jittered per level), all on the bank's centerlines
PAIRS_RECIPE = {"n": 20480, "jobs_seed": 77, "bank_seed": 1234, "N": 128, "jitter_frac": 0.3,
                "ensemble_frac": 0.5}


def pairs_path() -> str:
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), PAIRS_FILE)


def load_pairs():
    z = np.load(pairs_path())
    return z["X"], z["Y"], z["M"]


def pair_jobs(rc: dict):
    ne = int(round(rc["n"] * rc["ensemble_frac"]))
    jobs = ensemble_design_jobs(ne, rc["jobs_seed"]) + family_jobs(rc["n"] - ne, rc["jobs_seed"] + 1, rc["N"],
                                                                   rc["jitter_frac"])
    # interleaved, so every ensemble mixes both kinds
    order = np.random.default_rng(rc["jobs_seed"] + 2).permutation(len(jobs))
    return [jobs[i] for i in order]


def make_pairs(out: str, log=print, lo: int = 0, hi: int | None = None, **recipe):
    """Regenerates the shipped training set on the GPU (PAIRS_RECIPE), or the pairs lo..hi of it
    (parts written separately are concatenated in order by `merge_pairs`)."""
    rc = dict(PAIRS_RECIPE, **recipe)
    bank = CenterlineBank(rc["bank_seed"])
    jobs = pair_jobs(rc)[lo:hi]
    X, Y, M, Ns = generate_pairs(jobs, ensemble=32, log=log, bank=bank)
    np.savez_compressed(out, X=X, Y=Y, M=M, N=Ns)
    return X, Y, M


def merge_pairs(parts, out: str):
    z = [np.load(p) for p in parts]
    np.savez_compressed(out, **{k: np.concatenate([a[k] for a in z]) for k in ("X", "Y", "M", "N")})


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="regenerate the inverse-design training pairs on the GPU")
    ap.add_argument("--out", default=pairs_path())
    ap.add_argument("--n", type=int, default=PAIRS_RECIPE["n"])
    ap.add_argument("--lo", type=int, default=0)
    ap.add_argument("--hi", type=int, default=None)
    ap.add_argument("--merge", nargs="*", help="concatenate these part files into --out instead")
    a = ap.parse_args()
    if a.merge:
        merge_pairs(a.merge, a.out)
    else:
        make_pairs(a.out, n=a.n, lo=a.lo, hi=a.hi)
