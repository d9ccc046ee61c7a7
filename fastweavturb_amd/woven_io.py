"""Readers and writers for the reference's input and output file formats.

Input formats (the drop-in boundary):
  centerline_input.dat  - nline; then nline point counts; then one "x y z" row per point
                          (read list-directed at tube.f90:168-177, 341-379; written by
                          FBB_with_parameters.m:53-61, 122-126).
  parameters_input.dat  - label/value pairs N, N_downsample, line_s, line_e, sigma_ratio,
                          eta_t, a separator, ntf_times, ntf_pre, Rtube_ratio, dzeta_sup, a
                          separator, a "variables=" line, then one row per tube
                          "i sigma Gamma n_T if_filter" (tube.f90:191-223; written by
                          FBB_with_parameters.m:203-245).
Output formats: spectrum.dat rows "i E(i)" (subroutine.f90:1636), Sp.dat rows of ES16.5
(subroutine.f90:114-123).
"""
from __future__ import annotations

import dataclasses
import io
import os

import numpy as np


@dataclasses.dataclass
class RunParams:
    """The scalar header of parameters_input.dat, in file order (tube.f90:193-216)."""
    N: int
    n_d: int
    line_s: int
    line_e: int
    sigma_ratio: float
    eta_t: float
    ntf_times: int = 50
    ntf_pre: int = 1000
    Rtube_ratio: float = 3.0
    dzeta_sup: float = 1e-4


@dataclasses.dataclass
class Tubes:
    """Per-tube parameters plus the flat centerline point arrays (SoA, FP64)."""
    npoint: np.ndarray      # int32 [nline]
    cx: np.ndarray          # float64 [npoint_all]
    cy: np.ndarray
    cz: np.ndarray
    sigma: np.ndarray       # float64 [nline]  sigma_min of each tube
    gamma: np.ndarray       # float64 [nline]  circulation
    n_T: np.ndarray         # int32 [nline]
    if_filter: np.ndarray   # int32 [nline]

    @property
    def nline(self) -> int:
        return int(len(self.npoint))

    @property
    def npoint_all(self) -> int:
        return int(self.npoint.sum())

    def validate(self) -> None:
        n = self.nline
        if n <= 0:
            raise ValueError("no tubes")
        for name in ("sigma", "gamma", "n_T", "if_filter"):
            if len(getattr(self, name)) != n:
                raise ValueError(f"{name} has {len(getattr(self, name))} entries, expected {n}")
        P = self.npoint_all
        for name in ("cx", "cy", "cz"):
            if len(getattr(self, name)) != P:
                raise ValueError(f"{name} has {len(getattr(self, name))} points, expected {P}")
        if np.any(self.npoint < 3):
            raise ValueError("every tube needs at least 3 points")
        for name in ("cx", "cy", "cz"):
            if not np.all(np.isfinite(getattr(self, name))):
                raise ValueError(f"non-finite coordinate in {name}")


def _tokens(path_or_text):
    if isinstance(path_or_text, (str, os.PathLike)) and os.path.exists(path_or_text):
        with open(path_or_text, "r") as f:
            text = f.read()
    else:
        text = str(path_or_text)
    return text


def read_centerlines(path) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Parse centerline_input.dat (tube.f90:168-177, 341-379)."""
    text = _tokens(path)
    fields = text.split()
    if not fields:
        raise ValueError("empty centerline file")
    try:
        nline = int(float(fields[0]))
    except ValueError as e:
        raise ValueError("centerline file: first entry must be the number of lines") from e
    if nline <= 0:
        raise ValueError("centerline file: nline must be positive")
    if len(fields) < 1 + nline:
        raise ValueError("centerline file truncated in the point counts")
    npoint = np.array([int(float(t)) for t in fields[1:1 + nline]], dtype=np.int32)
    P = int(npoint.sum())
    body = fields[1 + nline:1 + nline + 3 * P]
    if len(body) < 3 * P:
        raise ValueError(f"centerline file truncated: expected {P} points, found {len(body) // 3}")
    xyz = np.array(body, dtype=np.float64).reshape(P, 3)
    return npoint, xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()


def read_parameters(path, nline: int) -> tuple[RunParams, np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Parse parameters_input.dat (tube.f90:191-223). Rows are placed by their own index i."""
    lines = _tokens(path).splitlines()
    it = iter(lines)

    def value(kind):
        next(it)  # label line
        v = next(it).split()[0]
        return kind(float(v)) if kind is int else kind(v)

    try:
        N = value(int)
        n_d = value(int)
        line_s = value(int)
        line_e = value(int)
        sigma_ratio = value(float)
        eta_t = value(float)
        next(it)  # separator
        ntf_times = value(int)
        ntf_pre = value(int)
        Rtube_ratio = value(float)
        dzeta_sup = value(float)
        next(it)  # separator
        next(it)  # variables=
    except StopIteration as e:
        raise ValueError("parameters file truncated in its header") from e
    sigma = np.zeros(nline)
    gamma = np.zeros(nline)
    n_T = np.zeros(nline, dtype=np.int32)
    if_filter = np.zeros(nline, dtype=np.int32)
    seen = 0
    for raw in it:
        tok = raw.split()
        if not tok:
            continue
        if seen == nline:
            break
        if len(tok) < 5:
            raise ValueError(f"parameters file: short tube row {raw!r}")
        i = int(float(tok[0]))
        if not 1 <= i <= nline:
            raise ValueError(f"parameters file: tube index {i} outside 1..{nline}")
        sigma[i - 1] = float(tok[1])
        gamma[i - 1] = float(tok[2])
        n_T[i - 1] = int(float(tok[3]))
        if_filter[i - 1] = int(float(tok[4]))
        seen += 1
    if seen < nline:
        raise ValueError(f"parameters file truncated: {seen} of {nline} tube rows")
    p = RunParams(N, n_d, line_s, line_e, sigma_ratio, eta_t, ntf_times, ntf_pre, Rtube_ratio, dzeta_sup)
    if N < 4 or N % 2:
        raise ValueError("N must be even and at least 4")
    return p, sigma, gamma, n_T, if_filter


def read_inputs(directory) -> tuple[RunParams, Tubes]:
    npoint, cx, cy, cz = read_centerlines(os.path.join(directory, "centerline_input.dat"))
    p, sigma, gamma, n_T, if_filter = read_parameters(
        os.path.join(directory, "parameters_input.dat"), len(npoint))
    t = Tubes(npoint, cx, cy, cz, sigma, gamma, n_T, if_filter)
    t.validate()
    return p, t


def write_inputs(directory, p: RunParams, t: Tubes, reference_digits: bool = False) -> None:
    """Write both input files in the layout FBB_with_parameters.m produces. With reference_digits
    the numbers are printed as the script prints them (%18g: six significant digits; counts,
    coordinates, sigma and Gamma, FBB_with_parameters.m:53-61, 122-126, 232-233); otherwise
    coordinates, sigma and Gamma keep ten digits."""
    os.makedirs(directory, exist_ok=True)
    buf = io.StringIO()
    fmt = "%18g" if reference_digits else "%18.10g"
    buf.write((f"{t.nline:18g}" if reference_digits else f"{t.nline:18d}") + "\n")
    for n in t.npoint:
        buf.write((f"{int(n):18g}" if reference_digits else f"{int(n):18d}") + "\n")
    xyz = np.stack([t.cx, t.cy, t.cz], axis=1)
    np.savetxt(buf, xyz, fmt=fmt, delimiter="")
    with open(os.path.join(directory, "centerline_input.dat"), "w") as f:
        f.write(buf.getvalue())
    rows = [
        "N", f"{p.N:8d}", "N_downsample", f"{p.n_d:8d}", "line_s", f"{p.line_s:8d}",
        "line_e", f"{p.line_e:8d}", "sigma_ratio", f"{p.sigma_ratio:18.5f}", "eta_t",
        f"{p.eta_t:18.5f}", "---------------", "ntf_times", f"{p.ntf_times:8d}", "ntf_pre",
        f"{p.ntf_pre:8d}", "Rtube_ratio", f"{p.Rtube_ratio:18.5f}", "dzeta_sup",
        f"{p.dzeta_sup:18g}", "---------------",
        'variables="i_line","sigma","Gamma","n_T","if_filter"',
    ]
    sg = "18g" if reference_digits else "18.10g"
    for i in range(t.nline):
        rows.append(f"{i + 1:8d}{t.sigma[i]:{sg}}{t.gamma[i]:{sg}}{int(t.n_T[i]):8d}{int(t.if_filter[i]):8d}")
    with open(os.path.join(directory, "parameters_input.dat"), "w") as f:
        f.write("\n".join(rows) + "\n")


def write_spectrum(path, E) -> None:
    """spectrum.dat: one row "i E(i)" per shell i = 1..nek (subroutine.f90:1636)."""
    with open(path, "w") as f:
        for i, e in enumerate(E, start=1):
            f.write(f"{i:12d} {float(e):.8E}\n")


def write_sp(path, r_list, Sp, N) -> None:
    """Sp.dat: per separation r*dr then S_1..S_order, each ES16.5 (subroutine.f90:114-123)."""
    dr = 2.0 * np.pi / N
    with open(path, "w") as f:
        for ir, r in enumerate(r_list):
            vals = [r * dr] + [Sp[p][ir] for p in range(len(Sp))]
            f.write("".join(f"{v:16.5E}" for v in vals) + "\n")
