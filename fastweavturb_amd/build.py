"""Build libwovengpu.so (HIP, gfx950) in-tree with hipcc.

Exact-arithmetic sources (geometry, binning, deposition tests) are compiled with
-ffp-contract=off so that FP64 statements round like the reference's; every source is compiled
with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwovengpu.so")
ARCH = os.environ.get("WT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = {
    "segments.hip": ["-ffp-contract=off"],
    "deposit.hip": ["-ffp-contract=off"],
    "bricks.hip": ["-ffp-contract=off"],
    "scan.hip": [],
    "fft.hip": [],
    "stats.hip": [],
    "fbb.hip": [],
    "mlp.hip": [],
    "wt_api.cpp": ["-x", "hip"],
}
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fno-fast-math"]


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    noise = "'-packed-fp32-ops' is not a recognized feature for this target"
    msgs = [ln for ln in r.stderr.splitlines() if ln.strip() and noise not in ln]
    if msgs:
        sys.stderr.write("\n".join(msgs) + "\n")
    return obj


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")
HOST_EXE = os.path.join(FORTRAN, "weave_host")


def build_fortran_host(verbose: bool = False) -> str | None:
    """Compile the Fortran drop-in host (ISO_C_BINDING) against libwovengpu.so with flang."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found; Fortran host not built")
        return None
    cmd = [FLANG, "-O2", "-J", BUILD, "-o", HOST_EXE + ".tmp",
           os.path.join(FORTRAN, "wovengpu_mod.f90"), os.path.join(FORTRAN, "weave.f90"),
           f"-L{HERE}", "-lwovengpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed:\n{r.stdout}\n{r.stderr}")
    os.replace(HOST_EXE + ".tmp", HOST_EXE)
    if verbose:
        print("built", HOST_EXE)
    return HOST_EXE


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
