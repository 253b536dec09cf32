"""Builds libhssc.so (the C-ABI library of include/hssc.h) for gfx950 with hipcc.

Usage: python -m hssc_directsolver_amd.build   (or build() from Python).
The library is written in-tree next to this file so that it travels with the repository
snapshot; nothing is installed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libhssc.so"
SOURCES = ["tree.cpp", "dr_weights.cpp", "stats.cpp", "pool.cpp", "kernels.hip", "id.hip", "solver.hip", "gridop.hip", "hss1d.hip", "mlp.hip"]
ARCH = os.environ.get("HSSC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libhssc.so")


def build(verbose: bool = False) -> Path:
    """Compiles every source and links libhssc.so (always rebuilds)."""
    cc = hipcc()
    objdir = PKG / "_objs"
    objdir.mkdir(exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
              "-I", str(ROOT / "include")]
    # packed FP32 ops are disabled for all device code (a correctness precaution on gfx950)
    dev = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    objs, cmds = [], []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        cmd = [cc, *common, *dev, "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            # host-only translation units (still compiled by hipcc for the HIP headers)
            cmd = [cc, *common, "-c", str(CSRC / src), "-o", str(obj)]
        cmds.append(cmd)
        objs.append(str(obj))
    # translation units compile independently: one worker each, bounded by the host's cores
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, cmd, verbose) for cmd in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l
                     and "not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{errs}")
    if verbose and errs.strip():
        print(errs, file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose=True))
