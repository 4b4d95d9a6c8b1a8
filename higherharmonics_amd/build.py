"""Builds libhhg_amd.so (gfx950) in-tree with hipcc.

    python -m higherharmonics_amd.build

Every translation unit is compiled for gfx950 only, with packed-FP32
instructions disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``);
the FP64 path does not use them and the flag keeps hipcc from generating them
anywhere in the library.  ``-ffp-contract=on`` fuses a*b + c into an FMA only
within one source expression, as the C standard's FP_CONTRACT does, instead of
wherever the backend finds a multiply feeding an add: the kernel variants of
one solver (with and without step logging, say) then round identically.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhhg_amd.so")
RUN = os.path.join(HERE, "hhg_run")

SOURCES = ["hhg_solve_dirac.hip", "hhg_solve_piflux.hip", "hhg_solve_honeycomb.hip", "hhg_solve_dirac3d.hip",
           "hhg_solve_piflux3d.hip", "hhg_solve_fused.hip", "hhg_kernels.hip", "hhg_mlp.hip", "hhg_plan.hip", "hhg_host.cpp"]
HEADERS = ["hhg_device.hpp", "hhg_internal.hpp", "hhg_kernels.hpp", "hhg_host.hpp", "hhg_solve.hpp", "hhg_mlp.hpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=on",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [hipcc(), *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _git_commit() -> str:
    """HEAD of the repository the sources come from (recorded in the driver's metadata)."""
    try:
        r = subprocess.run(["git", "rev-parse", "HEAD"], cwd=HERE, capture_output=True, text=True, timeout=10)
        dirty = subprocess.run(["git", "status", "--porcelain", "--untracked-files=no"], cwd=HERE,
                               capture_output=True, text=True, timeout=10)
        if r.returncode == 0:
            return r.stdout.strip() + ("-dirty" if dirty.stdout.strip() else "")
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def build(verbose: bool = False) -> str:
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lrocprofiler-sdk-roctx"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the C++ driver over the host interface (cpp/hhg.hpp): a host-only program
    cxx = shutil.which("g++") or shutil.which("c++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", f"-DHHG_GIT_COMMIT=\"{_git_commit()}\"",
           os.path.join(HERE, "cpp", "hhg_run.cpp"), "-o", RUN + ".tmp",
           "-L" + HERE, "-lhhg_amd", "-lz", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(RUN + ".tmp", RUN)
    if verbose:
        print("built", LIB, "and", RUN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
