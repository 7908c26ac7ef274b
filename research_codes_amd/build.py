"""Build libsphk.so in-tree for gfx950.

Device code: hipcc --offload-arch=gfx950, C++17, -O3, packed-FP32 ops
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).  Host tree code:
g++ with OpenMP.  Objects are compiled in parallel and linked with hipcc.

Usage: python -m research_codes_amd.build   (or research_codes_amd.build.build()).
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
INCLUDE = ROOT / "include"
LIB = PKG / "libsphk.so"
OBJ = PKG / "_build"

ARCH = os.environ.get("SPHK_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}", "-Wall",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", f"-I{INCLUDE}", f"-I{CSRC}",
             "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-Wall"]

HIP_SOURCES = ["sphk_kernels.hip", "sphk_runtime.hip", "sphk_step.hip", "sphk_dtree.hip", "sphk_adjoint.hip"]
CXX_SOURCES = ["sphk_tree.cpp", "sphk_fdps.cpp"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip() and verbose:
        print(out)


def sources() -> list[Path]:
    return [CSRC / s for s in HIP_SOURCES + CXX_SOURCES if (CSRC / s).exists()]


def build(verbose: bool = False) -> Path:
    """Compile every source and link libsphk.so; always rebuilds."""
    OBJ.mkdir(exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        src = CSRC / s
        if not src.exists():
            continue
        o = OBJ / (s + ".o")
        objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(o)])
    for s in CXX_SOURCES:
        src = CSRC / s
        if not src.exists():
            continue
        o = OBJ / (s + ".o")
        objs.append(o)
        jobs.append([CXX, *CXX_FLAGS, "-c", str(src), "-o", str(o)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lgomp"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
