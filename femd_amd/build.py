"""Build libfemd_hip.so (femd_amd/lib/) for gfx950 with hipcc, in-tree.

    python femd_amd/build.py            # compile everything
    python femd_amd/build.py --clean    # remove the build outputs first

(run as a script: importing the package itself needs the library this builds)

The device code is compiled with -ffp-contract=off (the exact kernels reproduce the reference's
two-rounding a*b + c) and without packed FP32 instructions.  The host setup (host.cpp) is plain
C++17 with OpenMP, compiled by g++.
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
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libfemd_hip.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = os.environ.get("FEMD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off", f"-I{CSRC}", f"-I{ROOT / 'include'}"]

SOURCES = [("host.cpp", "cxx"), ("sweep.cpp", "cxx"), ("sweep.hip", "hip"), ("kernels.hip", "hip"), ("capi.hip", "hip"), ("krylov.hip", "hip"), ("batch.hip", "hip"), ("spmv.hip", "hip"), ("mfma.hip", "hip"), ("cg_engine.hip", "hip"), ("setup_dev.hip", "hip"), ("timestep.hip", "hip"), ("serve.hip", "hip")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)

    def compile_one(item):
        src, kind = item
        obj = BUILD / (src + ".o")
        if kind == "hip":
            cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        else:
            cmd = ["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lgomp"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    if "--clean" in sys.argv:
        shutil.rmtree(BUILD, ignore_errors=True)
        if LIB.exists():
            LIB.unlink()
    print(build(verbose="-v" in sys.argv))
