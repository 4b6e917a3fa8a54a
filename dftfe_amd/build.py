"""Builds libchfsi.so (gfx950) in-tree with hipcc, and the program that drives
the C++ drop-in classes of include/dftfe_chfsi.hpp over it (bin/dropin_solve,
used by the GPU tests).

    python -m dftfe_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchfsi.so")
SOURCES = ["kernels.hip", "kernels_bloch.hip", "operator.hip", "solver.hip", "bloch.hip",
           "dense.hip"]
ROOT = os.path.dirname(HERE)
DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_solve.cpp")
DROPIN = os.path.join(HERE, "bin", "dropin_solve")
# the same classes built with USE_COMPLEX (complex wavefunctions, k-points)
DROPIN_C_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_solve_complex.cpp")
DROPIN_C = os.path.join(HERE, "bin", "dropin_solve_complex")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    if os.path.exists(DROPIN_SRC):
        build_dropin(verbose)
    return LIB


def build_dropin(verbose: bool = False) -> str:
    """The C++ drop-in drivers (real, and with USE_COMPLEX), linked against the
    in-tree libchfsi.so."""
    os.makedirs(os.path.dirname(DROPIN), exist_ok=True)
    for src, out, defs in ((DROPIN_SRC, DROPIN, []),
                           (DROPIN_C_SRC, DROPIN_C, ["-DUSE_COMPLEX"])):
        if not os.path.exists(src):
            continue
        cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *defs,
               "-I", os.path.join(ROOT, "include"), src, "-o", out + ".tmp",
               "-L", HERE, "-lchfsi", "-Wl,-rpath,$ORIGIN/..",
               "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"drop-in driver build failed ({src}):\n{res.stderr}")
        os.replace(out + ".tmp", out)
        if verbose:
            print("built", out)
    return DROPIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
