"""Build libprocell.so (gfx950) and the `procell` command in-tree.

Outputs go to cuda-pro-cell_amd/_build/.  Device code is compiled for gfx950
only, without floating-point contraction (the random stream's FP64 functions
must round exactly as written) and without packed-FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libprocell.so")
CLI = os.path.join(OUT, "procell")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC,
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE]

HIP_SOURCES = ["procell_kernels.hip", "procell_api.hip", "procell_relaxed.hip"]
CXX_SOURCES = ["procell_formats.cpp", "procell_setup.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    tmp_cli = CLI + ".tmp"
    _run(["g++"] + CXX_FLAGS + [os.path.join(CSRC, "procell_cli.cpp"), "-o", tmp_cli,
                                "-L", OUT, "-lprocell", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp_cli, CLI)
    return LIB


def clean() -> None:
    shutil.rmtree(OUT, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
