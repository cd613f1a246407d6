"""Build recipe for the oracle (test infrastructure only).

build(): compiles procell_oracle.c into oracle/_build/libprocell_oracle.so
with gcc.  -ffp-contract=off keeps each floating-point operation as written,
which is what makes the oracle bit-comparable with the GPU kernels.

build_ref(): where the reference's sources are present, compiles its argument
parser (src/io/cmdargs.cpp, plain C++ with no dependency beyond the standard
library) from where it lies, together with cmdargs_harness.cpp, into
oracle/_ref/cmdargs_ref.  Nothing of the reference is copied: g++ reads its
files in place.  The rest of the reference is CUDA (src/main.cu,
src/simulation/*.cu, src/io/parser.cu) and is not built here.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libprocell_oracle.so")
REF_ROOT = "/root/reference"
REF_DIR = os.path.join(HERE, "_ref")
REF_CMDARGS = os.path.join(REF_DIR, "cmdargs_ref")

CFLAGS = ["-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-mfma",
          "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    src = os.path.join(HERE, "procell_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + CFLAGS + ["-o", tmp, src, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_ref(verbose=False):
    """Build oracle/_ref/cmdargs_ref; returns its path, or None without the reference sources."""
    src = os.path.join(REF_ROOT, "src", "io", "cmdargs.cpp")
    if not os.path.isfile(src):
        return None
    os.makedirs(REF_DIR, exist_ok=True)
    tmp = REF_CMDARGS + ".tmp"
    cmd = ["g++", "-std=c++11", "-O1", "-I", os.path.join(REF_ROOT, "src"), src,
           os.path.join(HERE, "cmdargs_harness.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, REF_CMDARGS)
    return REF_CMDARGS


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_ref(verbose="-v" in sys.argv))
