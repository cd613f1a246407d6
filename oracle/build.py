"""Build recipe for the oracle (test infrastructure only).

Compiles procell_oracle.c into oracle/_build/libprocell_oracle.so with gcc.
-ffp-contract=off keeps each floating-point operation as written, which is
what makes the oracle bit-comparable with the GPU kernels.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libprocell_oracle.so")

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


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
