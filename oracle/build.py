"""Build recipe for the oracle (TEST INFRASTRUCTURE ONLY).

Compiles oracle/fr_oracle.c with gcc into oracle/liboracle.so.

The reference itself is not built: every reference translation unit includes
include/global.hpp, which includes "cblas.h" (global.hpp:18), and the image has
no CBLAS header or library (the reference makefile links -lopenblas,
makefile:11-12).  Building it would need a stand-in header and library, so
oracle/_ref stays empty and the oracle is pinned by the checks in tests/.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

CFLAGS = ["-O3", "-std=c99", "-fPIC", "-shared", "-fopenmp", "-mavx2", "-mfma", "-Wall", "-Wno-unused-function"]


def build(verbose=False):
    src = os.path.join(HERE, "fr_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + CFLAGS + [src, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("gcc failed building the oracle")
    if r.stderr.strip() and verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
