"""Build the CPU oracle (test infrastructure only) in place.

    python oracle/build.py

The oracle is compiled without floating-point contraction so that its float
path evaluates every expression in the written order.

The reference itself (kranzowe/Kino-PAX-Fork) is CUDA C++ that needs nvcc,
cuRAND, Thrust and CUB; none is present here, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_kpax.so")
SRC = os.path.join(HERE, "kpax_oracle.cpp")


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-mfma", "-fopenmp", "-fPIC", "-shared",
           "-Wall", SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
