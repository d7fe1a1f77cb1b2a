"""Build libkpax.so (HIP kernels for gfx950 + host C++) in place.

    python kino-pax-fork_amd/build.py

Flags that matter for results:
  -ffp-contract=off   no a*b+c fusion the source did not ask for, so every
                      child state is reproducible bit for bit on the host
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keeps the compiler from emitting packed FP32 VALU ops
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libkpax.so")
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = ["kpax_planner.hip", "kpax_ops.hip", "kpax_mlp.hip"]
HOST_SOURCES = ["kpax_host.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KPAX_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def _wait(procs, verbose):
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError(f"build step failed: {' '.join(cmd)}")
        err = "\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if verbose and err:
            sys.stderr.write(err + "\n")


def build(verbose=False):
    tmp = LIB + ".tmp"
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs, procs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, _run(cmd, verbose)))
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(bdir, src.replace(".cpp", ".o"))
        cmd = ["g++", *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, _run(cmd, verbose)))
        objs.append(obj)
    _wait(procs, verbose)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    _wait([(cmd, _run(cmd, verbose))], verbose)
    os.replace(tmp, LIB)
    # command-line planner over the C++ mirror (host/kpax.hpp)
    exe = os.path.join(HERE, "plan_main")
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(HERE, "host", "plan_main.cpp"), "-o", exe + ".tmp",
           "-L" + HERE, "-lkpax", "-Wl,-rpath,$ORIGIN"]
    _wait([(cmd, _run(cmd, verbose))], verbose)
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
