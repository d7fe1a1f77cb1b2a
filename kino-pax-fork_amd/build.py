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
SOURCES = [os.path.join(HERE, "csrc", "kpax_device.hip"), os.path.join(HERE, "csrc", "kpax_host.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KPAX_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def build(verbose=False):
    objs = []
    tmp = LIB + ".tmp"
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    dev_obj = os.path.join(HERE, "_build", "kpax_device.o")
    host_obj = os.path.join(HERE, "_build", "kpax_host.o")
    cmds = [
        [HIPCC, f"--offload-arch={ARCH}", *COMMON,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-c", SOURCES[0], "-o", dev_obj],
        ["g++", *COMMON, "-c", SOURCES[1], "-o", host_obj],
        [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", dev_obj, host_obj, "-o", tmp],
    ]
    objs += [dev_obj, host_obj]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"build step failed: {' '.join(cmd)}")
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if verbose and err:
            sys.stderr.write(err + "\n")
    os.replace(tmp, LIB)
    # command-line planner over the C++ mirror (host/kpax.hpp)
    exe = os.path.join(HERE, "plan_main")
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(HERE, "host", "plan_main.cpp"), "-o", exe + ".tmp",
           "-L" + HERE, "-lkpax", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("plan_main build failed")
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
