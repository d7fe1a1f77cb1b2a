"""The HIP kernels compiled for the host (tests/emu/hip/hip_runtime.h) and run
under AddressSanitizer against the oracle: every kernel index is checked on
the CPU before the GPU ever sees the code."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

EMU = os.path.join(ROOT, "tests", "emu")
SRC = os.path.join(ROOT, "kino-pax-fork_amd", "csrc")


@pytest.fixture(scope="module")
def emu_lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    out = str(tmp_path_factory.mktemp("emu") / "libkpax_emu.so")
    cmd = ["g++", "-std=c++20", "-x", "c++", "-I", EMU, "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-ffp-contract=off", "-fPIC", "-shared", "-pthread", os.path.join(SRC, "kpax_planner.hip"),
           os.path.join(SRC, "kpax_ops.hip"), os.path.join(SRC, "kpax_mlp.hip"), "-x", "none",
           os.path.join(SRC, "kpax_host.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def test_kernels_under_asan_match_oracle(emu_lib):
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, KPAX_LIB=emu_lib, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_emu_checks.py")], capture_output=True, text=True,
                       env=env, timeout=1800)
    print(r.stdout[-4000:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
