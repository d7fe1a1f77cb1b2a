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
    """The check sections run as concurrent child processes (each emulated
    work-group is a host thread; the sections share nothing)."""
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, KPAX_LIB=emu_lib, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    groups = [("ops", "plans", "prune", "dump"), ("batch",), ("mlp",)]
    procs = [subprocess.Popen([sys.executable, os.path.join(EMU, "run_emu_checks.py"), *g], stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True, env=env) for g in groups]
    failed = []
    for g, p in zip(groups, procs):
        try:
            out, err = p.communicate(timeout=1800)
        except subprocess.TimeoutExpired:
            p.kill()
            out, err = p.communicate()
        print(out[-4000:])
        if p.returncode != 0:
            failed.append((g, out[-2000:], err[-4000:]))
    assert not failed, failed
