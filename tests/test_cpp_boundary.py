"""The C++ side of the drop-in boundary: INTEGRATION.md's rewrite of the
reference's examples/gpu/main.cu, compiled as printed there against
kino-pax-fork_amd/host/kpax.hpp and libkpax.so, and the command-line planner
host/plan_main.cpp (the counterpart of examples/gpu/executionTimeMain.cu).

CPU: both programs build, and a failure (no device, or no obstacle file)
reaches the caller as a kpax::Error with a message and an exit status; the
library never calls exit() (the reference's reader does, helper.cu:18-22).
GPU: the example plans the quadcopter forest from an obstacle CSV and finds a
verified path; plan_main plans the double integrator three times."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from workloads import workload

PKG = os.path.join(ROOT, "kino-pax-fork_amd")


def _example_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"`examples/gpu/main\.cu` becomes.*?```cpp\n(.*?)```", text, re.S)
    assert m, "INTEGRATION.md holds no main.cpp example"
    return m.group(1)


@pytest.fixture(scope="module")
def example(tmp_path_factory, kpax):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("cpp")
    src = d / "main.cpp"
    src.write_text(_example_source())
    exe = d / "main"
    r = subprocess.run(["g++", "-std=c++17", "-O1", str(src), "-I", os.path.join(ROOT, "include"), "-I",
                        os.path.join(PKG, "host"), "-L", PKG, "-lkpax", f"-Wl,-rpath,{PKG}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return d, exe


def test_example_reports_failures_as_errors(example):
    d, exe = example
    run = d / "no_csv"
    run.mkdir(exist_ok=True)
    r = subprocess.run([str(exe)], cwd=run, capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "kpax:" in r.stderr and "failed" in r.stderr


def test_plan_main_reports_failures_as_errors(kpax, tmp_path):
    exe = os.path.join(PKG, "plan_main")
    assert os.path.exists(exe), "build.py builds kino-pax-fork_amd/plan_main"
    r = subprocess.run([exe, "2", "1", str(tmp_path / "missing.csv")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "error:" in r.stderr, (r.returncode, r.stderr)


@pytest.mark.gpu
def test_example_plans_the_quadcopter_forest(example, kpax):
    d, exe = example
    run = d / "forest"
    run.mkdir(exist_ok=True)
    _, obs, _, _ = workload(2, seed=0)
    np.savetxt(run / "obstacles.csv", obs, delimiter=",", fmt="%.9g")
    r = subprocess.run([str(exe)], cwd=run, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert r.stdout.startswith("status 0,") and "verified 1" in r.stdout


@pytest.mark.gpu
def test_plan_main_plans_the_double_integrator(kpax):
    r = subprocess.run([os.path.join(PKG, "plan_main"), "1", "3"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    runs = [l for l in r.stdout.splitlines() if l.startswith("run ")]
    assert len(runs) == 3 and all("status 0" in l and "verified 1" in l for l in runs)
    assert "p50" in r.stdout
