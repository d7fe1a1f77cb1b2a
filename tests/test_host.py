"""Host side of libkpax on the CPU: the C ABI exports, built-in workloads,
the obstacle CSV reader (reference fixtures), environment generators and the
FP64 trajectory re-verifier."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from workloads import workload


def test_library_exports_every_declared_symbol(kpax):
    hdr = open(os.path.join(ROOT, "include", "kpax.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(kpax_\w+)\s*\(", hdr, re.M)))
    assert len(names) >= 20
    L = ctypes.CDLL(kpax.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_struct_layouts_match_header(kpax, tmp_path):
    """Every ctypes mirror of a struct in include/kpax.h has the C compiler's
    size and field offsets (a C program built against the header prints them)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    mirrors = {"kpax_config": kpax.Config, "kpax_result": kpax._Result, "kpax_stats": kpax._Stats,
               "kpax_launch": kpax._Launch, "kpax_query": kpax._Query}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "kpax.h"', 'int main(void) {']
    for cname, cls in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, cls in mirrors.items():
        assert ctypes.sizeof(cls) == got[(cname, "size")], cname
        for f in cls._fields_:
            assert getattr(cls, f[0]).offset == got[(cname, f[0])], (cname, f[0])


@pytest.mark.parametrize("k,model,S,C,N,R1,R2,G,dt", [
    (1, 1, 6, 3, 65536, 8, 4, 16, 0.01),
    (2, 3, 12, 4, 1 << 20, 16, 4, 64, 0.02),
    (3, 2, 6, 3, 1 << 21, 16, 4, 128, 0.01),
])
def test_builtin_workloads(kpax, k, model, S, C, N, R1, R2, G, dt):
    c = kpax.config_workload(k)
    assert c.model == model and (c.S, c.C) == (S, C) and c.max_tree == N
    assert list(c.r1_len) == [R1] * 3 and list(c.r2_len) == [R2] * 3 and list(c.grid_dims) == [G] * 3
    assert abs(c.dt - dt) < 1e-9
    assert (c.steps_min, c.steps_max) == (5, 20)
    assert (c.fanout, c.root_repeat, c.sparse_repeat, c.dense_repeat, c.sparse_valid_below) == (32, 5, 15, 1, 10)
    assert kpax.model_dims(model) == (S, C)
    if k == 2:
        assert list(c.ws_max) == [10, 10, 5] and abs(c.ctrl_max[0] - 19.62) < 1e-6
        assert [round(x, 6) for x in c.inertia] == [0.01, 0.01, 0.02] and c.mass == 1.0
        assert abs(c.ctrl_min[1] + 0.1) < 1e-7 and abs(c.ctrl_max[3] - 0.1) < 1e-7


def test_unknown_workload_rejected(kpax):
    with pytest.raises(kpax.KpaxError):
        kpax.config_workload(9)


@pytest.mark.parametrize("name,rows", [("quadTrees", 158), ("narrowPassage", 2), ("quadNarrowPassage", 2),
                                       ("office", 496), ("clutter", 17)])
def test_read_reference_obstacle_csv(kpax, name, rows):
    """Fixtures are the reference's own obstacle files (include/config/obstacles/*/obstacles.csv)."""
    o = kpax.read_obstacles_csv(os.path.join(GOLDEN, "obstacles", f"{name}.csv"))
    assert o.shape == (rows, 6)
    assert np.all(o[:, 3:] >= o[:, :3])
    raw = open(os.path.join(GOLDEN, "obstacles", f"{name}.csv")).read().replace(",", " ").split()
    assert np.array_equal(o.ravel(), np.array([float(x) for x in raw], np.float32))


def test_read_csv_errors(kpax, tmp_path):
    with pytest.raises(kpax.KpaxError):
        kpax.read_obstacles_csv(tmp_path / "missing.csv")
    bad = tmp_path / "bad.csv"
    bad.write_text("0,0,0,1,1,x\n")
    with pytest.raises(kpax.KpaxError):
        kpax.read_obstacles_csv(bad)
    short = tmp_path / "short.csv"
    short.write_text("0,0,0,1,1\n")
    with pytest.raises(kpax.KpaxError):
        kpax.read_obstacles_csv(short)
    empty = tmp_path / "empty.csv"
    empty.write_text("")
    assert kpax.read_obstacles_csv(empty).shape == (0, 6)


def _box_point_dist(b, p):
    lo, hi = b[:3], b[3:]
    e = np.maximum(0, np.maximum(lo - p, p - hi))
    return np.linalg.norm(e)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_environment_generators(kpax, k):
    cfg, obs, start, goal = workload(k, seed=4)
    cfg2, obs2, _, _ = workload(k, seed=4)
    _, obs3, _, _ = workload(k, seed=5)
    assert np.array_equal(obs, obs2) and not np.array_equal(obs, obs3)
    assert obs.shape == (dict({1: 20, 2: 400, 3: 2000})[k], 6)
    lo, hi = np.array(cfg.ws_min), np.array(cfg.ws_max)
    assert np.all(obs[:, :3] >= lo - 1e-6) and np.all(obs[:, 3:] <= hi + 1e-6)
    assert np.all(obs[:, 3:] >= obs[:, :3])
    for b in obs:
        assert _box_point_dist(b, start[:3]) > 0.01
        assert _box_point_dist(b, goal[:3]) > cfg.goal_radius
    if k == 1:
        e = obs[:, 3:] - obs[:, :3]
        assert e.max() <= 0.15 + 1e-6
    if k == 2:
        assert np.all(obs[:, 2] == 0) and np.all(obs[:, 5] == 5)
        fp = obs[:, 3:5] - obs[:, :2]
        assert fp.min() >= 0.1 - 1e-6 and fp.max() <= 0.3 + 1e-6


def test_maze_has_one_narrow_passage(kpax):
    cfg, obs, start, goal = workload(3, seed=0)
    xm = 0.5
    wall = obs[(obs[:, 0] <= xm) & (obs[:, 3] >= xm)]
    ys = np.linspace(0.0005, 0.9995, 1000)
    Y, Z = np.meshgrid(ys, ys, indexing="ij")
    P = np.stack([Y.ravel(), Z.ravel()], 1)
    blocked = np.zeros(len(P), bool)
    for b in wall:
        blocked |= (P[:, 0] > b[1]) & (P[:, 0] < b[4]) & (P[:, 1] > b[2]) & (P[:, 1] < b[5])
    free = P[~blocked]
    assert len(free) > 0
    assert np.ptp(free[:, 0]) <= 0.1 + 2e-3 and np.ptp(free[:, 1]) <= 0.1 + 2e-3
    assert np.ptp(free[:, 0]) >= 0.1 - 2e-3


def _one_edge_path(kpax, oracle, cfg, obs, rng):
    x0, u, steps = np.zeros((cfg.S, 1), np.float32), None, np.array([12], np.int32)
    x0[:3, 0] = [5.0, 5.0, 2.0]
    u = np.array([[10.5], [0.02], [-0.01], [0.005]], np.float32)
    x1, valid, _ = oracle.propagate(cfg, obs, x0, u, steps)
    assert valid.all()
    D = cfg.S + cfg.C + 1
    rows = np.zeros((2, D), np.float32)
    rows[0, :cfg.S] = x0[:, 0]
    rows[1, :cfg.S] = x1[:, 0]
    rows[1, cfg.S:cfg.S + cfg.C] = u[:, 0]
    rows[1, -1] = np.float32(cfg.dt) * 12
    return rows, np.array([0, 12], np.int32), x1[:3, 0]


def test_verify_path_accepts_and_rejects(kpax, oracle, rng):
    cfg = kpax.config_workload(2)
    obs = np.zeros((0, 6), np.float32)
    rows, steps, end = _one_edge_path(kpax, oracle, cfg, obs, rng)
    ok, dev = kpax.verify_path(cfg, obs, rows, steps, end)
    assert ok and dev < 1e-5
    # goal not reached
    assert not kpax.verify_path(cfg, obs, rows, steps, end + 1.0)[0]
    # stored state inconsistent with the control
    bad = rows.copy()
    bad[1, 0] += 0.01
    assert not kpax.verify_path(cfg, obs, bad, steps, bad[1, :3])[0]
    # an obstacle across the edge
    mid = 0.5 * (rows[0, :3] + rows[1, :3])
    wall = np.array([[mid[0] - 0.5, mid[1] - 0.5, mid[2] - 0.5, mid[0] + 0.5, mid[1] + 0.5, mid[2] + 0.5]], np.float32)
    assert not kpax.verify_path(cfg, wall, rows, steps, end)[0]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_fails_cleanly(kpax):
    with pytest.raises(kpax.KpaxError, match="NO_DEVICE"):
        kpax.KPAX(kpax.config_workload(1))


def test_invalid_config_rejected(kpax):
    c = kpax.config_workload(1)
    c.steps_min = 0
    with pytest.raises(kpax.KpaxError):
        kpax.KPAX(c)


def test_dubins_rows_use_the_reference_control_order(kpax, oracle):
    """Result rows hold the Dubins airplane's controls as (yawRate, pitchRate,
    a), the reference's x1[6..8] (statePropagatorSpatialHash.cu:223-225).  An
    oracle path written in that order passes the FP64 re-verifier; the same
    path with the controls in draw order (a, yawRate, pitchRate) does not."""
    from workloads import oracle_path_rows, workload
    assert kpax.row_control_order(kpax.MODEL_DUBINS_AIRPLANE) == [1, 2, 0]
    assert kpax.row_control_order(kpax.MODEL_QUADCOPTER) == [0, 1, 2, 3]
    cfg, obs, start, _ = workload(3, seed=0, max_tree=1 << 14, max_iter=30)
    goal = np.array([0.22, 0.5, 0.5], np.float32)
    o = oracle.plan(cfg, obs, start, goal, seed=3)
    assert o["status"] == kpax.OK
    rows = oracle_path_rows(o, cfg)
    j, nodes = o["goal_index"], []
    while j != -1:
        nodes.append(j)
        j = int(o["parents"][j])
    steps = o["steps"][nodes[::-1]]
    assert rows.shape[0] >= 3
    S = cfg.S
    assert np.array_equal(rows[1:, S:S + 3], o["controls"][[1, 2, 0]][:, nodes[::-1][1:]].T)
    ok, dev = kpax.verify_path(cfg, obs, rows, steps, goal)
    assert ok and dev < 1e-4, dev
    draw_order = rows.copy()
    draw_order[:, S:S + 3] = rows[:, S:S + 3][:, [2, 0, 1]]
    ok2, dev2 = kpax.verify_path(cfg, obs, draw_order, steps, goal)
    assert not ok2 and dev2 > 1e-3, dev2
