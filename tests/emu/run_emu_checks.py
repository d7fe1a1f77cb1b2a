"""Runs the HIP kernels' CPU build (tests/emu) against the oracle.

Started by tests/test_emu_kernels.py in a child process with
KPAX_LIB=<the CPU build> and AddressSanitizer preloaded, so every
out-of-bounds access a kernel makes aborts with a report.  Exits 0 when all
outputs equal the oracle's."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "kino-pax-fork_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import kpax  # noqa: E402
import oracle  # noqa: E402
from workloads import random_states, workload  # noqa: E402


def check(cond, what):
    if not cond:
        print(f"FAIL: {what}", flush=True)
        sys.exit(1)
    print(f"ok: {what}", flush=True)


def small(cfg):
    # a coarse grid: the same kernels, far fewer emulated work-groups
    for d in range(3):
        cfg.grid_dims[d] = 16
    return cfg


def main():
    rng = np.random.default_rng(7)
    for k in (1, 2, 3):
        cfg, obs, start, goal = workload(k, seed=1, max_tree=1024)
        small(cfg)
        with kpax.KPAX(cfg) as p:
            p.set_obstacles(obs)
            x0, u, steps = random_states(cfg, 128, rng)
            x1, v, f = p.propagate_batch(x0, u, steps)
            o1, ov, of = oracle.propagate(cfg, obs, x0, u, steps)
            check(np.array_equal(x1.view(np.uint32), o1.view(np.uint32)) and np.array_equal(v, ov)
                  and np.array_equal(f, of), f"propagate workload {k}")
            start_, items = p.grid()
            check(np.array_equal(np.diff(start_), oracle.grid_counts(cfg, obs)), f"grid workload {k}")
            lo = np.array(cfg.ws_min, np.float32)[:, None]
            ext = (np.array(cfg.ws_max, np.float32) - np.array(cfg.ws_min, np.float32))[:, None]
            a = (lo + ext * rng.uniform(-0.1, 1.1, (3, 256))).astype(np.float32)
            b = (a + 0.05 * ext * rng.uniform(-1, 1, (3, 256))).astype(np.float32)
            bl, bh = np.minimum(a, b), np.maximum(a, b)
            check(np.array_equal(p.collide_batch(bl, bh), oracle.collide(obs, bl, bh)), f"collide workload {k}")
            r1, r2 = p.region_batch(a)
            o1, o2 = oracle.regions(cfg, a)
            check(np.array_equal(r1, o1) and np.array_equal(r2, o2), f"regions workload {k}")
            valid = rng.uniform(0, 1, 256) < 0.7
            got = p.region_histogram(a, valid)
            want = oracle.histogram(cfg, a, valid)
            check(all(np.array_equal(g, w) for g, w in zip(got, want[:3])), f"histogram workload {k}")
            vc, tot, cov = want[1], want[0], want[2]
            check(np.array_equal(p.scores(vc, tot, cov).view(np.uint32),
                                 oracle.scores(cfg, vc, tot, cov)[0].view(np.uint32)), f"scores workload {k}")
    with kpax.KPAX(small(kpax.config_workload(1))) as p:
        for n in (1, 1500, 5 * 1024 + 3):
            w = rng.integers(0, 4, n).astype(np.uint32)
            out, total = p.repeat_expand(w)
            check(total == w.sum() and np.array_equal(out, np.repeat(np.arange(n), w)), f"repeat_expand n={n}")
    # whole plans: budget exhausted, goal found, max_iter reached
    cases = ((1, 0, dict(max_tree=1500, max_iter=6), None),
             (2, 0, dict(max_tree=4096, max_iter=12), [1.6, 0.5, 1.0]),
             (3, 0, dict(max_tree=2048, max_iter=4), None))
    for k, seed, kw, near_goal in cases:
        cfg, obs, start, goal = workload(k, seed=seed, **kw)
        small(cfg)
        if near_goal is not None:
            goal = np.array(near_goal, np.float32)
        with kpax.KPAX(cfg) as p:
            r = p.plan(start, goal, obs, seed=seed + 77)
            tree = p.tree()
        o = oracle.plan(cfg, obs, start, goal, seed=seed + 77, threads=4)
        same = (r.status == o["status"] and r.iterations == o["iterations"] and r.tree_size == o["tree_size"]
                and np.array_equal(tree["parents"], o["parents"])
                and np.array_equal(tree["states"].view(np.uint32), o["states"].view(np.uint32)))
        check(same, f"plan workload {k}: status {r.status} iterations {r.iterations} tree {r.tree_size}")
        if near_goal is not None:
            check(r.status == kpax.OK and r.verified == 1 and r.goal_index == o["goal_index"],
                  f"plan workload {k} reaches the goal with a verified path")

if __name__ == "__main__":
    main()
