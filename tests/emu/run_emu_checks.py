"""Runs the HIP kernels' CPU build (tests/emu) against the oracle.

Started by tests/test_emu_kernels.py in a child process with
KPAX_LIB=<the CPU build> and AddressSanitizer preloaded, so every
out-of-bounds access a kernel makes aborts with a report.  Exits 0 when all
outputs equal the oracle's."""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "kino-pax-fork_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import kpax  # noqa: E402
import oracle  # noqa: E402
from workloads import oracle_path_rows, random_states, workload  # noqa: E402


T0 = time.perf_counter()


def check(cond, what):
    if not cond:
        print(f"FAIL: {what}", flush=True)
        sys.exit(1)
    print(f"ok: {what} ({time.perf_counter() - T0:.1f} s)", flush=True)


def small(cfg):
    # a coarse grid: the same kernels, far fewer emulated work-groups
    for d in range(3):
        cfg.grid_dims[d] = 16
    return cfg


def ops_checks():
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


def plan_checks():
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
    # the double integrator and the Dubins airplane (in a generated forest:
    # its maze has no LDS image) on both grid paths
    for model in (1, 3):
        cfg, obs, start, goal = workload(model, seed=0, max_tree=2000, max_iter=6 if model == 1 else 4)
        small(cfg)
        if model == 3:
            obs = kpax.generate_environment(2, 5, 40, cfg.ws_min, cfg.ws_max, start[:3],
                                            list(goal[:3]) + [cfg.goal_radius])
        o = oracle.plan(cfg, obs, start, goal, seed=41, threads=4)
        with kpax.KPAX(cfg) as p:
            p.set_obstacles(obs)
            has_image = p.lds_grid()[1] > 0
            for lds in (1, 0):
                p.set_launch(lds_grid=lds)
                r = p.plan(start, goal, seed=41)
                t = p.tree()
                check(has_image and r.tree_size == o["tree_size"] and np.array_equal(t["parents"], o["parents"])
                      and np.array_equal(t["states"].view(np.uint32), o["states"].view(np.uint32)),
                      f"model {model} with lds_grid={lds}: tree {r.tree_size} vs oracle {o['tree_size']}")
    # profiled phases count the launches that ran: a single-query handle
    # computes its scores inside the frontier pass and runs no proposal pass
    cfg, obs, start, goal = workload(1, seed=0, max_tree=1500, max_iter=6)
    small(cfg)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        p.set_profiling(True)
        p.reset_stats()
        r = p.plan(start, goal, seed=77)
        n = p.stats()["phase_launches"]
    check(n["score"] == 0 and n["propose"] == 0 and n["init"] == 1
          and n["prepare"] == n["expand"] == n["append"] == n["frontier"] == r.iterations,
          f"profiled phases of a {r.iterations}-iteration plan: {n}")


def prune_checks():
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1500, max_iter=8)
    small(cfg)
    cfg.prune = 1
    with kpax.KPAX(cfg) as p:
        r = p.plan(start, goal, obs, seed=21)
        tree = p.tree()
    o = oracle.plan(cfg, obs, start, goal, seed=21, threads=4)
    check(r.tree_size == o["tree_size"] and np.array_equal(tree["parents"], o["parents"])
          and np.array_equal(tree["states"].view(np.uint32), o["states"].view(np.uint32)),
          f"pruned plan: tree {r.tree_size} vs oracle {o['tree_size']}")


def dump_checks():
    """planBench-compatible dump: one samples/parents file per iteration, the
    last one equal to the returned tree."""
    import tempfile
    cfg, obs, start, goal = workload(2, seed=0, max_tree=4096, max_iter=12)
    small(cfg)
    goal = np.array([1.6, 0.5, 1.0], np.float32)
    with tempfile.TemporaryDirectory() as d, kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        r = p.plan_dump(start, goal, d, seed=77, bench_itr=3)
        tree = p.tree()
        last = np.loadtxt(os.path.join(d, "Samples", "Samples3", f"samples{r.iterations}.csv"), delimiter=",", ndmin=2)
        par = np.loadtxt(os.path.join(d, "Parents", "Parents3", f"parents{r.iterations}.csv"), ndmin=1)
        ts = np.loadtxt(os.path.join(d, "TreeSize", "TreeSize3", "treeSize.csv"), ndmin=1)
        n_files = len(os.listdir(os.path.join(d, "Samples", "Samples3")))
        ok = (last.shape == (r.tree_size, cfg.S + cfg.C + 1) and np.array_equal(par, tree["parents"])
              and np.allclose(last[:, :3], tree["states"][:3].T, atol=1e-6) and ts[-1] == r.tree_size
              and n_files == r.iterations and len(ts) == r.iterations)
        if r.status == kpax.OK:
            path = np.loadtxt(os.path.join(d, "ControlPathToGoal", "ControlPathToGoal3", "controlPathToGoal.csv"),
                              delimiter=",", ndmin=2)
            ok = ok and path.shape[0] == r.rows.shape[0] and np.allclose(path[::-1], r.rows, atol=1e-6)
    check(ok and r.status == kpax.OK, f"planBench dump: {n_files} iterations, tree {r.tree_size}, status {r.status}")


def mlp_checks():
    """The learned proposal's kernels: features and sampler bit-exact against
    the oracle, network forward and gradients within bf16 tolerance, and a
    planner run with the proposal mixed in."""
    import mlp_oracle
    rng = np.random.default_rng(3)
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1500, max_iter=6)
    small(cfg)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        x0, _, _ = random_states(cfg, 100, rng)
        x0[:, :5] = np.array([[-1, 5, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0]]).T  # outside the workspace
        g = np.tile(np.array([[9.5], [9.5], [1.0]], np.float32), (1, 100))
        f = p.mlp_features(0, x0, g)
        check(np.array_equal(f.view(np.uint32), oracle.mlp_features(cfg, obs, x0, g).view(np.uint32)),
              "proposal features bit-exact")
        ms = np.concatenate([rng.uniform(-1.5, 1.5, (300, 4)), rng.uniform(0.01, 1.0, (300, 4))], 1).astype(np.float32)
        w = rng.integers(0, 2**32, (300, 4), dtype=np.uint64).astype(np.uint32)
        w[0] = [0, 0, 0, 0]
        w[1] = [0xFFFFFFFF] * 4
        u = p.propose_sample(ms, w)
        check(np.array_equal(u.view(np.uint32), oracle.propose_sample(cfg, ms, w).view(np.uint32)),
              "proposal sampler bit-exact")
        params = kpax.mlp_init_params(5)
        out = p.mlp_infer(params, f)
        ref, _ = mlp_oracle.forward(params, f)
        err = np.abs(out - ref).max() / max(1e-6, np.abs(ref).max())
        check(err < 1e-2, f"network forward within bf16 tolerance (rel err {err:.2e})")
        p.set_proposal(params, 0.5)
        r1 = p.plan(start, goal, seed=9)
        r2 = p.plan(start, goal, seed=9)
        tree, table = p.tree(), p.proposal_table()
        used = np.unique(tree["parents"][1:])  # nodes the plan expanded: their Gaussians were read
        gg = np.tile(np.asarray(goal[:3], np.float32).reshape(3, 1), (1, used.size))
        raw = p.mlp_infer(params, p.mlp_features(0, tree["states"][:, used], gg)).T
        check(np.array_equal(table[:4, used].view(np.uint32), raw[:4].view(np.uint32))
              and np.allclose(table[4:, used], np.exp(np.clip(raw[4:], -4.0, 0.5)), rtol=1e-6, atol=0),
              f"proposal pass equals the network on the features of {used.size} expanded nodes")
        p.set_proposal(None)
        r0 = p.plan(start, goal, seed=9)
    o = oracle.plan(cfg, obs, start, goal, seed=9, threads=4)
    check(r1.tree_size == r2.tree_size and r1.expansions == r2.expansions, "proposal planner is deterministic")
    check(r0.tree_size == o["tree_size"] and r0.expansions == o["expansions"],
          "proposal off gives the plain planner again")
    check(r1.tree_size != r0.tree_size or r1.expansions != r0.expansions, "proposal changes the tree")
    X = f[:100]
    Y = rng.uniform(-1, 1, (100, 4)).astype(np.float32)
    with kpax.Trainer(max_batch=256) as t:
        t.set_params(params)
        loss, grad = t.gradients(X, Y)
        rl, rg = mlp_oracle.loss_and_grads(params, X, Y)
        gerr = np.linalg.norm(grad - rg) / np.linalg.norm(rg)
        check(abs(loss - rl) < 1e-4 * (1 + abs(rl)) and gerr < 2e-2,
              f"gradients match the oracle (loss {loss:.5f} vs {rl:.5f}, rel err {gerr:.2e})")
        losses = t.fit(np.tile(X, (3, 1)), np.tile(Y, (3, 1)), epochs=3, batch=64, lr=1e-3, seed=1)
        check(np.all(np.isfinite(losses)), f"fit runs: losses {losses}")


def batch_checks():
    """Uneven budgets over two environments on three slots: queries wait for
    node ranges, reuse the ranges and slots of finished ones, and each result
    equals the oracle's plan of that query alone."""
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1500, max_iter=10)
    small(cfg)
    obs1 = kpax.generate_environment(2, 99, 60, cfg.ws_min, cfg.ws_max, start[:3], list(goal[:3]) + [cfg.goal_radius])
    envs = [obs, obs1]
    near = np.array([1.6, 0.5, 1.0], np.float32)
    qs = [kpax.Query(start, near, seed=11, budget=1024, env=0),
          kpax.Query(start, goal, seed=12, budget=700, env=1),
          kpax.Query(start, near, seed=13, budget=1500, env=1),
          kpax.Query(start, goal, seed=14, budget=1, env=0),
          kpax.Query(start, near, seed=15, budget=900, env=0)]
    with kpax.KPAX(cfg, max_queries=3, node_capacity=3000) as p:
        p.set_environment(0, obs)
        p.set_environment(1, obs1)
        res = p.plan_batch(qs)
        p.set_launch(expand_blocks=3, expand_chunk=192, scan_blocks=2, order_children=1, refill_lanes=1,
                     scan_tile=1024)
        res2 = p.plan_batch(qs)
        p.set_launch(refill_lanes=64)
        res3 = p.plan_batch(qs)
    for other, what in ((res2, "another launch shape"), (res3, "servicing a wave only when all its lanes idle")):
        check(all(a.status == b.status and a.tree_size == b.tree_size and a.expansions == b.expansions
                  and a.integrated_steps == b.integrated_steps and np.array_equal(a.rows, b.rows)
                  for a, b in zip(res, other)), f"{what} gives the same batch results")
    for i, (q, r) in enumerate(zip(qs, res)):
        c = cfg.copy()
        c.max_tree = q.budget
        o = oracle.plan(c, envs[q.env], q.start, q.goal, seed=q.seed, threads=4)
        same = (r.status == o["status"] and r.iterations == o["iterations"] and r.tree_size == o["tree_size"]
                and r.goal_index == o["goal_index"] and r.expansions == o["expansions"])
        if same and r.status == kpax.OK:
            same = np.array_equal(r.rows.view(np.uint32), oracle_path_rows(o, c).view(np.uint32)) and r.verified == 1
        check(same, f"batch query {i} (budget {q.budget}, env {q.env}): status {r.status} iterations "
                    f"{r.iterations} tree {r.tree_size} vs oracle {o['status']} {o['iterations']} {o['tree_size']}")
    check(any(r.status == kpax.OK for r in res) and any(r.status != kpax.OK for r in res),
          "batch holds both found and unfound queries")
    # the batched kernels of the other two models
    for model in (1, 3):
        cfg, obs, start, goal = workload(model, seed=0, max_tree=3000, max_iter=8 if model == 1 else 5)
        small(cfg)
        mid = ((np.asarray(start[:3], np.float32) + np.asarray(goal[:3], np.float32)) / 2).astype(np.float32)
        mq = [kpax.Query(start, mid, seed=31, budget=2000, env=0), kpax.Query(start, goal, seed=32, budget=900, env=0),
              kpax.Query(start, goal, seed=33, budget=1, env=0), kpax.Query(start, mid, seed=34, budget=1500, env=0)]
        with kpax.KPAX(cfg, max_queries=2, node_capacity=5000) as p:
            p.set_environment(0, obs)
            mres = p.plan_batch(mq)
        for i, (q, r) in enumerate(zip(mq, mres)):
            c = cfg.copy()
            c.max_tree = q.budget
            o = oracle.plan(c, obs, q.start, q.goal, seed=q.seed, threads=4)
            same = (r.status, r.iterations, r.tree_size, r.expansions) == (o["status"], o["iterations"],
                                                                           o["tree_size"], o["expansions"])
            if same and r.status == kpax.OK:
                same = np.array_equal(r.rows.view(np.uint32), oracle_path_rows(o, c).view(np.uint32))
            check(same, f"model {model} batch query {i} (budget {q.budget}): status {r.status} tree {r.tree_size} "
                        f"vs oracle {o['status']} {o['tree_size']}")


SECTIONS = {"ops": ops_checks, "plans": plan_checks, "batch": batch_checks, "mlp": mlp_checks,
            "prune": prune_checks, "dump": dump_checks}


if __name__ == "__main__":
    # sections named on the command line, or all of them
    for name in sys.argv[1:] or list(SECTIONS):
        SECTIONS[name]()
