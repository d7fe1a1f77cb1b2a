"""GPU: batched queries on one handle (kpax_plan_batch).

Each query of a batch must return exactly what a plan of that query alone
returns (same seed, environment and tree cap): the status, iteration count,
tree size, goal node, expansion and step counters, and the path rows bit for
bit.  Checked against the CPU oracle for small budgets and against single
GPU plans for larger ones.
"""
import os

import numpy as np
import pytest

from workloads import oracle_path_rows, workload

pytestmark = pytest.mark.gpu


def _envs(kpax, cfg, start, goal, n_env):
    return [kpax.generate_environment(2, 100 + e, 100 + 150 * e, cfg.ws_min, cfg.ws_max, start[:3],
                                      list(goal[:3]) + [cfg.goal_radius]) for e in range(n_env)]


def _same(r, o):
    return (r.status == o.status and r.iterations == o.iterations and r.tree_size == o.tree_size
            and r.goal_index == o.goal_index and r.expansions == o.expansions
            and r.integrated_steps == o.integrated_steps and r.rows.shape == o.rows.shape
            and np.array_equal(r.rows.view(np.uint32), o.rows.view(np.uint32)) and np.array_equal(r.steps, o.steps))


def test_batch_matches_oracle_per_query(kpax, oracle):
    cfg, obs, start, goal = workload(2, seed=0, max_iter=40)
    envs = [obs] + _envs(kpax, cfg, start, goal, 2)
    near = np.array([3.0, 3.0, 1.0], np.float32)
    qs = [kpax.Query(start, near, seed=21, budget=1 << 15, env=0),
          kpax.Query(start, goal, seed=22, budget=20000, env=1),
          kpax.Query(start, near, seed=23, budget=1 << 16, env=2),
          kpax.Query(start, goal, seed=24, budget=1, env=0),      # the root alone fills the budget
          kpax.Query(start, near, seed=25, budget=50000, env=1),
          kpax.Query(start, goal, seed=26, budget=1 << 14, env=2)]
    print("batch vs oracle: 6 queries on 3 slots, 3 environments", flush=True)
    with kpax.KPAX(cfg, max_queries=3, node_capacity=100000) as p:
        for e, o in enumerate(envs):
            p.set_environment(e, o)
        res = p.plan_batch(qs)
    statuses = set()
    for i, (q, r) in enumerate(zip(qs, res)):
        c = cfg.copy()
        c.max_tree = q.budget
        o = oracle.plan(c, envs[q.env], q.start, q.goal, seed=q.seed, threads=8)
        print(f"  query {i}: gpu {r.status} it {r.iterations} ts {r.tree_size}; oracle {o['status']} "
              f"it {o['iterations']} ts {o['tree_size']}", flush=True)
        assert (r.status, r.iterations, r.tree_size, r.expansions) == \
            (o["status"], o["iterations"], o["tree_size"], o["expansions"])
        assert r.goal_index == o["goal_index"]
        if r.status == kpax.OK:
            assert np.array_equal(r.rows.view(np.uint32), oracle_path_rows(o, c).view(np.uint32))
            assert r.verified == 1
        statuses.add(r.status)
    assert kpax.OK in statuses and kpax.BUDGET_EXHAUSTED in statuses


@pytest.mark.parametrize("model", [1, 3], ids=["double_integrator", "dubins_airplane"])
def test_batch_of_the_other_models_matches_oracle_per_query(kpax, oracle, model):
    """The batched expansion, preparation and scans compiled for the 6D double
    integrator and the 6D Dubins airplane (the quadcopter's are covered
    above): each query of a batch with uneven budgets over two slots equals
    the oracle's plan of that query alone, counters and path rows bit for
    bit."""
    cfg, obs, start, goal = workload(model, seed=0, max_iter=25 if model == 1 else 12)
    mid = ((np.asarray(start[:3], np.float32) + np.asarray(goal[:3], np.float32)) / 2).astype(np.float32)
    qs = [kpax.Query(start, mid, seed=31, budget=1 << 14, env=0),
          kpax.Query(start, goal, seed=32, budget=6000, env=0),
          kpax.Query(start, goal, seed=33, budget=1, env=0),
          kpax.Query(start, mid, seed=34, budget=20000, env=0),
          kpax.Query(start, goal, seed=35, budget=1 << 13, env=0)]
    print(f"model {model}: 5 queries on 2 slots vs the oracle", flush=True)
    with kpax.KPAX(cfg, max_queries=2, node_capacity=40000) as p:
        p.set_environment(0, obs)
        res = p.plan_batch(qs)
    for i, (q, r) in enumerate(zip(qs, res)):
        c = cfg.copy()
        c.max_tree = q.budget
        o = oracle.plan(c, obs, q.start, q.goal, seed=q.seed, threads=8)
        print(f"  query {i}: gpu {r.status} it {r.iterations} ts {r.tree_size}; oracle {o['status']} "
              f"it {o['iterations']} ts {o['tree_size']}", flush=True)
        assert (r.status, r.iterations, r.tree_size, r.expansions) == \
            (o["status"], o["iterations"], o["tree_size"], o["expansions"])
        assert r.goal_index == o["goal_index"]
        if r.status == kpax.OK:
            assert np.array_equal(r.rows.view(np.uint32), oracle_path_rows(o, c).view(np.uint32))
            assert r.verified == 1
    assert res[2].status == kpax.BUDGET_EXHAUSTED


def test_batch_matches_single_plans(kpax):
    cfg, obs, start, goal = workload(2, seed=0, max_iter=60)
    envs = [obs] + _envs(kpax, cfg, start, goal, 3)
    rng = np.random.default_rng(5)
    budgets = np.exp(rng.uniform(np.log(16384), np.log(1 << 18), 20)).astype(np.uint32)
    qs = []
    for i, b in enumerate(budgets):
        g = goal if i % 3 else np.array([5.0, 5.0, 1.0], np.float32)
        qs.append(kpax.Query(start, g, seed=1000 + i, budget=int(b), env=i % 4))
    print(f"batch vs single: 20 queries, budgets {budgets.min()}..{budgets.max()}", flush=True)
    with kpax.KPAX(cfg, max_queries=8, node_capacity=1 << 20) as p:
        for e, o in enumerate(envs):
            p.set_environment(e, o)
        res = p.plan_batch(qs)
    for i, (q, r) in enumerate(zip(qs, res)):
        c = cfg.copy()
        c.max_tree = q.budget
        with kpax.KPAX(c) as s:
            s.set_obstacles(envs[q.env])
            o = s.plan(q.start, q.goal, seed=q.seed)
        assert _same(r, o), (i, r.status, r.iterations, r.tree_size, o.status, o.iterations, o.tree_size)
    assert sum(r.status == kpax.OK for r in res) >= 5


def test_batch_graph_and_direct_launches_agree(kpax):
    cfg, obs, start, goal = workload(2, seed=0, max_iter=30)
    qs = [kpax.Query(start, goal if i % 2 else np.array([4.0, 4.0, 1.0], np.float32), seed=50 + i,
                     budget=30000 + 7000 * i) for i in range(10)]
    with kpax.KPAX(cfg, max_queries=4, node_capacity=200000) as p:
        p.set_obstacles(obs)
        p.set_launch(use_graph=1)
        graphed = p.plan_batch(qs)
        n_graph = p.stats()["graph_launches"]
        p.set_launch(use_graph=0)
        direct = p.plan_batch(qs)
        p.set_launch(use_graph=1)
        p.set_profiling(True)  # profiling launches the kernels one by one
        profiled = p.plan_batch(qs)
        p.set_profiling(False)
        again = p.plan_batch(qs)  # the handle's state carries over between calls
    assert n_graph > 0, "the iteration was not captured as a HIP graph"
    for a, b, c, d in zip(graphed, direct, profiled, again):
        assert _same(a, b) and _same(a, c) and _same(a, d)


def test_batch_rejects_impossible_queries_and_runs_the_rest(kpax):
    cfg, obs, start, goal = workload(2, seed=0, max_iter=20)
    bad_start = np.array(start, np.float32)
    bad_start[0] = np.nan
    qs = [kpax.Query(start, goal, seed=1, budget=20000),
          kpax.Query(start, goal, seed=2, budget=300000),      # above the node capacity
          kpax.Query(start, goal, seed=3, budget=20000, env=7),  # environment never set
          kpax.Query(bad_start, goal, seed=4, budget=20000),
          kpax.Query(start, goal, seed=5, budget=1)]            # the root alone fills the budget
    with kpax.KPAX(cfg, max_queries=2, node_capacity=100000) as p:
        p.set_obstacles(obs)
        res = p.plan_batch(qs)
        single = p.plan(start, goal, seed=1)  # a batch handle also plans one query at a time
    assert [r.status for r in res[1:4]] == [kpax.INVALID_ARG] * 3
    assert res[0].status in (kpax.NOT_FOUND, kpax.BUDGET_EXHAUSTED, kpax.OK)
    assert res[4].status == kpax.BUDGET_EXHAUSTED and res[4].tree_size == 1
    c = cfg.copy()
    c.max_tree = 20000
    with kpax.KPAX(c) as s:
        s.set_obstacles(obs)
        assert _same(res[0], s.plan(start, goal, seed=1))
    assert single.status in (kpax.OK, kpax.NOT_FOUND, kpax.BUDGET_EXHAUSTED)


def test_autotune_keeps_results_and_applies_the_shape(kpax):
    cfg, obs, start, goal = workload(2, seed=0, max_iter=30)
    qs = [kpax.Query(start, goal, seed=300 + i, budget=40000) for i in range(6)]
    with kpax.KPAX(cfg, max_queries=6, node_capacity=240000) as p:
        p.set_obstacles(obs)
        before = p.plan_batch(qs)
        best = p.autotune(qs, repeats=1)
        assert p.launch() == best
        assert best["expand_chunk"] % 16 == 0 and best["expand_blocks"] > 0 and best["scan_blocks"] > 0
        after = p.plan_batch(qs)
    for a, b in zip(before, after):
        assert _same(a, b)


def test_launch_knobs_leave_results_unchanged(kpax):
    """A wave writes out finished children and takes new ones once
    `refill_lanes` of its lanes are idle, the append and frontier scans run
    over tiles of `scan_tile` items, and children are handed out in index
    order or longest edge first (`order_children`, off by default for a
    batch handle), in chunks of `expand_chunk` children per work ticket;
    every child is a function of its index and the scans are exact, so the
    knobs change only speed.  Checked on a single plan's tree and on a
    batch."""
    cfg, obs, start, goal = workload(2, seed=0, max_iter=25)
    qs = [kpax.Query(start, goal, seed=70 + i, budget=50000 + 9000 * i) for i in range(6)]
    shapes = [(1, 1024, 0, 64), (24, 4096, 1, 16), (64, 2048, 1, 32), (16, 4096, 0, 48)]
    out = {}
    for refill, tile, order, chunk in shapes:
        with kpax.KPAX(cfg, max_queries=3, node_capacity=300000) as p:
            p.set_obstacles(obs)
            assert p.launch()["order_children"] == 0
            p.set_launch(refill_lanes=refill, scan_tile=tile, order_children=order, expand_chunk=chunk)
            assert p.launch()["refill_lanes"] == refill and p.launch()["scan_tile"] == tile
            assert p.launch()["expand_chunk"] == chunk
            res = p.plan_batch(qs)
            one = p.plan(start, goal, seed=5)
            tree = p.tree()
        out[(refill, tile, order, chunk)] = (res, one, tree)
    ref_res, ref_one, ref_tree = out[shapes[0]]
    for shape in shapes[1:]:
        res, one, tree = out[shape]
        for a, b in zip(ref_res, res):
            assert _same(a, b), shape
        assert _same(ref_one, one)
        for k in ref_tree:
            assert np.array_equal(ref_tree[k].view(np.uint8), tree[k].view(np.uint8)), (shape, k)


def test_serving_batch_at_its_shape(kpax, oracle):
    """Batched serving at its stated shape: 256 quadcopter queries over 32
    environments (16 forests of 100-600 columns, 16 mazes of 2,000 boxes),
    budgets log-uniform in [16k, 1M] nodes, in one kpax_plan_batch call.
    Every query returns what it returns planned alone; eight of them (per
    environment kind the two largest budgets and two budgets in [2^17, 2^19])
    also equal the oracle's plan, path rows included."""
    import proposal
    cfg = workload(2)[0]
    envs, qs = proposal.serving_workload(cfg)
    assert len(qs) == 256 and len(envs) == 32
    assert sum(o.shape[0] == 2000 for o in envs) == 16
    assert max(q.budget for q in qs) > (1 << 19) and min(q.budget for q in qs) < (1 << 15)
    T = sum(q.budget for q in qs) + 1024
    print(f"serving batch: 256 queries, 32 environments, {T} nodes", flush=True)
    with kpax.KPAX(cfg, max_queries=len(qs), node_capacity=T) as p:
        for e, o in enumerate(envs):
            p.set_environment(e, o)
        res = p.plan_batch(qs)
    print("single plans of the same queries", flush=True)
    with kpax.KPAX(cfg, max_queries=1, node_capacity=1 << 20) as s:
        for e, o in enumerate(envs):
            s.set_environment(e, o)
        seq = [s.plan_batch([q])[0] for q in qs]
    diff = [i for i, (a, b) in enumerate(zip(res, seq)) if not _same(a, b)]
    assert not diff, diff
    found = [r.status == kpax.OK and r.verified == 1 for r in res]
    assert sum(found) / len(found) >= 0.95, sum(found)
    assert any(f for f, q in zip(found, qs) if q.env % 2 == 1)  # mazes solved too
    order = sorted(range(len(qs)), key=lambda i: -qs[i].budget)
    mid = [i for i in range(len(qs)) if (1 << 17) <= qs[i].budget <= (1 << 19)]
    picks = []
    for kind in (0, 1):
        picks += [i for i in order if qs[i].env % 2 == kind][:2] + [i for i in mid if qs[i].env % 2 == kind][:2]
    for i in picks:
        q, r = qs[i], res[i]
        c = cfg.copy()
        c.max_tree = q.budget
        o = oracle.plan(c, envs[q.env], q.start, q.goal, seed=q.seed, threads=16)
        print(f"  query {i} (env {q.env}, budget {q.budget}): gpu {r.status} it {r.iterations} ts {r.tree_size}; "
              f"oracle {o['status']} it {o['iterations']} ts {o['tree_size']}", flush=True)
        assert (r.status, r.iterations, r.tree_size, r.expansions, r.goal_index) == \
            (o["status"], o["iterations"], o["tree_size"], o["expansions"], o["goal_index"])
        if r.status == kpax.OK:
            assert np.array_equal(r.rows.view(np.uint32), oracle_path_rows(o, c).view(np.uint32))


@pytest.mark.gpu
def test_batch_at_the_slot_and_environment_limits(kpax):
    """A handle at its limits: 1,024 query slots (the most kpax_create_batch
    takes) over 64 environments (the most a handle holds), every slot busy at
    once, with and without the learned proposal (its per-slot table then has
    all 1,024 entries).  Every query returns what it returns planned alone; a
    query naming environment 64 is rejected while the rest run; a 1,025-slot
    handle is refused."""
    import proposal
    cfg = workload(2)[0]
    cfg.max_iter = 60
    rng = np.random.default_rng(11)
    envs = [kpax.generate_environment(2, 7000 + e, int(rng.integers(40, 121)), cfg.ws_min, cfg.ws_max)
            for e in range(64)]
    qs = proposal.serving_queries(cfg, envs, 1024, seed=13, min_dist=1.0, max_dist=4.0,
                                  budget_range=(4096, 16384))
    params = np.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kino-pax-fork_amd",
                                  "data", "proposal_forest.npy"))
    T = sum(q.budget for q in qs) + 1024
    with pytest.raises(kpax.KpaxError):
        kpax.KPAX(cfg, max_queries=1025, node_capacity=T)
    bad = kpax.Query(qs[0].start, qs[0].goal, seed=1, budget=4096, env=64)
    for mix in (0.0, 0.5):
        print(f"1024 slots, 64 environments, mix {mix}", flush=True)
        with kpax.KPAX(cfg, max_queries=1024, node_capacity=T) as p:
            for e, o in enumerate(envs):
                p.set_environment(e, o)
            if mix:
                p.set_proposal(params, mix)
            res = p.plan_batch(qs + [bad])
        assert res[-1].status == kpax.INVALID_ARG
        res = res[:-1]
        with kpax.KPAX(cfg, max_queries=1, node_capacity=16384) as s:
            for e, o in enumerate(envs):
                s.set_environment(e, o)
            if mix:
                s.set_proposal(params, mix)
            seq = [s.plan_batch([q])[0] for q in qs]
        diff = [i for i, (a, b) in enumerate(zip(res, seq)) if not _same(a, b)]
        assert not diff, diff[:10]
        found = sum(r.status == kpax.OK for r in res)
        print(f"  found {found} of 1024", flush=True)
        assert found > 512 and any(r.status != kpax.OK for r in res)  # both outcomes at the limits
