"""GPU parity: every hot-path operator of libkpax against the CPU oracle.

Integer, index and verdict outputs must be bit-exact.  Propagated FP32 states
must equal the oracle's float path bit for bit (same operation order, no
contraction, the same sin/cos routine) and lie within 1e-5 relative of the
oracle's FP64 path (|x32 - x64| <= 1e-5 * max(1, |x64|) per component).
"""
import numpy as np
import pytest

from workloads import oracle_path_rows, random_states, workload

pytestmark = pytest.mark.gpu

MODELS = [1, 2, 3]  # workload ids: double integrator, quadcopter, Dubins


def _planner(kpax, k, **kw):
    cfg, obs, start, goal = workload(k, **kw)
    p = kpax.KPAX(cfg)
    p.set_obstacles(obs)
    return p, cfg, obs, start, goal


@pytest.mark.parametrize("k", MODELS)
def test_propagate_bit_exact_vs_oracle_float(kpax, oracle, rng, k):
    p, cfg, obs, _, _ = _planner(kpax, k, max_tree=4096)
    print(f"propagate parity workload {k}", flush=True)
    x0, u, steps = random_states(cfg, 20000, rng)
    x1, valid, fail = p.propagate_batch(x0, u, steps)
    o1, ovalid, ofail = oracle.propagate(cfg, obs, x0, u, steps)
    assert np.array_equal(valid, ovalid)
    assert np.array_equal(fail, ofail)
    assert np.array_equal(x1.view(np.uint32), o1.view(np.uint32)), \
        f"max |diff| {np.max(np.abs(x1 - o1))}"
    assert valid.any() and (~valid).any()


@pytest.mark.parametrize("k", MODELS)
def test_propagate_within_1e5_of_fp64(kpax, oracle, rng, k):
    p, cfg, _, _, _ = _planner(kpax, k, max_tree=4096, n_obs=0)
    x0, u, steps = random_states(cfg, 4000, rng)
    x1, valid, fail = p.propagate_batch(x0, u, steps)
    d1, dvalid, dfail = oracle.propagate(cfg, None, x0, u, steps, double=True)
    both = valid & dvalid & (fail == dfail)
    assert both.mean() > 0.5
    err = np.abs(x1[:, both].astype(np.float64) - d1[:, both]) / np.maximum(1.0, np.abs(d1[:, both]))
    assert err.max() <= 1e-5, err.max()


def test_collide_matches_brute_force(kpax, oracle, rng):
    cfg = kpax.config_workload(2)
    # 2,000 boxes, half packed into a 1 x 1 x 0.5 corner: many cells hold more
    # than 16 obstacles (the reference's grid keeps only 16, spatialHash.cu:44)
    c = rng.uniform(0, 1, (2000, 3)) * np.array([10, 10, 5])
    c[:1000] = 4.0 + rng.uniform(0, 1, (1000, 3)) * np.array([1, 1, 0.5])
    e = rng.uniform(0.05, 0.6, (2000, 3))
    obs = np.concatenate([c - e / 2, c + e / 2], 1).astype(np.float32)
    obs[:, :3] = np.maximum(obs[:, :3], 0)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        start, items = p.grid()
        assert np.diff(start).max() > 16
        n = 100000
        a = (rng.uniform(-0.5, 1.5, (3, n)) * np.array([[10], [10], [5]])).astype(np.float32)
        b = a + rng.uniform(-0.3, 0.3, (3, n)).astype(np.float32)
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        # exact face contact and zero-length segments
        lo[:, :500] = obs[:500, 3:].T
        hi[:, :500] = obs[:500, 3:].T + 0.01
        hi[:, 500:1000] = lo[:, 500:1000]
        got = p.collide_batch(lo, hi)
        want = oracle.collide(obs, lo, hi)
        assert np.array_equal(got, want)
        assert (~got).sum() > 1000
        # touching an obstacle's face counts as free (collisionCheck.cu:7): each
        # face-touching box is separated from the obstacle it touches
        assert oracle.collide(obs[:1], lo[:, :1], hi[:, :1])[0]


@pytest.mark.parametrize("k", MODELS)
def test_grid_csr_matches_oracle_counts(kpax, oracle, k):
    p, cfg, obs, _, _ = _planner(kpax, k, max_tree=4096)
    start, items = p.grid()
    counts = oracle.grid_counts(cfg, obs)
    assert np.array_equal(np.diff(start), counts)
    assert start[0] == 0 and start[-1] == items.size
    for cidx in np.flatnonzero(counts)[:2000]:
        seg = items[start[cidx]:start[cidx + 1]]
        assert np.all(np.diff(seg) > 0)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_lds_grid_verdicts_equal_brute_force(kpax, oracle, rng, k):
    """The collision grid staged in LDS (a coarse grid of 16-bit obstacle
    lists) and the CSR grid in global memory give the brute-force verdicts for
    random segments, segments ending on an obstacle face (free) and
    zero-length segments, and propagated edges are bit-identical both ways.
    The 2,000-box maze gets no image: the coarse grids that fit the LDS
    budget list ~17x more obstacles per cell than its 128^3 CSR grid."""
    cfg, obs, _, _ = workload(k, seed=1)
    lo_ws = np.array(cfg.ws_min, np.float32)[:, None]
    ext = (np.array(cfg.ws_max, np.float32) - np.array(cfg.ws_min, np.float32))[:, None]
    n = 200000
    a = (lo_ws + ext * rng.uniform(-0.05, 1.05, (3, n))).astype(np.float32)
    b = (a + ext * rng.uniform(-0.03, 0.03, (3, n))).astype(np.float32)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    pick = rng.integers(0, obs.shape[0], 2000)
    hi[0, :2000] = obs[pick, 0]
    lo[0, :2000] = np.minimum(lo[0, :2000], hi[0, :2000])
    lo[:, 2000:4000] = hi[:, 2000:4000]
    x0, u, steps = random_states(cfg, 20000, rng)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        dims, nbytes = p.lds_grid()
        print(f"workload {k}: LDS grid {dims}, {nbytes} bytes", flush=True)
        assert p.launch()["lds_grid"] == 1
        assert (nbytes == 0) if k == 3 else (nbytes > 0)
        staged = p.collide_batch(lo, hi)
        e1 = p.propagate_batch(x0, u, steps)
        p.set_launch(lds_grid=0)
        csr = p.collide_batch(lo, hi)
        e0 = p.propagate_batch(x0, u, steps)
    want = oracle.collide(obs, lo, hi)
    assert np.array_equal(staged, want) and np.array_equal(csr, want)
    assert 0.01 < want.mean() < 0.999
    assert np.array_equal(e1[0].view(np.uint32), e0[0].view(np.uint32)) and np.array_equal(e1[2], e0[2])


def test_plan_identical_with_and_without_lds_grid(kpax):
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1 << 18)
    with kpax.KPAX(cfg) as p:
        r1 = p.plan(start, goal, obs, seed=4)
        t1 = p.tree()
        p.set_launch(lds_grid=0)
        r0 = p.plan(start, goal, obs, seed=4)
        t0 = p.tree()
    assert (r1.status, r1.tree_size, r1.iterations, r1.expansions) == (r0.status, r0.tree_size, r0.iterations,
                                                                       r0.expansions)
    assert np.array_equal(t1["states"].view(np.uint32), t0["states"].view(np.uint32))


@pytest.mark.parametrize("k", [1, 3], ids=["double_integrator", "dubins_airplane"])
def test_other_models_with_and_without_lds_grid_bit_exact_vs_oracle(kpax, oracle, k):
    """The single-query expansion of the double integrator and of the Dubins
    airplane compiled for both grid paths (its LDS image, and the CSR grid in
    global memory) plans the oracle's tree; the Dubins airplane flies a
    generated column forest here, since its maze has no LDS image."""
    cfg, obs, start, goal = workload(k, seed=0, max_tree=1 << 15, max_iter=10 if k == 1 else 8)
    if k == 3:
        obs = kpax.generate_environment(2, 5, 40, cfg.ws_min, cfg.ws_max, start[:3],
                                        list(goal[:3]) + [cfg.goal_radius])
    o = oracle.plan(cfg, obs, start, goal, seed=41, threads=8)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        assert p.lds_grid()[1] > 0  # the environment has an LDS image
        for lds in (1, 0):
            p.set_launch(lds_grid=lds)
            r = p.plan(start, goal, seed=41)
            t = p.tree()
            assert (r.status, r.iterations, r.tree_size) == (o["status"], o["iterations"], o["tree_size"]), lds
            assert np.array_equal(t["parents"], o["parents"])
            assert np.array_equal(t["states"].view(np.uint32), o["states"].view(np.uint32))


@pytest.mark.parametrize("k", MODELS)
def test_plan_identical_with_and_without_wave_priority(kpax, k):
    """Issue priorities reorder the expansion's instruction issue only: the
    tree is the same bit for bit with them on and off, for every model."""
    cfg, obs, start, goal = workload(k, seed=0, max_tree=1 << 17, max_iter=12)
    out = {}
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        for w in (1, 0):
            p.set_launch(wave_priority=w)
            assert p.launch()["wave_priority"] == w
            r = p.plan(start, goal, seed=11)
            out[w] = ((r.status, r.tree_size, r.iterations, r.expansions, r.integrated_steps), p.tree())
    assert out[1][0] == out[0][0]
    assert np.array_equal(out[1][1]["states"].view(np.uint32), out[0][1]["states"].view(np.uint32))
    assert np.array_equal(out[1][1]["parents"], out[0][1]["parents"])


@pytest.mark.parametrize("k", MODELS)
def test_regions_match_oracle(kpax, oracle, rng, k):
    p, cfg, _, _, _ = _planner(kpax, k, max_tree=4096)
    lo = np.array(cfg.ws_min, np.float32)[:, None]
    ext = (np.array(cfg.ws_max, np.float32) - np.array(cfg.ws_min, np.float32))[:, None]
    n = 50000
    st = (lo + ext * rng.uniform(-0.1, 1.1, (3, n))).astype(np.float32)
    # exact region and sub-region boundaries
    L = np.array(cfg.r1_len)[:, None]
    st[:, :1000] = (lo + ext * rng.integers(0, 4 * L.max() + 1, (3, 1000)) / (4 * L)).astype(np.float32)
    st[0, 1000] = np.nan
    r1, r2 = p.region_batch(st)
    o1, o2 = oracle.regions(cfg, st)
    assert np.array_equal(r1, o1) and np.array_equal(r2, o2)


@pytest.mark.parametrize("k", MODELS)
def test_histogram_matches_oracle(kpax, oracle, rng, k):
    p, cfg, _, _, _ = _planner(kpax, k, max_tree=4096)
    lo = np.array(cfg.ws_min, np.float32)[:, None]
    ext = (np.array(cfg.ws_max, np.float32) - np.array(cfg.ws_min, np.float32))[:, None]
    n = 60000
    # clustered states: heavy atomic contention on a few regions
    centre = lo + ext * rng.uniform(0.2, 0.8, (3, 1))
    st = (centre + 0.05 * ext * rng.standard_normal((3, n))).astype(np.float32)
    valid = rng.uniform(0, 1, n) < 0.7
    t0 = rng.integers(0, 5, cfg.r1_count).astype(np.int32)
    got = p.region_histogram(st, valid, total=t0 + 3, valid_count=t0, covered=np.zeros_like(t0))
    want = oracle.histogram(cfg, st, valid, 1, total=t0 + 3, valid_count=t0, covered=np.zeros_like(t0))
    for g, w in zip(got, want[:3]):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("k", MODELS)
def test_scores_bit_exact(kpax, oracle, rng, k):
    p, cfg, _, _, _ = _planner(kpax, k, max_tree=4096)
    R1 = cfg.r1_count
    valid = rng.integers(0, 50, R1).astype(np.int32)
    valid[rng.uniform(0, 1, R1) < 0.5] = 0
    total = valid + rng.integers(0, 30, R1).astype(np.int32)
    covered = np.minimum(valid, rng.integers(0, cfg.r2_per + 1, R1)).astype(np.int32)
    got = p.scores(valid, total, covered)
    want, _ = oracle.scores(cfg, valid, total, covered)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("n", [1, 1023, 1024, 1025, 300000, 2_100_000])
def test_repeat_expand(kpax, rng, n):
    with kpax.KPAX(kpax.config_workload(1)) as p:
        w = rng.integers(0, 4, n).astype(np.uint32)
        w[rng.uniform(0, 1, n) < 0.3] = 15
        out, total = p.repeat_expand(w)
        assert total == int(w.sum())
        assert np.array_equal(out, np.repeat(np.arange(n, dtype=np.uint32), w))


def test_repeat_expand_zeros_and_cap(kpax):
    with kpax.KPAX(kpax.config_workload(1)) as p:
        out, total = p.repeat_expand(np.zeros(5000, np.uint32))
        assert total == 0 and out.size == 0
        w = np.full(5000, 3, np.uint32)
        out, total = p.repeat_expand(w, cap=100)
        assert total == 15000 and np.array_equal(out, np.repeat(np.arange(34, dtype=np.uint32), 3)[:100])


@pytest.mark.parametrize("n", [1, 1024, 1025, 70000])
def test_scan_compact(kpax, rng, n):
    """Frontier compaction: the indices of the set flags in increasing order,
    the last element included when set (the reference's next-frontier
    compaction drops it, KPAX.cu:523-525); all-clear and all-set masks."""
    with kpax.KPAX(kpax.config_workload(1)) as p:
        for flags in (rng.uniform(0, 1, n) < 0.4, np.zeros(n, bool), np.ones(n, bool)):
            flags = flags.copy()
            flags[-1] = True if flags.any() else flags[-1]
            out, count = p.scan_compact(flags)
            want = np.flatnonzero(flags).astype(np.uint32)
            assert count == want.size and np.array_equal(out, want)


def _compare_plans(kpax, oracle, k, seed, **kw):
    cfg, obs, start, goal = workload(k, seed=seed, **kw)
    with kpax.KPAX(cfg) as p:
        r = p.plan(start, goal, obs, seed=seed + 77)
        tree = p.tree()
        reg = p.regions()
    o = oracle.plan(cfg, obs, start, goal, seed=seed + 77, threads=8)
    print(f"plan workload {k} seed {seed}: gpu status {r.status} it {r.iterations} ts {r.tree_size}; "
          f"oracle status {o['status']} it {o['iterations']} ts {o['tree_size']}", flush=True)
    assert r.status == o["status"]
    assert r.iterations == o["iterations"]
    assert r.tree_size == o["tree_size"]
    assert r.expansions == o["expansions"]
    assert np.array_equal(tree["parents"], o["parents"])
    assert np.array_equal(tree["steps"], o["steps"])
    assert np.array_equal(tree["regions"], o["regions"])
    assert np.array_equal(tree["states"].view(np.uint32), o["states"].view(np.uint32))
    assert np.array_equal(tree["controls"].view(np.uint32), o["controls"].view(np.uint32))
    for key in ("valid", "total", "covered"):
        assert np.array_equal(reg[key], o[key]), key
    assert np.array_equal(reg["scores"].view(np.uint32), o["scores"].view(np.uint32))
    if r.status == kpax.OK:
        assert r.goal_index == o["goal_index"]
        assert r.verified == 1
    return r


@pytest.mark.parametrize("k,seed,kw", [
    (1, 0, dict(max_tree=20000, max_iter=12)),
    (1, 1, dict(max_tree=65536, max_iter=8)),
    (1, 2, dict()),  # the double-integrator configuration at its full size
    (2, 0, dict(max_tree=1 << 17)),
    (2, 3, dict(max_tree=1 << 16, max_iter=10)),
    (3, 0, dict(max_tree=1 << 17, max_iter=15)),
])
def test_plan_tree_bit_exact_vs_oracle(kpax, oracle, k, seed, kw):
    _compare_plans(kpax, oracle, k, seed, **kw)


@pytest.mark.parametrize("k,kw", [(1, dict(max_tree=30000, max_iter=12)), (2, dict(max_tree=1 << 17, max_iter=25))])
def test_plan_with_goal_progress_pruning_bit_exact(kpax, oracle, k, kw):
    cfg, obs, start, goal = workload(k, seed=2, **kw)
    cfg.prune = 1  # PruneKPAX defaults: exploration 0.3, goal bias 0.7, progress scale 10
    with kpax.KPAX(cfg) as p:
        r = p.plan(start, goal, obs, seed=11)
        tree = p.tree()
    o = oracle.plan(cfg, obs, start, goal, seed=11, threads=8)
    print(f"pruned plan workload {k}: gpu {r.status} it {r.iterations} ts {r.tree_size}; oracle {o['status']} "
          f"it {o['iterations']} ts {o['tree_size']}", flush=True)
    assert (r.status, r.iterations, r.tree_size, r.expansions) == (o["status"], o["iterations"], o["tree_size"],
                                                                   o["expansions"])
    assert np.array_equal(tree["parents"], o["parents"])
    assert np.array_equal(tree["states"].view(np.uint32), o["states"].view(np.uint32))


def test_plan_full_size_tree_bit_exact_vs_oracle(kpax, oracle):
    """The headline workload at its full size (12D quadcopter, 400-column
    forest, 2^20-node cap, default iteration limit): tree, controls, step
    counts, regions and scores equal the oracle's bit for bit."""
    _compare_plans(kpax, oracle, 2, 5)


def test_plan_dubins_full_size_tree_bit_exact_vs_oracle(kpax, oracle):
    """The Dubins maze at its full size (2,000 boxes with a 0.1-wide passage,
    128^3 grid, 2^21-node cap, default iteration limit): this plan passes the
    passage and reaches the goal after ~140 iterations and 144 M expansions;
    tree, regions and scores equal the oracle's bit for bit."""
    cfg, obs, start, goal = workload(3, seed=0)
    with kpax.KPAX(cfg) as p:
        r = p.plan(start, goal, obs, seed=5)
        tree = p.tree()
        reg = p.regions()
    print(f"full-size Dubins plan: gpu status {r.status} it {r.iterations} ts {r.tree_size}", flush=True)
    o = oracle.plan(cfg, obs, start, goal, seed=5, threads=16)
    print(f"  oracle status {o['status']} it {o['iterations']} ts {o['tree_size']}", flush=True)
    assert (r.status, r.iterations, r.tree_size, r.expansions) == (o["status"], o["iterations"], o["tree_size"],
                                                                   o["expansions"])
    assert r.status == kpax.OK and r.goal_index == o["goal_index"] and r.verified == 1
    assert np.array_equal(tree["parents"], o["parents"])
    assert np.array_equal(tree["states"].view(np.uint32), o["states"].view(np.uint32))
    assert np.array_equal(tree["controls"].view(np.uint32), o["controls"].view(np.uint32))
    for key in ("valid", "total", "covered"):
        assert np.array_equal(reg[key], o[key]), key
    assert np.array_equal(reg["scores"].view(np.uint32), o["scores"].view(np.uint32))
    assert np.array_equal(r.rows.view(np.uint32), oracle_path_rows(o, cfg).view(np.uint32))


def test_dubins_rows_and_dump_in_reference_control_order(kpax, oracle, tmp_path):
    """Result rows and planBench sample rows carry the Dubins controls as
    (yawRate, pitchRate, a), x1[6..8] of statePropagatorSpatialHash.cu:223-225."""
    cfg, obs, start, _ = workload(3, seed=0, max_tree=1 << 14, max_iter=30)
    goal = np.array([0.22, 0.5, 0.5], np.float32)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        r = p.plan_dump(start, goal, tmp_path, seed=3, bench_itr=0)
        tree = p.tree()
    o = oracle.plan(cfg, obs, start, goal, seed=3)
    assert r.status == kpax.OK and r.verified == 1
    assert np.array_equal(r.rows.view(np.uint32), oracle_path_rows(o, cfg).view(np.uint32))
    assert np.array_equal(r.rows[1:, 6:9], tree["controls"][[1, 2, 0]][:, r.node_index[1:]].T)
    last = np.loadtxt(tmp_path / "Samples" / "Samples0" / f"samples{r.iterations}.csv", delimiter=",", ndmin=2)
    assert last.shape == (r.tree_size, 10)
    assert np.allclose(last[:, 6:9], tree["controls"][[1, 2, 0]].T, atol=1e-6)


def test_plan_quadcopter_full_size_finds_verified_path(kpax):
    cfg, obs, start, goal = workload(2, seed=0)
    with kpax.KPAX(cfg) as p:
        r = p.plan(start, goal, obs, seed=1)
        assert r.status == kpax.OK, r
        assert r.verified == 1
        assert np.linalg.norm(r.rows[-1, :3] - goal) < cfg.goal_radius
        assert np.array_equal(r.rows[0, :12], start)
        ok, dev = kpax.verify_path(cfg, obs, r.rows, r.steps, list(goal))
        assert ok and dev < 1e-3


def test_plan_dump_matches_the_tree(kpax, tmp_path):
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1 << 16, max_iter=30)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        plain = p.plan(start, goal, seed=8)
        r = p.plan_dump(start, goal, tmp_path, seed=8, bench_itr=0)
        tree = p.tree()
    assert (r.status, r.tree_size, r.iterations, r.expansions) == (plain.status, plain.tree_size, plain.iterations,
                                                                   plain.expansions)
    d = tmp_path
    last = np.loadtxt(d / "Samples" / "Samples0" / f"samples{r.iterations}.csv", delimiter=",", ndmin=2)
    par = np.loadtxt(d / "Parents" / "Parents0" / f"parents{r.iterations}.csv", ndmin=1)
    assert last.shape == (r.tree_size, 17) and np.array_equal(par, tree["parents"])
    assert np.allclose(last[:, :12], tree["states"].T, atol=5e-6 * (1 + np.abs(tree["states"].T)))
    assert len(np.loadtxt(d / "TreeSize" / "TreeSize0" / "treeSize.csv", ndmin=1)) == r.iterations


def test_failed_run_leaves_the_handle_usable(kpax, tmp_path):
    """A plan stopped by an error mid-run (the planBench dump cannot write its
    first samples file, forced from the host) leaves no slot running: the
    handle's next plan equals the same plan on a fresh handle."""
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1 << 16, max_iter=30)
    (tmp_path / "Samples" / "Samples0" / "samples1.csv").mkdir(parents=True)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        with pytest.raises(kpax.KpaxError):
            p.plan_dump(start, goal, tmp_path, seed=8, bench_itr=0)
        r = p.plan(start, goal, seed=9)
    with kpax.KPAX(cfg) as q:
        fresh = q.plan(start, goal, obs, seed=9)
    assert (r.status, r.iterations, r.tree_size, r.expansions) == (fresh.status, fresh.iterations,
                                                                   fresh.tree_size, fresh.expansions)
    assert np.array_equal(r.rows.view(np.uint32), fresh.rows.view(np.uint32))


def test_environment_rebuilds_reuse_buffers(kpax, oracle):
    """Replacing a handle's obstacles (more, then fewer, then none, then the
    first set again) rebuilds the CSR grid and the LDS image in the kept
    buffers: each build equals the oracle's counts, and a plan after it equals
    the same plan on a fresh handle."""
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1 << 16, max_iter=30)
    big = kpax.generate_environment(3, 7, 2000, cfg.ws_min, cfg.ws_max, start[:3], list(goal[:3]) + [cfg.goal_radius])
    with kpax.KPAX(cfg) as p:
        for o in (obs, big, np.zeros((0, 6), np.float32), obs[:50], obs):
            p.set_obstacles(o)
            st, items = p.grid()
            assert np.array_equal(np.diff(st), oracle.grid_counts(cfg, o)) and st[-1] == items.size
            r = p.plan(start, goal, seed=12)
            with kpax.KPAX(cfg) as q:
                f = q.plan(start, goal, o, seed=12)
            assert (r.status, r.tree_size, r.expansions) == (f.status, f.tree_size, f.expansions), o.shape
            print(f"{o.shape[0]} obstacles: LDS grid {p.lds_grid()}, status {r.status}, tree {r.tree_size}",
                  flush=True)
