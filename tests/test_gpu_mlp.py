"""GPU: the learned control proposal.

Features and the sampler are integer/float work the CPU oracle restates
exactly, so they must match bit for bit.  The network runs on the matrix
cores with bf16 operands and fp32 accumulation; its outputs and gradients are
checked against the numpy oracle, which rounds to bf16 at the same points, to
a relative error of 1e-2 (the accumulation order and the last bit of tanh
differ; a wrong lane layout or operand gives errors of order 1).
"""
import os

import numpy as np
import pytest

import mlp_oracle
from conftest import ROOT
from workloads import random_states, workload

pytestmark = pytest.mark.gpu


def _planner(kpax, **kw):
    cfg, obs, start, goal = workload(2, seed=0, **kw)
    p = kpax.KPAX(cfg)
    p.set_obstacles(obs)
    return p, cfg, obs, start, goal


def _states(cfg, n, rng):
    x0, _, _ = random_states(cfg, n, rng)
    x0[:3, :50] += np.array([[12.0], [0.0], [0.0]], np.float32)  # outside the workspace
    return x0


def test_features_bit_exact(kpax, oracle, rng):
    p, cfg, obs, _, goal = _planner(kpax, max_tree=4096)
    # the forest, then on the same handle 1,000 small boxes (the occupancy
    # bits are ORed over 32-obstacle chunks: 31 full chunks and a partial one),
    # a single box and none
    small = rng.uniform(0, 9.5, (1000, 3)).astype(np.float32)
    small = np.concatenate([small, small + rng.uniform(0.05, 0.5, (1000, 3)).astype(np.float32)], 1)
    with p:
        for o in (obs, small, small[:1], np.zeros((0, 6), np.float32)):
            p.set_obstacles(o)
            x0 = _states(cfg, 5000, rng)
            g = rng.uniform(0, 10, (3, 5000)).astype(np.float32)
            f = p.mlp_features(0, x0, g)
            want = oracle.mlp_features(cfg, o, x0, g)
            assert np.array_equal(f.view(np.uint32), want.view(np.uint32)), o.shape


def test_sampler_bit_exact(kpax, oracle, rng):
    p, cfg, _, _, _ = _planner(kpax, max_tree=4096)
    n = 50000
    ms = np.concatenate([rng.uniform(-1.5, 1.5, (n, 4)), rng.uniform(0.018, 1.65, (n, 4))], 1).astype(np.float32)
    w = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    w[:2] = [[0] * 4, [0xFFFFFFFF] * 4]
    with p:
        u = p.propose_sample(ms, w)
    assert np.array_equal(u.view(np.uint32), oracle.propose_sample(cfg, ms, w).view(np.uint32))


def test_network_forward_matches_oracle(kpax, oracle, rng):
    p, cfg, obs, _, _ = _planner(kpax, max_tree=4096)
    params = kpax.mlp_init_params(11)
    params[mlp_oracle.B1:mlp_oracle.W2] = rng.normal(0, 0.3, 128)
    with p:
        x0 = _states(cfg, 3000, rng)
        g = rng.uniform(0, 10, (3, 3000)).astype(np.float32)
        X = p.mlp_features(0, x0, g)
        out = p.mlp_infer(params, X)
    ref, _ = mlp_oracle.forward(params, X)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"network forward: max rel err {err:.2e}", flush=True)
    assert err < 1e-2


def test_trainer_gradients_match_oracle(kpax, rng):
    X = rng.normal(0, 1, (1000, 80)).astype(np.float32)
    X[:, 15:79] = rng.uniform(0, 1, (1000, 64)) < 0.2
    X[:, 79] = 0
    Y = rng.uniform(-1, 1, (1000, 4)).astype(np.float32)
    params = kpax.mlp_init_params(7)
    params[mlp_oracle.W3:mlp_oracle.B3] *= 5
    with kpax.Trainer(max_batch=2048) as t:
        t.set_params(params)
        loss, g = t.gradients(X, Y)
    rl, rg = mlp_oracle.loss_and_grads(params, X, Y)
    assert abs(loss - rl) <= 1e-4 * (1 + abs(rl))
    for name, a, b in (("W1", mlp_oracle.W1, mlp_oracle.B1), ("b1", mlp_oracle.B1, mlp_oracle.W2),
                       ("W2", mlp_oracle.W2, mlp_oracle.B2), ("b2", mlp_oracle.B2, mlp_oracle.W3),
                       ("W3", mlp_oracle.W3, mlp_oracle.B3), ("b3", mlp_oracle.B3, mlp_oracle.NPARAMS)):
        e = np.linalg.norm(g[a:b] - rg[a:b]) / np.linalg.norm(rg[a:b])
        print(f"gradient {name}: rel err {e:.2e}", flush=True)
        assert e < 1e-2, name


def test_training_fits_a_known_conditional(kpax, rng):
    # targets are a fixed function of the features plus noise of sigma 0.1:
    # the trained network must beat the best constant predictor by far
    n = 32768
    X = rng.normal(0, 1, (n, 80)).astype(np.float32)
    X[:, 79] = 0
    A = rng.normal(0, 0.3, (80, 4))
    Y = (np.tanh(X @ A) + 0.1 * rng.standard_normal((n, 4))).clip(-1, 1).astype(np.float32)
    with kpax.Trainer(max_batch=4096) as t:
        t.set_params(kpax.mlp_init_params(3))
        losses = t.fit(X, Y, epochs=12, batch=1024, lr=3e-3, seed=5)
        params = t.params()
    print("epoch losses", np.round(losses, 3), flush=True)
    out, _ = mlp_oracle.forward(params, X[:4096])
    mse = np.mean((out[:, :4] - Y[:4096]) ** 2)
    assert losses[-1] < losses[0] - 1.0
    assert mse < 0.1 * Y.var()


def test_planner_with_proposal(kpax, oracle):
    p, cfg, obs, start, goal = _planner(kpax, max_tree=1 << 18, max_iter=60)
    params = kpax.mlp_init_params(2)
    # a network that always asks for full thrust: mean of control 0 at +1, small sigma
    params[mlp_oracle.W3:mlp_oracle.B3] = 0
    params[mlp_oracle.B3:mlp_oracle.B3 + 8] = [3, 0, 0, 0, -4, -1, -1, -1]
    with p:
        plain = p.plan(start, goal, seed=4)
        p.set_proposal(params, 0.5)
        a = p.plan(start, goal, seed=4)
        ta = p.tree()
        b = p.plan(start, goal, seed=4)
        p.set_proposal(None)
        again = p.plan(start, goal, seed=4)
    assert (a.status, a.tree_size, a.expansions) == (b.status, b.tree_size, b.expansions)
    assert (again.status, again.tree_size, again.expansions, again.iterations) == \
        (plain.status, plain.tree_size, plain.expansions, plain.iterations)
    full = np.mean(ta["controls"][0, 1:] == np.float32(cfg.ctrl_max[0]))
    print(f"nodes at full thrust with the proposal: {full:.2f}", flush=True)
    assert full > 0.1  # the uniform sampler never hits the bound exactly
    if a.status == kpax.OK:
        assert a.verified == 1


def test_batch_with_proposal_matches_single_plans(kpax):
    cfg, obs, start, goal = workload(2, seed=0, max_iter=40)
    params = kpax.mlp_init_params(9)
    qs = [kpax.Query(start, goal, seed=60 + i, budget=40000 + 9000 * i) for i in range(6)]
    with kpax.KPAX(cfg, max_queries=3, node_capacity=250000) as p:
        p.set_obstacles(obs)
        p.set_proposal(params, 0.5)
        res = p.plan_batch(qs)
    for q, r in zip(qs, res):
        c = cfg.copy()
        c.max_tree = q.budget
        with kpax.KPAX(c) as s:
            s.set_obstacles(obs)
            s.set_proposal(params, 0.5)
            o = s.plan(q.start, q.goal, seed=q.seed)
        assert (r.status, r.tree_size, r.expansions, r.iterations) == (o.status, o.tree_size, o.expansions,
                                                                       o.iterations)
        assert np.array_equal(r.rows.view(np.uint32), o.rows.view(np.uint32))


def test_proposal_budget_policy(kpax):
    """Queries under the proposal's minimum budget sample uniformly (the plain
    planner's results bit for bit); the others use the mixture, in the same
    batch."""
    cfg, obs, start, goal = workload(2, seed=0, max_iter=40)
    params = kpax.mlp_init_params(9)
    budgets = [30000, 80000, 45000, 95000]
    qs = [kpax.Query(start, goal, seed=90 + i, budget=b) for i, b in enumerate(budgets)]
    with kpax.KPAX(cfg, max_queries=2, node_capacity=200000) as p:
        p.set_obstacles(obs)
        p.set_proposal(params, 0.5, min_budget=60000)
        res = p.plan_batch(qs)
    for q, r in zip(qs, res):
        c = cfg.copy()
        c.max_tree = q.budget
        with kpax.KPAX(c) as s:
            s.set_obstacles(obs)
            if q.budget >= 60000:
                s.set_proposal(params, 0.5)
            o = s.plan(q.start, q.goal, seed=q.seed)
        assert (r.status, r.tree_size, r.expansions, r.iterations) == (o.status, o.tree_size, o.expansions,
                                                                       o.iterations), q.budget
        assert np.array_equal(r.rows.view(np.uint32), o.rows.view(np.uint32))


PARAMS = os.path.join(ROOT, "kino-pax-fork_amd", "data", "proposal_forest.npy")


@pytest.mark.parametrize("budget,seed", [(1 << 17, 21), (1 << 18, 22), (1 << 20, 24)])
def test_proposal_plan_matches_oracle_with_injected_gaussians(kpax, oracle, budget, seed):
    """A plan with the shipped network mixed in (50/50) equals the oracle's plan
    tree for tree when the oracle is given each node's Gaussian as the GPU's
    network produced it (the network's own bf16 arithmetic is pinned against
    its float64 oracle in the tests above): controls drawn from the proposal,
    the mixture draws and everything downstream agree bit for bit."""
    cfg, obs, start, goal = workload(2, seed=0, max_tree=budget)
    params = np.load(PARAMS)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        p.set_proposal(params, 0.5)
        r = p.plan(start, goal, seed=seed)
        tree = p.tree()
        table = p.proposal_table()
    o = oracle.plan(cfg, obs, start, goal, seed=seed, threads=16, proposal=table, mix=0.5)
    print(f"proposal plan {budget}: gpu {r.status} it {r.iterations} ts {r.tree_size}; oracle {o['status']} "
          f"it {o['iterations']} ts {o['tree_size']}", flush=True)
    assert (r.status, r.iterations, r.tree_size, r.expansions) == (o["status"], o["iterations"], o["tree_size"],
                                                                   o["expansions"])
    assert np.array_equal(tree["parents"], o["parents"])
    assert np.array_equal(tree["states"].view(np.uint32), o["states"].view(np.uint32))
    assert np.array_equal(tree["controls"].view(np.uint32), o["controls"].view(np.uint32))
    # the plan really used the proposal: the uniform sampler never hits the thrust bound exactly
    assert np.mean(tree["controls"][0, 1:] == np.float32(cfg.ctrl_max[0])) > 0.005
    # and the oracle with no proposal differs
    plain = oracle.plan(cfg, obs, start, goal, seed=seed, threads=16, want_tree=False)
    assert (plain["tree_size"], plain["expansions"]) != (o["tree_size"], o["expansions"])


def test_proposal_pass_is_the_network_on_each_node(kpax):
    """The planner's proposal pass (features in registers, network on the
    matrix cores, 16 nodes per wave) gives every expanded node of a plan the
    Gaussian that the feature operator and the inference operator give for that
    node's state: means bit for bit, sigmas = exp(clamp(log sigma, -4, 1/2))
    to float rounding."""
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1 << 17)
    params = np.load(PARAMS)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        p.set_proposal(params, 0.5)
        r = p.plan(start, goal, seed=23)
        tree, table = p.tree(), p.proposal_table()
        used = np.unique(tree["parents"][1:])  # every node the plan expanded
        gg = np.tile(np.asarray(goal[:3], np.float32).reshape(3, 1), (1, used.size))
        raw = p.mlp_infer(params, p.mlp_features(0, tree["states"][:, used], gg)).T
    print(f"plan {r.status} it {r.iterations} tree {r.tree_size}: {used.size} expanded nodes", flush=True)
    assert used.size > 10000
    assert np.array_equal(table[:4, used].view(np.uint32), raw[:4].view(np.uint32))
    assert np.allclose(table[4:, used], np.exp(np.clip(raw[4:], -4.0, 0.5)), rtol=1e-6, atol=0)


def test_trained_proposal_gains_over_uniform_sampling(kpax):
    """The shipped network (trained on plain-planner paths in other forests),
    mixed 50/50 with uniform sampling, on 96 held-out queries in 8 held-out
    forests at a 2^18-node budget: it solves at least as many queries as the
    uniform sampler and, on the queries both solve, needs fewer iterations and
    fewer expansions on average."""
    import proposal
    cfg = workload(2)[0]
    cfg.max_tree = 1 << 18
    cfg.max_iter = 100
    envs = proposal.forest_environments(cfg, 8, seed=2)
    qs = proposal.serving_queries(cfg, envs, 96, seed=4)
    params = np.load(PARAMS)
    with kpax.KPAX(cfg) as p:
        for e, o in enumerate(envs):
            p.set_environment(e, o)
        plain = proposal.evaluate(p, qs, None, phases=False)
        learned = proposal.evaluate(p, qs, params, 0.5, phases=False)
    both = [(a, b) for a, b in zip(plain["results"], learned["results"]) if a.status == 0 and b.status == 0]
    it_p = np.mean([a.iterations for a, _ in both])
    it_l = np.mean([b.iterations for _, b in both])
    ex_p = np.mean([a.expansions for a, _ in both])
    ex_l = np.mean([b.expansions for _, b in both])
    print(f"success plain {plain['success']:.3f} learned {learned['success']:.3f}; on {len(both)} both solve: "
          f"iterations {it_p:.2f} -> {it_l:.2f}, expansions {ex_p:.0f} -> {ex_l:.0f}", flush=True)
    assert learned["success"] >= plain["success"] and learned["success"] >= 0.95
    assert len(both) >= 80
    assert it_l < 0.95 * it_p and ex_l < ex_p
