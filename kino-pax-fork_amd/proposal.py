"""Learned control proposal: training data from the plain planner, training on
the GPU, and the comparison with the plain sampler.

The reference samples every control uniformly (statePropagatorSpatialHash.cu:
236-241); it has no counterpart to any of this.  The pipeline:

1. `serving_queries` draws start/goal pairs in free space of seeded forests.
2. The plain planner (`KPAX.plan_batch`, uniform controls) solves them; every
   edge of every returned path is a training example: the features of the
   edge's parent node (state, goal-relative position, 4x4x4 occupancy window)
   and the control that edge applied, normalised to [-1, 1].
3. `train` fits the network (80 -> 128 -> 128 -> 8, Gaussian negative
   log-likelihood) with the GPU trainer's forward/backward kernels and Adam.
4. `evaluate` plans held-out queries with the plain sampler and with the
   50/50 mixture of learned and uniform controls, on the same GPU, seeds and
   budgets, and reports success rate, iterations and time to solution.
"""
import statistics
import time

import numpy as np

import kpax


def forest_environments(cfg, n_env, seed, n_min=100, n_max=600):
    """Seeded forests over the configured workspace."""
    rng = np.random.default_rng(seed)
    return [kpax.generate_environment(2, int(seed) * 1000 + e, int(rng.integers(n_min, n_max + 1)), cfg.ws_min,
                                      cfg.ws_max) for e in range(n_env)]


def _free(p, obs, margin):
    if obs.shape[0] == 0:
        return True
    lo, hi = obs[:, :3] - margin, obs[:, 3:] + margin
    return not np.any(np.all((p >= lo) & (p <= hi), axis=1))


def budget_distance(budget, lo_budget=16384, hi_budget=1 << 20, near=2.0, far=9.0):
    """Largest start-goal distance a query of this tree budget asks for: `near`
    metres at `lo_budget` nodes up to `far` at `hi_budget`, linear in the
    log of the budget."""
    f = np.clip(np.log(budget / lo_budget) / np.log(hi_budget / lo_budget), 0.0, 1.0)
    return near + (far - near) * f


def serving_queries(cfg, envs, n, seed, budget=0, min_dist=4.0, budget_range=None, clearance=0.2, tries=100000,
                    max_dist=None):
    """n start/goal pairs `min_dist` to `max_dist` apart (max_dist a number or
    a function of the query's budget), both `clearance` away from every
    obstacle and 0.5 inside the workspace faces, each in environment
    i % len(envs).  With `budget_range`, budgets are log-uniform in it."""
    rng = np.random.default_rng(seed)
    lo = np.array(cfg.ws_min, np.float64) + 0.5
    hi = np.array(cfg.ws_max, np.float64) - 0.5
    out = []
    S = 12 if cfg.model == kpax.MODEL_QUADCOPTER else 6
    for i in range(n):
        env = i % len(envs)
        bud = budget
        if budget_range is not None:
            bud = int(np.exp(rng.uniform(np.log(budget_range[0]), np.log(budget_range[1]))))
        dmax = np.inf if max_dist is None else (max_dist(bud) if callable(max_dist) else max_dist)
        for _ in range(tries):
            a, b = rng.uniform(lo, hi), rng.uniform(lo, hi)
            d = np.linalg.norm(a - b)
            if min_dist <= d <= dmax and _free(a, envs[env], clearance) and _free(b, envs[env], clearance):
                break
        else:
            raise RuntimeError(f"no free start/goal pair in environment {env} after {tries} draws")
        st = np.zeros(S, np.float32)
        st[:3] = a
        out.append(kpax.Query(st, b.astype(np.float32), seed=int(seed) * 100003 + i, budget=bud, env=env))
    return out


def serving_workload(cfg, n_queries=256, n_env=32, seed=0):
    """The batched-serving workload: n_env environments over the configured
    workspace, alternately a forest of 100-600 columns and a maze of 2,000
    boxes, and n_queries quadcopter queries with budgets log-uniform in
    [16k, 1M] nodes, start and goal 1 m apart at least and at most
    budget_distance(budget).  Returns (environments, queries)."""
    rng = np.random.default_rng(seed)
    envs = []
    for e in range(n_env):
        if e % 2 == 0:
            envs.append(kpax.generate_environment(2, 500 + e, int(rng.integers(100, 601)), cfg.ws_min, cfg.ws_max))
        else:
            envs.append(kpax.generate_environment(3, 500 + e, 2000, cfg.ws_min, cfg.ws_max))
    qs = serving_queries(cfg, envs, n_queries, seed=7, min_dist=1.0, budget_range=(16384, 1 << 20),
                         max_dist=budget_distance)
    return envs, qs


def normalise_controls(cfg, u):
    lo = np.array(cfg.ctrl_min[:4], np.float32)
    hi = np.array(cfg.ctrl_max[:4], np.float32)
    return (2.0 * (u - lo) / (hi - lo) - 1.0).astype(np.float32)


def dataset(planner, cfg, queries, results):
    """Features X (m, 80) and normalised controls Y (m, 4) of every edge on the
    returned paths: the parent row's state and the child row's control."""
    S, C = planner.S, planner.C
    st, gl, ev, ct = [], [], [], []
    for q, r in zip(queries, results):
        if r.status != kpax.OK or r.rows.shape[0] < 2:
            continue
        st.append(r.rows[:-1, :S])
        ct.append(r.rows[1:, S:S + C])
        gl.append(np.tile(np.asarray(q.goal, np.float32)[:3], (r.rows.shape[0] - 1, 1)))
        ev.append(np.full(r.rows.shape[0] - 1, q.env))
    if not st:
        return np.zeros((0, kpax.MLP_FEATURES), np.float32), np.zeros((0, 4), np.float32)
    st, gl, ev, ct = np.concatenate(st), np.concatenate(gl), np.concatenate(ev), np.concatenate(ct)
    X = np.zeros((st.shape[0], kpax.MLP_FEATURES), np.float32)
    for e in np.unique(ev):
        sel = ev == e
        X[sel] = planner.mlp_features(int(e), st[sel].T.copy(), gl[sel].T.copy())
    return X, normalise_controls(cfg, ct)


def train(X, Y, epochs=20, batch=4096, lr=2e-3, seed=0, init_seed=1):
    """Fit the network on the GPU; returns (parameters, per-epoch loss)."""
    batch = min(batch, max(32, (X.shape[0] // 32) * 32))
    with kpax.Trainer(max_batch=batch) as t:
        t.set_params(kpax.mlp_init_params(init_seed))
        losses = t.fit(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return t.params(), losses


def evaluate(planner, queries, params=None, mix=0.5, phases=True, min_budget=0):
    """Plan `queries` one at a time (the single-query form) with the plain
    sampler (params None) or the mixture (off for queries whose budget is
    below `min_budget`); success = goal reached within the budget and the FP64
    re-check passed.  With `phases`, a second pass over the same queries with
    HIP events on gives the device time per phase."""
    planner.set_proposal(params, mix if params is not None else 0.0, min_budget=min_budget)
    res = planner.plan_batch(queries[:1])  # warm-up
    t0 = time.perf_counter()
    res = [planner.plan_batch([q])[0] for q in queries]
    wall = time.perf_counter() - t0
    phase_ms = None
    if phases:
        planner.set_profiling(True)
        planner.reset_stats()
        for q in queries:
            planner.plan_batch([q])
        st = planner.stats()
        planner.set_profiling(False)
        phase_ms = {k: v / max(1, st["plans"]) for k, v in st["phase_ms"].items()}
    planner.set_proposal(None)
    ok = [r for r in res if r.status == kpax.OK and r.verified == 1]
    return {
        "success": len(ok) / len(res),
        "plans_per_sec": len(res) / wall,
        "p50_ms_found": statistics.median(r.plan_ms for r in ok) if ok else None,
        "mean_iterations_found": statistics.mean(r.iterations for r in ok) if ok else None,
        "mean_expansions": statistics.mean(r.expansions for r in res),
        "phase_ms_per_plan": phase_ms,
        "results": res,
    }
