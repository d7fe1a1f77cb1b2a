#!/usr/bin/env python3
"""Benchmark of the Kino-PAX planner on one MI355X.

A step is one complete plan() of the 12D quadcopter forest workload (400
columns in [0,10]^2 x [0,5], RK4 dt = 0.02, 2^20-node tree cap, 16^3 regions,
64^3 collision grid), each step with its own seed.  Inputs (obstacles, grid)
are resident in HBM before timing starts.  The timed plans run as they ship
(the default launch shape, no profiling events); a second,
profiled pass over the same seeds gives the per-phase device times and the
expansion kernel's roofline figures (the plans are deterministic, so both
passes do identical work).

Also measured, each on the same GPU and stated beside the headline:
  batched  256 queries over 32 environments (16 forests of 100-600 columns,
           16 mazes of 2,000 boxes), budgets log-uniform in [16k, 1M] nodes,
           start-goal distance from 1 m up to 2 m (16k) .. 9 m (1M) with the
           log of the budget, in one kpax_plan_batch call, against the same
           queries planned one at a time; every query's result is identical
           in both forms, so the success rates match by construction
  proposal the learned control proposal (50/50 mixture with uniform) against
           the plain sampler on held-out forests and queries, at 2^18-, 2^17-
           and 2^16-node budgets, then at budgets log-uniform in [2^15, 2^19]
           with and without the budget policy (uniform sampling below 2^17)
Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-batch] [--no-proposal] [--no-others]
                    [--cpu-seconds S] [--batched-only] [--proposal-only]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kino-pax-fork_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import kpax  # noqa: E402
import proposal  # noqa: E402
from workloads import workload  # noqa: E402

METRIC = "p50 time-to-first-solution & plans/sec at ≥95% success; node-expansions/sec"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
PARAMS = os.path.join(ROOT, "kino-pax-fork_amd", "data", "proposal_forest.npy")


# Algorithmic bytes per expansion: parent state read (4S), child sample write
# (4(S+C+1)), parent index (4), admission flag (1), two region counters (8).
def bytes_per_expansion(S, C):
    return 8 * S + 4 * C + 17


# HBM bytes per expansion that k_expand<quadcopter> moves on this workload,
# measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (average per
# dispatch over the expansions per dispatch; FETCH scaled by 2, WRITE by 1,
# the factors that make a 1 GiB device copy read and write 1 GiB): 175 B read
# (the child's 32-byte prepared record; parents read in longest-edge-first
# order, a cache line per state word; each work-group's 22 KB grid image) and
# 96 B written (the 80-byte child and its share of partial lines).
MEASURED_EXPAND_TRAFFIC_PER_EXPANSION = 175 + 96


# Algorithmic FP32 operations per integration step: the reference's step
# formula counted as written (add, sub, mul, div and negation 1, a fused
# multiply-add 2, each sin/cos/tan/sqrt call 1; compile-time constants such
# as STEP_SIZE / 6.0f and (IY - IZ) / IX folded; compares not counted):
#   double integrator  statePropagatorSpatialHash.cu:92-97: 3 x (12 + 8) = 60
#   Dubins airplane    :160-178: x 32 + y 32 + z 22 arithmetic, yaw, pitch 2 + 2,
#                      v 6, plus 20 sin/cos calls = 116
#   quadcopter         ode (statePropagator.cu:341-369) 109 arithmetic + 18 for
#                      the stage inputs + 44 calls = 171 per evaluation, 4 per
#                      step, plus the RK4 combine (statePropagatorSpatialHash.cu:
#                      262-265) 7 x 12 = 768
# The kernel itself does less arithmetic per step than the formula (shared
# sin/cos, angle-shifted stage sines, fused multiply-adds).
FLOP_PER_STEP = {kpax.MODEL_DOUBLE_INTEGRATOR: 60, kpax.MODEL_DUBINS_AIRPLANE: 116, kpax.MODEL_QUADCOPTER: 768}


MODEL_NAME = {kpax.MODEL_DOUBLE_INTEGRATOR: ("6D double integrator", "double_integrator"),
              kpax.MODEL_DUBINS_AIRPLANE: ("6D Dubins airplane", "dubins"),
              kpax.MODEL_QUADCOPTER: ("12D quadcopter", "quadcopter")}
ENV_NAME = {1: "random boxes", 2: "column forest", 3: "box maze"}


def describe(k, cfg, n_obs):
    """The workload of `bench.py --workload k` in words, from its config."""
    ext = lambda a, b: f"[{a:g},{b:g}]"  # noqa: E731
    ws = "x".join(ext(a, b) for a, b in zip(cfg.ws_min, cfg.ws_max))
    l1, l2, g = cfg.r1_len[0], cfg.r2_len[0], cfg.grid_dims[0]
    return (f"{MODEL_NAME[cfg.model][0]}, {n_obs}-obstacle {ENV_NAME[k]}, {ws}, RK4 dt={cfg.dt:g}, "
            f"2^{int(cfg.max_tree).bit_length() - 1}-node tree cap, {l1}^3 regions x {l2}^3 sub-regions, {g}^3 grid")


def dubins_environments(launch, plans=10):
    """Outcomes of the Dubins maze in four seeded mazes, at the reference's
    300-iteration limit and with the 2^21-node tree cap as the only budget."""
    out = {}
    for env in range(4):
        row = {}
        for max_iter in (300, 3000):
            cfg, obs, start, goal = workload(3, seed=env, max_iter=max_iter)
            with kpax.KPAX(cfg) as p:
                p.set_obstacles(obs)
                if launch:
                    p.set_launch(**launch)
                res = [p.plan(start, goal, seed=7000 + s) for s in range(plans)]
            row[f"max_iter_{max_iter}"] = {name: sum(r.status == code for r in res) for name, code in
                                           (("found", kpax.OK), ("iteration_limit", kpax.NOT_FOUND),
                                            ("tree_cap", kpax.BUDGET_EXHAUSTED))}
        out[f"env_seed_{env}"] = row
    return out


def host_cpu():
    """CPU model, online CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1, len(os.sched_getaffinity(0))


def cpu_baseline(k, env_seed, seeds, budget_s=20.0, min_plans=10, max_plans=50):
    """The CPU restatement of the same planner (OpenMP over expansions) on the
    host's cores: seeded full plans of the bench workload until `budget_s`
    seconds have passed and at least `min_plans` ran (at most `max_plans`).
    The box's CPU share is 16 cores per GPU, so at most 16 threads run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker only: times the CPU restatement of the same planner
    model, nproc, allowed = host_cpu()
    threads = max(1, min(allowed, 16))
    cfg, obs, start, goal = workload(k, seed=env_seed)
    exp, ms, found = 0, [], 0
    t0 = time.perf_counter()
    for s in seeds:
        t = time.perf_counter()
        r = oracle.plan(cfg, obs, start, goal, seed=s, threads=threads, want_tree=False)
        ms.append((time.perf_counter() - t) * 1e3)
        exp += r["expansions"]
        found += r["status"] == kpax.OK
        el = time.perf_counter() - t0
        if len(ms) >= max_plans or (len(ms) >= min_plans and el >= budget_s):
            break
    dt = time.perf_counter() - t0
    return {"value": exp / dt, "unit": "node-expansions/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": nproc, "cpus_allowed": allowed,
            "plans": len(ms), "success_rate": found / len(ms),
            "p50_time_to_solution_ms": statistics.median(ms), "p99_time_to_solution_ms": float(np.percentile(ms, 99)),
            "plans_per_sec": len(ms) / dt,
            "sample": f"{len(ms)} full seeded plans of the same workload (seeds {seeds[0]}..{seeds[len(ms) - 1]}), "
                      f"{dt:.1f} s, {threads} OpenMP threads over expansions on {model} "
                      f"(nproc {nproc}, {allowed} CPUs allowed); same iteration semantics and draws as the GPU "
                      f"planner, collision through the same kind of uniform grid"}


def other_workload(k, plans=20, env_seed=0, launch=None, max_iter=None):
    """Success, time-to-solution and expansion rate of workload k (single-query form)."""
    cfg, obs, start, goal = workload(k, seed=env_seed, max_iter=max_iter)
    with kpax.KPAX(cfg) as p:
        p.set_obstacles(obs)
        if launch:
            p.set_launch(**launch)
        for s in range(3):
            p.plan(start, goal, seed=5000 + s)
        res = []
        t0 = time.perf_counter()
        for s in range(plans):
            res.append(p.plan(start, goal, seed=6000 + s))
        wall = time.perf_counter() - t0
    ok = [r for r in res if r.status == kpax.OK and r.verified == 1]
    status = {name: sum(r.status == code for r in res) for name, code in
              (("found", kpax.OK), ("iteration_limit", kpax.NOT_FOUND), ("tree_cap", kpax.BUDGET_EXHAUSTED))}
    return {"workload": describe(k, cfg, obs.shape[0]), "plans": plans, "success_rate": len(ok) / plans,
            "outcomes": status,
            "p50_time_to_solution_ms": statistics.median(r.plan_ms for r in ok) if ok else None,
            "p50_plan_ms_all": statistics.median(r.plan_ms for r in res),
            "plans_per_sec": plans / wall, "node_expansions_per_sec": sum(r.expansions for r in res) / wall,
            "mean_iterations": statistics.mean(r.iterations for r in res),
            "mean_tree_size": statistics.mean(r.tree_size for r in res)}


def batched(cfg):
    envs, qs = proposal.serving_workload(cfg)
    T = sum(q.budget for q in qs) + 1024
    with kpax.KPAX(cfg, max_queries=len(qs), node_capacity=T) as p:
        for e, o in enumerate(envs):
            p.set_environment(e, o)
        p.plan_batch(qs[:4])
        p.reset_stats()
        t0 = time.perf_counter()
        res = p.plan_batch(qs)
        wall = time.perf_counter() - t0
        st = p.stats()
        launch = p.launch()
        # the same batch with the learned proposal (queries from 2^17 nodes)
        prop = None
        if os.path.exists(PARAMS):
            p.set_proposal(np.load(PARAMS), 0.5, min_budget=1 << 17)
            p.plan_batch(qs[:4])
            t0 = time.perf_counter()
            res_p = p.plan_batch(qs)
            wall_p = time.perf_counter() - t0
            p.set_proposal(None)
            ok_p = [r for r in res_p if r.status == kpax.OK and r.verified == 1]
            lat_p = [r.plan_ms for r in res_p]
            prop = {"plans_per_sec": len(qs) / wall_p, "wall_s": wall_p, "success_rate": len(ok_p) / len(qs),
                    "p50_latency_ms": statistics.median(lat_p), "p99_latency_ms": float(np.percentile(lat_p, 99)),
                    "queries_with_proposal": sum(q.budget >= (1 << 17) for q in qs),
                    "note": "mixture 0.5 for queries with budgets >= 2^17 (forest-trained network; half of the "
                            "environments are mazes)"}
    with kpax.KPAX(cfg, max_queries=1, node_capacity=1 << 20) as s:
        for e, o in enumerate(envs):
            s.set_environment(e, o)
        s.plan_batch(qs[:2])
        t0 = time.perf_counter()
        seq = [s.plan_batch([q])[0] for q in qs]
        wall_seq = time.perf_counter() - t0
    same = all((a.status, a.tree_size, a.expansions, a.iterations) == (b.status, b.tree_size, b.expansions,
                                                                       b.iterations) for a, b in zip(res, seq))
    ok = [r for r in res if r.status == kpax.OK and r.verified == 1]
    lat = [r.plan_ms for r in res]
    return {"queries": len(qs), "environments": len(envs), "budget_nodes_total": T - 1024,
            "plans_per_sec": len(qs) / wall, "wall_s": wall, "success_rate": len(ok) / len(qs),
            "p50_latency_ms": statistics.median(lat), "p99_latency_ms": float(np.percentile(lat, 99)),
            "node_expansions_per_sec": sum(r.expansions for r in res) / wall,
            "iterations": st["iterations"], "graph_launches": st["graph_launches"], "launch": launch,
            "single_query_form": {"plans_per_sec": len(qs) / wall_seq, "wall_s": wall_seq,
                                  "p50_time_to_solution_ms": statistics.median(r.plan_ms for r in seq),
                                  "success_rate": sum(r.status == kpax.OK and r.verified == 1 for r in seq) / len(seq)},
            "batched_over_single_plans_per_sec": (len(qs) / wall) / (len(qs) / wall_seq),
            "results_identical_to_single_query_form": same,
            "with_learned_proposal": prop}


def learned_proposal(cfg):
    if not os.path.exists(PARAMS):
        return {"skipped": f"{PARAMS} not found"}
    params = np.load(PARAMS)
    strip = lambda d: {k: v for k, v in d.items() if k != "results"}  # noqa: E731
    out = {"workload": "96 held-out queries (start/goal >= 4 m apart) over 8 held-out forests of 100-600 columns, "
                       "max_iter 100, single-query form; parameters trained by train_proposal.py on the 2.0M "
                       "path edges of 180,000 plain-planner runs in other forests; per budget, then budgets "
                       "log-uniform in "
                       "[2^15, 2^19] with and without the budget policy (uniform sampling below 2^17)"}
    for budget in (1 << 18, 1 << 17, 1 << 16):
        c = cfg.copy()
        c.max_tree = budget
        c.max_iter = 100
        envs = proposal.forest_environments(c, 8, seed=2)
        qs = proposal.serving_queries(c, envs, 96, seed=4)
        with kpax.KPAX(c) as p:
            for e, o in enumerate(envs):
                p.set_environment(e, o)
            plain = proposal.evaluate(p, qs, None)
            learned = proposal.evaluate(p, qs, params, 0.5)
        both = [(a, b) for a, b in zip(plain["results"], learned["results"]) if a.status == 0 and b.status == 0]
        out[f"budget_{budget}"] = {
            "plain": strip(plain), "learned_mix_0.5": strip(learned),
            "on_queries_both_solve": {
                "n": len(both),
                "mean_iterations_plain": float(np.mean([a.iterations for a, _ in both])) if both else None,
                "mean_iterations_learned": float(np.mean([b.iterations for _, b in both])) if both else None,
                "mean_expansions_plain": float(np.mean([a.expansions for a, _ in both])) if both else None,
                "mean_expansions_learned": float(np.mean([b.expansions for _, b in both])) if both else None}}
    # budgets log-uniform in [2^15, 2^19]: the mixture for every query, and
    # with the budget policy (uniform sampling below 2^17)
    c = cfg.copy()
    c.max_tree = 1 << 19
    c.max_iter = 100
    envs = proposal.forest_environments(c, 8, seed=2)
    qs = proposal.serving_queries(c, envs, 96, seed=5, budget_range=(1 << 15, 1 << 19))
    with kpax.KPAX(c) as p:
        for e, o in enumerate(envs):
            p.set_environment(e, o)
        out["mixed_budgets_32k_512k"] = {
            "plain": strip(proposal.evaluate(p, qs, None)),
            "learned_mix_0.5": strip(proposal.evaluate(p, qs, params, 0.5)),
            "learned_mix_0.5_from_128k": strip(proposal.evaluate(p, qs, params, 0.5, min_budget=1 << 17))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--env-seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU baseline: seeded plans until this many seconds (at least 10 plans)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true", help="skip the double-integrator and Dubins sections")
    ap.add_argument("--launch", default="", metavar="K=V,...",
                    help="launch-shape overrides for the single-query planners (kpax_launch fields)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-proposal", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched-serving section (for profiling it alone); prints its JSON object")
    ap.add_argument("--proposal-only", action="store_true",
                    help="only the learned-proposal section (for profiling it alone); prints its JSON object")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates the byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch  # device memory plumbing for the calibration copy only
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (each GPU box here has one MI355X); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    cfg, obs, start, goal = workload(args.workload, seed=args.env_seed)
    if args.batched_only:
        print(json.dumps({"batched": batched(cfg)}), flush=True)
        return
    if args.proposal_only:
        print(json.dumps({"learned_proposal": learned_proposal(cfg)}), flush=True)
        return
    planner = kpax.KPAX(cfg)
    planner.set_obstacles(obs)
    overrides = {k: int(v) for k, v in (kv.split("=") for kv in args.launch.split(",") if kv)}
    if overrides:
        planner.set_launch(**overrides)
    seeds = [1000 + i for i in range(args.warmup + args.steps)]
    for s in seeds[: args.warmup]:
        planner.plan(start, goal, seed=s)

    # timed: the plans as they ship
    planner.reset_stats()
    results = []
    t0 = time.perf_counter()
    for s in seeds[args.warmup:]:
        results.append(planner.plan(start, goal, seed=s))
    wall = time.perf_counter() - t0
    graph_launches = planner.stats()["graph_launches"]
    # profiled pass over the same seeds: per-phase device time from HIP events
    planner.set_profiling(True)
    planner.reset_stats()
    for s in seeds[args.warmup:]:
        planner.plan(start, goal, seed=s)
    st = planner.stats()
    # time-to-solution with the environment's grid built inside the timed call
    # (the reference rebuilds its grid inside plan(), KPAX.cu:306)
    planner.set_profiling(False)
    with_grid = []
    for s in seeds[args.warmup:args.warmup + 10]:
        t = time.perf_counter()
        planner.set_obstacles(obs)
        r = planner.plan(start, goal, seed=s)
        with_grid.append((time.perf_counter() - t) * 1e3)
    planner.close()

    ok = [r for r in results if r.status == kpax.OK and r.verified == 1]
    success = len(ok) / len(results)
    expansions = sum(r.expansions for r in results)
    steps_int = sum(r.integrated_steps for r in results)
    p50 = statistics.median(r.plan_ms for r in ok) if ok else None
    p99 = float(np.percentile([r.plan_ms for r in ok], 99)) if ok else None

    S, C = cfg.S, cfg.C
    exp_ms = st["phase_ms"]["expand"]
    launches = max(1, st["phase_launches"]["expand"])
    per_launch_exp = st["expansions"] / launches
    per_launch_ms = exp_ms / launches
    achieved_gbs = per_launch_exp * bytes_per_expansion(S, C) / (per_launch_ms * 1e-3) / 1e9
    flops = st["integrated_steps"] * FLOP_PER_STEP[cfg.model]
    achieved_tflops = flops / (exp_ms * 1e-3) / 1e12
    kernel = f"k_expand<{MODEL_NAME[cfg.model][1]}>"
    compute = {"bound": "valu", "achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": achieved_tflops / FP32_PEAK_TFLOPS,
               "traffic": (per_launch_exp * MEASURED_EXPAND_TRAFFIC_PER_EXPANSION if args.workload == 2 else None),
               "kernel": kernel, "ceiling": "FP32 vector (VALU) issue: MI355X's published fp32 157.3 TFLOP/s",
               "flop_per_step": FLOP_PER_STEP[cfg.model],
               "flop_convention": "reference step formula as written: add/sub/mul/div/neg 1, fma 2, "
                                  "sin/cos/tan/sqrt 1",
               "flop_per_launch": flops / launches, "avg_launch_ms": per_launch_ms}
    memory = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": achieved_gbs / HBM_PEAK_GBS,
              "traffic": (per_launch_exp * MEASURED_EXPAND_TRAFFIC_PER_EXPANSION if args.workload == 2 else None),
              "traffic_basis": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE bytes per expansion x expansions "
                               "per launch (measured on workload 2)",
              "kernel": kernel, "bytes_per_expansion": bytes_per_expansion(S, C),
              "expansions_per_launch": per_launch_exp, "avg_launch_ms": per_launch_ms}
    # the quadcopter forest's binding ceiling is FP32 issue (about 80 FLOP per
    # algorithmic byte); the double integrator's is bytes
    valu_bound = cfg.model != kpax.MODEL_DOUBLE_INTEGRATOR

    line = {
        "metric": METRIC,
        "value": expansions / wall,
        "unit": "node-expansions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": describe(args.workload, cfg, obs.shape[0]),
                   "plans_per_step": 1, "env_seed": args.env_seed, "max_tree": int(cfg.max_tree),
                   "max_iter": int(cfg.max_iter), "start": [float(x) for x in start], "goal": [float(x) for x in goal],
                   "goal_radius": float(cfg.goal_radius), "steps_per_edge": [int(cfg.steps_min), int(cfg.steps_max)],
                   "velocity_box": [float(cfg.vel_min), float(cfg.vel_max)],
                   "controls": [[float(a), float(b)] for a, b in zip(cfg.ctrl_min[:C], cfg.ctrl_max[:C])]},
        "p50_time_to_first_solution_ms": p50,
        "p99_time_to_first_solution_ms": p99,
        "p50_time_to_first_solution_ms_incl_grid_build": statistics.median(with_grid),
        "plans_per_sec": len(results) / wall,
        "success_rate": success,
        "mean_iterations": statistics.mean(r.iterations for r in results),
        "expansions_per_plan": expansions / len(results),
        "integrated_steps": steps_int,
        "graph_launches": graph_launches,
        "phase_ms_per_plan": {k: v / max(1, st["plans"]) for k, v in st["phase_ms"].items()},
        "roofline": compute if valu_bound else memory,
        "roofline_hbm" if valu_bound else "roofline_valu": memory if valu_bound else compute,
    }
    if not args.no_others:
        line["other_workloads"] = {"double_integrator": other_workload(1, launch=overrides),
                                   "dubins_maze": other_workload(3, launch=overrides),
                                   "dubins_maze_by_environment": dubins_environments(overrides)}
    if not args.no_batch:
        line["batched"] = batched(cfg)
    if not args.no_proposal:
        line["learned_proposal"] = learned_proposal(cfg)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.env_seed, [2000 + i for i in range(50)],
                                            budget_s=args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
