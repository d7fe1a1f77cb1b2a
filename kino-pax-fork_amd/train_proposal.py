#!/usr/bin/env python3
"""Train the learned control proposal for the quadcopter forest workload.

    python kino-pax-fork_amd/train_proposal.py [--runs 180000] [--epochs 20] [--batch 4096] [--lr 1e-3]
                                               [--out data/proposal_forest.npy]

1. Draws `runs` start/goal pairs (at least 4 m apart) in 32 seeded forests of
   100-600 columns over [0,10]^2 x [0,5].
2. Solves them with the plain planner (uniform controls), 256 queries per
   batched call, tree budget 2^18, 100 iterations.
3. Takes every edge of every returned path as a training example: features of
   the parent node and the edge's control, normalised to [-1, 1].  The
   default 180,000 runs give about 2.0 million edges (11 per path).
4. Fits the network with the GPU trainer (Adam, Gaussian negative
   log-likelihood) and saves the fp32 parameters (flat layout of
   include/kpax.h) with numpy.

The held-out evaluation (bench.py's `learned_proposal`) uses other forests
(seed 2) and other queries (seed 4).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import kpax  # noqa: E402
import proposal  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=180000)
    ap.add_argument("--envs", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "data", "proposal_forest.npy"))
    args = ap.parse_args()

    cfg = kpax.config_workload(2)
    cfg.max_iter = 100
    envs = proposal.forest_environments(cfg, args.envs, seed=1)
    qs = proposal.serving_queries(cfg, envs, args.runs, seed=3, budget=1 << 18)
    t0 = time.perf_counter()
    res = []
    with kpax.KPAX(cfg, max_queries=256, node_capacity=256 << 18) as p:
        for e, o in enumerate(envs):
            p.set_environment(e, o)
        for b in range(0, len(qs), 256):
            res += p.plan_batch(qs[b:b + 256])
            if (b // 256) % 40 == 0 or len(res) == len(qs):
                print(f"planned {len(res)}/{len(qs)} ({time.perf_counter() - t0:.1f} s)", flush=True)
        X, Y = proposal.dataset(p, cfg, qs, res)
    t_data = time.perf_counter() - t0
    print(f"{X.shape[0]} edges from {sum(r.status == kpax.OK for r in res)} solved runs", flush=True)
    t0 = time.perf_counter()
    params, losses = proposal.train(X, Y, epochs=args.epochs, batch=args.batch, lr=args.lr)
    t_train = time.perf_counter() - t0
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params)
    print(json.dumps({"runs": len(qs), "edges": int(X.shape[0]), "data_s": t_data, "train_s": t_train,
                      "epochs": args.epochs, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                      "out": args.out}), flush=True)


if __name__ == "__main__":
    main()
