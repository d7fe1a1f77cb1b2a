"""The built-in workloads with their start/goal and environment recipe.

Workload 1: 6D double integrator, 20 random boxes in [0,1]^3.
Workload 2: 12D quadcopter, 400-column forest in [0,10]^2 x [0,5].
Workload 3: 6D Dubins airplane, 2,000-box maze with a 0.1-wide passage.
"""
import numpy as np

import kpax

START_GOAL = {
    1: ([0.05, 0.05, 0.05, 0, 0, 0], [0.95, 0.95, 0.95]),
    2: ([0.5, 0.5, 1.0] + [0.0] * 9, [9.5, 9.5, 1.0]),
    3: ([0.1, 0.5, 0.5, 0.0, 0.0, 0.15], [0.9, 0.5, 0.5]),
}
ENV = {1: (1, 20), 2: (2, 400), 3: (3, 2000)}


def workload(k, seed=0, max_tree=None, max_iter=None, n_obs=None):
    """(config, obstacles, start, goal) of workload k with environment seed."""
    cfg = kpax.config_workload(k)
    if max_tree is not None:
        cfg.max_tree = max_tree
    if max_iter is not None:
        cfg.max_iter = max_iter
    start, goal = START_GOAL[k]
    kind, n = ENV[k]
    if n_obs is not None:
        n = n_obs
    obs = kpax.generate_environment(kind, seed, n, list(cfg.ws_min), list(cfg.ws_max), start[:3],
                                    list(goal) + [cfg.goal_radius])
    return cfg, obs, np.array(start, np.float32), np.array(goal, np.float32)


def random_states(cfg, n, rng, margin=0.05):
    """Random start states inside the workspace with small velocities/angles."""
    S, C = cfg.S, cfg.C
    lo = np.array(cfg.ws_min, np.float32)
    hi = np.array(cfg.ws_max, np.float32)
    ext = hi - lo
    x0 = np.zeros((S, n), np.float32)
    x0[:3] = (lo[:, None] + ext[:, None] * rng.uniform(margin, 1 - margin, (3, n))).astype(np.float32)
    if cfg.model == kpax.MODEL_DOUBLE_INTEGRATOR:
        x0[3:6] = rng.uniform(-0.5, 0.5, (3, n))
    elif cfg.model == kpax.MODEL_DUBINS_AIRPLANE:
        x0[3] = rng.uniform(-np.pi, np.pi, n)
        x0[4] = rng.uniform(-0.5, 0.5, n)
        x0[5] = rng.uniform(0.05, 0.25, n)
    else:
        x0[3:6] = rng.uniform(-0.4, 0.4, (3, n))
        x0[6:9] = rng.uniform(-1.0, 1.0, (3, n))
        x0[9:12] = rng.uniform(-1.0, 1.0, (3, n))
    cmin = np.array(cfg.ctrl_min[:C], np.float32)[:, None]
    cmax = np.array(cfg.ctrl_max[:C], np.float32)[:, None]
    u = (cmin + (cmax - cmin) * rng.uniform(0, 1, (C, n))).astype(np.float32)
    steps = rng.integers(cfg.steps_min, cfg.steps_max + 1, n).astype(np.int32)
    return x0.astype(np.float32), u, steps


def oracle_path_rows(o, cfg):
    """Root -> goal rows [state, control, dt * steps] of an oracle plan, in the
    result's row format (controls in the reference's storage order)."""
    j, nodes = o["goal_index"], []
    while j != -1:
        nodes.append(j)
        j = int(o["parents"][j])
    nodes.reverse()
    S, C = cfg.S, cfg.C
    order = kpax.row_control_order(cfg.model)
    rows = np.zeros((len(nodes), S + C + 1), np.float32)
    for r, j in enumerate(nodes):
        rows[r, :S] = o["states"][:, j]
        rows[r, S:S + C] = o["controls"][order, j]
        rows[r, S + C] = np.float32(cfg.dt) * np.float32(o["steps"][j])
    return rows
