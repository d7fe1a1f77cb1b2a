"""Python host mirror of the Kino-PAX planner over the libkpax.so C ABI.

The class `KPAX` keeps the reference's planner interface
(include/planners/KPAX.cuh:14: plan(h_initial, h_goal, d_obstacles, n)) and
exposes the hot-path operators the C ABI exports.  All arrays are numpy;
state/control batches are SoA (dimension-major, shape (S, n)).

The library is loaded from this directory and nothing falls back to another
implementation: if libkpax.so is missing or a call fails, an exception says so.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KPAX_LIB selects another build of the same library (for instance the CPU
# build of the kernels that the test suite runs under AddressSanitizer).
LIB_PATH = os.environ.get("KPAX_LIB", os.path.join(HERE, "libkpax.so"))

STATUS = {0: "OK", 1: "NOT_FOUND", 2: "BUDGET_EXHAUSTED", -1: "INVALID_ARG", -2: "DEVICE_ERROR",
          -3: "NO_MEMORY", -4: "NO_DEVICE", -5: "IO_ERROR"}
OK, NOT_FOUND, BUDGET_EXHAUSTED, INVALID_ARG = 0, 1, 2, -1
MODEL_DOUBLE_INTEGRATOR, MODEL_DUBINS_AIRPLANE, MODEL_QUADCOPTER = 1, 2, 3
N_PHASES = 7
PHASES = ("expand", "append", "score", "frontier", "init", "propose", "prepare")


class Config(ctypes.Structure):
    """kpax_config (include/kpax.h)."""
    _fields_ = [
        ("model", ctypes.c_int32), ("di_textbook", ctypes.c_int32), ("dt", ctypes.c_float),
        ("steps_min", ctypes.c_int32), ("steps_max", ctypes.c_int32),
        ("ws_min", ctypes.c_float * 3), ("ws_max", ctypes.c_float * 3),
        ("ctrl_min", ctypes.c_float * 4), ("ctrl_max", ctypes.c_float * 4),
        ("vel_min", ctypes.c_float), ("vel_max", ctypes.c_float),
        ("pitch_min", ctypes.c_float), ("pitch_max", ctypes.c_float),
        ("mass", ctypes.c_float), ("inertia", ctypes.c_float * 3), ("nu", ctypes.c_float),
        ("mu", ctypes.c_float), ("gravity", ctypes.c_float),
        ("r1_len", ctypes.c_int32 * 3), ("r2_len", ctypes.c_int32 * 3), ("grid_dims", ctypes.c_int32 * 3),
        ("max_tree", ctypes.c_uint32), ("max_iter", ctypes.c_uint32), ("goal_radius", ctypes.c_float),
        ("fanout", ctypes.c_uint32), ("root_repeat", ctypes.c_uint32), ("sparse_repeat", ctypes.c_uint32),
        ("dense_repeat", ctypes.c_uint32), ("sparse_valid_below", ctypes.c_int32), ("eps", ctypes.c_float),
        ("verify", ctypes.c_int32), ("device", ctypes.c_int32),
        ("prune", ctypes.c_int32), ("prune_exploration", ctypes.c_float), ("prune_goal_bias", ctypes.c_float),
        ("prune_progress_scale", ctypes.c_float),
    ]

    def copy(self):
        c = Config()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(Config))
        return c

    @property
    def S(self):
        return 12 if self.model == MODEL_QUADCOPTER else 6

    @property
    def C(self):
        return 4 if self.model == MODEL_QUADCOPTER else 3

    @property
    def r1_count(self):
        return int(np.prod(list(self.r1_len)))

    @property
    def r2_per(self):
        return int(np.prod(list(self.r2_len)))


class _Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("n_rows", ctypes.c_uint32), ("state_dim", ctypes.c_uint32),
        ("control_dim", ctypes.c_uint32), ("rows", ctypes.POINTER(ctypes.c_float)),
        ("steps", ctypes.POINTER(ctypes.c_int32)), ("node_index", ctypes.POINTER(ctypes.c_int32)),
        ("goal_index", ctypes.c_int32), ("iterations", ctypes.c_uint32), ("tree_size", ctypes.c_uint32),
        ("expansions", ctypes.c_uint64), ("integrated_steps", ctypes.c_uint64), ("plan_ms", ctypes.c_double),
        ("verified", ctypes.c_int32),
        ("verify_max_dev", ctypes.c_float),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [("phase_ms", ctypes.c_double * N_PHASES), ("phase_launches", ctypes.c_uint64 * N_PHASES),
                ("expansions", ctypes.c_uint64), ("integrated_steps", ctypes.c_uint64), ("plans", ctypes.c_uint64),
                ("iterations", ctypes.c_uint64), ("graph_launches", ctypes.c_uint64)]


class _Launch(ctypes.Structure):
    _fields_ = [("expand_blocks", ctypes.c_uint32), ("expand_chunk", ctypes.c_uint32),
                ("scan_blocks", ctypes.c_uint32), ("use_graph", ctypes.c_int32), ("order_children", ctypes.c_int32),
                ("refill_lanes", ctypes.c_uint32), ("scan_tile", ctypes.c_uint32), ("lds_grid", ctypes.c_int32),
                ("expand_threads", ctypes.c_uint32), ("wave_priority", ctypes.c_int32)]


class _Query(ctypes.Structure):
    _fields_ = [("start", ctypes.c_void_p), ("goal", ctypes.c_void_p), ("seed", ctypes.c_uint64),
                ("budget", ctypes.c_uint32), ("env", ctypes.c_uint32)]


class KpaxError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libkpax.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KpaxError(f"{LIB_PATH} is missing: build it with kino-pax-fork_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32, i32, u64 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
    sig = {
        "kpax_config_workload": (i32, [i32, P]),
        "kpax_model_dims": (i32, [i32, P, P]),
        "kpax_create": (i32, [P, P]),
        "kpax_create_batch": (i32, [P, u32, u64, P]),
        "kpax_set_environment": (i32, [P, u32, P, u32]),
        "kpax_plan_batch": (i32, [P, P, u32, P]),
        "kpax_plan_dump": (i32, [P, P, P, u64, ctypes.c_char_p, i32, P]),
        "kpax_mlp_init_params": (i32, [u64, P]),
        "kpax_set_proposal": (i32, [P, P, ctypes.c_float]),
        "kpax_mlp_features": (i32, [P, u32, P, P, u32, P]),
        "kpax_mlp_infer": (i32, [P, P, P, u32, P]),
        "kpax_propose_sample": (i32, [P, P, P, u32, P]),
        "kpax_trainer_create": (i32, [i32, u32, P]),
        "kpax_trainer_destroy": (i32, [P]),
        "kpax_trainer_last_error": (ctypes.c_char_p, [P]),
        "kpax_trainer_set_params": (i32, [P, P]),
        "kpax_trainer_get_params": (i32, [P, P]),
        "kpax_trainer_gradients": (i32, [P, P, P, u32, P, P]),
        "kpax_trainer_fit": (i32, [P, P, P, u32, u32, u32, ctypes.c_float, u64, P]),
        "kpax_get_launch": (i32, [P, P]),
        "kpax_set_launch": (i32, [P, P]),
        "kpax_autotune": (i32, [P, P, u32, u32, P]),
        "kpax_destroy": (i32, [P]),
        "kpax_last_error": (ctypes.c_char_p, [P]),
        "kpax_set_obstacles": (i32, [P, P, u32]),
        "kpax_plan": (i32, [P, P, P, u64, P]),
        "kpax_result_free": (None, [P]),
        "kpax_get_tree": (i32, [P, u32, P, P, P, P, P, P]),
        "kpax_get_regions": (i32, [P, P, P, P, P]),
        "kpax_get_grid": (i32, [P, P, P, u32, P, P]),
        "kpax_get_lds_grid": (i32, [P, u32, P, P]),
        "kpax_get_proposal": (i32, [P, u32, P, P]),
        "kpax_repeat_expand_noalloc": (i32, [P, P, u32, P, u32, P]),
        "kpax_set_profiling": (i32, [P, i32]),
        "kpax_get_stats": (i32, [P, P]),
        "kpax_reset_stats": (i32, [P]),
        "kpax_propagate_batch": (i32, [P, P, P, P, u32, P, P, P]),
        "kpax_collide_batch": (i32, [P, P, P, u32, P]),
        "kpax_region_batch": (i32, [P, P, u32, P, P]),
        "kpax_region_histogram": (i32, [P, P, P, u32, P, P, P]),
        "kpax_scores": (i32, [P, P, P, P, P]),
        "kpax_repeat_expand": (i32, [P, P, u32, P, u32, P]),
        "kpax_read_obstacles_csv": (i32, [ctypes.c_char_p, P, u32, P]),
        "kpax_generate_environment": (i32, [i32, u64, u32, P, P, P, P, P, u32, P]),
        "kpax_verify_path": (i32, [P, P, u32, P, P, u32, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def config_workload(k):
    """Built-in workload 1 (double integrator), 2 (quadcopter) or 3 (Dubins)."""
    c = Config()
    rc = lib().kpax_config_workload(int(k), ctypes.byref(c))
    if rc != OK:
        raise KpaxError(f"unknown workload {k}")
    return c


def model_dims(model):
    s, c = ctypes.c_int32(), ctypes.c_int32()
    if lib().kpax_model_dims(int(model), ctypes.byref(s), ctypes.byref(c)) != OK:
        raise KpaxError(f"unknown model {model}")
    return s.value, c.value


def read_obstacles_csv(path):
    """readObstaclesFromCSV (src/helper/helper.cu:13-38) -> (n, 6) float32."""
    n = ctypes.c_uint32()
    rc = lib().kpax_read_obstacles_csv(str(path).encode(), None, 0, ctypes.byref(n))
    if rc != OK:
        raise KpaxError(f"cannot read obstacles from {path}: {STATUS.get(rc, rc)}")
    out = np.zeros((n.value, 6), np.float32)
    rc = lib().kpax_read_obstacles_csv(str(path).encode(), _p(out), n.value, ctypes.byref(n))
    if rc != OK:
        raise KpaxError(f"cannot read obstacles from {path}: {STATUS.get(rc, rc)}")
    return out


def generate_environment(kind, seed, n, ws_min, ws_max, start=None, goal=None):
    """Seeded obstacle sets: 1 random boxes, 2 forest, 3 maze (see include/kpax.h)."""
    wmin, wmax = _f32(ws_min), _f32(ws_max)
    st = None if start is None else _f32(start)
    gl = None if goal is None else _f32(goal)
    cnt = ctypes.c_uint32()
    args = (int(kind), int(seed), int(n), _p(wmin), _p(wmax), _p(st), _p(gl))
    rc = lib().kpax_generate_environment(*args, None, 0, ctypes.byref(cnt))
    if rc != OK:
        raise KpaxError(f"generate_environment failed: {STATUS.get(rc, rc)}")
    out = np.zeros((cnt.value, 6), np.float32)
    rc = lib().kpax_generate_environment(*args, _p(out), cnt.value, ctypes.byref(cnt))
    if rc != OK:
        raise KpaxError(f"generate_environment failed: {STATUS.get(rc, rc)}")
    return out


def verify_path(cfg, obstacles, rows, steps, goal):
    """Host FP64 re-propagation and re-collision-check of a trajectory."""
    obs = _f32(obstacles).reshape(-1, 6)
    rows = _f32(rows)
    steps = np.ascontiguousarray(steps, np.int32)
    g = _f32(goal)
    dev = ctypes.c_float()
    rc = lib().kpax_verify_path(ctypes.byref(cfg), _p(obs), obs.shape[0], _p(rows), _p(steps), steps.shape[0],
                                _p(g), ctypes.byref(dev))
    if rc not in (OK, NOT_FOUND):
        raise KpaxError(f"verify_path: {STATUS.get(rc, rc)}")
    return rc == OK, float(dev.value)


@dataclass
class PlanResult:
    """What KPAX::plan leaves in h_pathToGoal_ / h_controlPathToGoal_ / h_treeSize_ / h_itr_
    (include/planners/Planner.cuh:31-34), plus timing and verification."""
    status: int
    rows: np.ndarray          # (n_rows, S + C + 1) root -> goal
    steps: np.ndarray         # (n_rows,) int32
    node_index: np.ndarray    # (n_rows,) tree index of each row
    goal_index: int
    iterations: int
    tree_size: int
    expansions: int
    integrated_steps: int
    plan_ms: float
    verified: int
    verify_max_dev: float
    state_dim: int = 0
    control_dim: int = 0

    @property
    def found(self):
        return self.status == OK

    @property
    def path_to_goal(self):
        """h_pathToGoal_: tree index of the goal node, 0 when none (the reference's convention)."""
        return self.goal_index if self.goal_index > 0 else 0

    @property
    def control_path_to_goal(self):
        """h_controlPathToGoal_ rows, goal -> root, each [state, control, duration]."""
        return self.rows[::-1].copy()


def row_control_order(model):
    """Planner control index held by each control column of a result row.
    Rows keep the reference's storage order, (yawRate, pitchRate, a) for the
    Dubins airplane (statePropagatorSpatialHash.cu:223-225); configs and
    kpax_get_tree keep its draw order (a, yawRate, pitchRate, :141-143)."""
    return [1, 2, 0] if model == MODEL_DUBINS_AIRPLANE else list(range(model_dims(model)[1]))


def _plan_result(r, S, C):
    D = S + C + 1
    n = r.n_rows
    rows = np.ctypeslib.as_array(r.rows, (n * D,)).reshape(n, D).copy() if n else np.zeros((0, D), np.float32)
    steps = np.ctypeslib.as_array(r.steps, (n,)).copy() if n else np.zeros(0, np.int32)
    idx = np.ctypeslib.as_array(r.node_index, (n,)).copy() if n else np.zeros(0, np.int32)
    return PlanResult(r.status, rows, steps, idx, r.goal_index, r.iterations, r.tree_size, r.expansions,
                      r.integrated_steps, r.plan_ms, r.verified, r.verify_max_dev, S, C)


MLP_FEATURES, MLP_OUTPUTS, MLP_PARAMS = 80, 8, 27912


def mlp_init_params(seed=1):
    """Initial network parameters (flat fp32, layout of include/kpax.h)."""
    p = np.zeros(MLP_PARAMS, np.float32)
    if lib().kpax_mlp_init_params(int(seed), _p(p)) != OK:
        raise KpaxError("mlp_init_params failed")
    return p


class Trainer:
    """The proposal network's forward/backward passes and Adam on the GPU."""

    def __init__(self, max_batch=8192, device=0):
        h = ctypes.c_void_p()
        rc = lib().kpax_trainer_create(int(device), int(max_batch), ctypes.byref(h))
        if rc != OK:
            raise KpaxError(f"kpax_trainer_create failed: {STATUS.get(rc, rc)}")
        self._t = h

    def close(self):
        if getattr(self, "_t", None):
            lib().kpax_trainer_destroy(self._t)
            self._t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != OK:
            msg = lib().kpax_trainer_last_error(self._t)
            raise KpaxError(f"{what}: {STATUS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def set_params(self, params):
        p = _f32(params, (MLP_PARAMS,))
        self._check(lib().kpax_trainer_set_params(self._t, _p(p)), "set_params")

    def params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        self._check(lib().kpax_trainer_get_params(self._t, _p(p)), "get_params")
        return p

    def gradients(self, X, Y):
        X = _f32(X).reshape(-1, MLP_FEATURES)
        Y = _f32(Y).reshape(-1, 4)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        self._check(lib().kpax_trainer_gradients(self._t, _p(X), _p(Y), X.shape[0], _p(g), ctypes.byref(loss)),
                    "gradients")
        return float(loss.value), g

    def fit(self, X, Y, epochs=10, batch=4096, lr=1e-3, seed=0):
        X = _f32(X).reshape(-1, MLP_FEATURES)
        Y = _f32(Y).reshape(-1, 4)
        losses = np.zeros(epochs, np.float32)
        self._check(lib().kpax_trainer_fit(self._t, _p(X), _p(Y), X.shape[0], int(epochs), int(batch),
                                           float(lr), int(seed), _p(losses)), "fit")
        return losses


@dataclass
class Query:
    """One query of plan_batch: start state, goal centre, seed, tree budget (0 = config max_tree), environment."""
    start: np.ndarray
    goal: np.ndarray
    seed: int = 0
    budget: int = 0
    env: int = 0


class KPAX:
    """MI355X Kino-PAX planner (mirrors the reference's KPAX class).

    KPAX(config) plans one query at a time; KPAX(config, max_queries=Q,
    node_capacity=T) holds Q query slots over a pool of T tree nodes for
    plan_batch."""

    def __init__(self, config=2, max_queries=None, node_capacity=None):
        cfg = config_workload(config) if isinstance(config, int) else config.copy()
        self.config = cfg
        self.S, self.C = model_dims(cfg.model)
        h = ctypes.c_void_p()
        if max_queries is None and node_capacity is None:
            rc = lib().kpax_create(ctypes.byref(cfg), ctypes.byref(h))
        else:
            q = int(max_queries or 1)
            t = int(node_capacity or cfg.max_tree * q)
            rc = lib().kpax_create_batch(ctypes.byref(cfg), q, t, ctypes.byref(h))
        if rc != OK:
            raise KpaxError(f"kpax_create failed: {STATUS.get(rc, rc)}")
        self._h = h
        self._obs = None
        self._obs_src = None

    def close(self):
        if getattr(self, "_h", None):
            lib().kpax_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != OK:
            msg = lib().kpax_last_error(self._h)
            raise KpaxError(f"{what}: {STATUS.get(rc, rc)}: {msg.decode() if msg else ''}")

    # ---- planner ----------------------------------------------------------
    def set_obstacles(self, obstacles):
        obs = _f32(obstacles).reshape(-1, 6)
        self._check(lib().kpax_set_obstacles(self._h, _p(obs), obs.shape[0]), "set_obstacles")
        self._obs = obs

    def plan(self, initial, goal, obstacles=None, seed=0):
        """plan(h_initial, h_goal, obstacles) -> PlanResult (KPAX.cu:51-115).

        `initial` holds at least S floats (the reference passes SAMPLE_DIM);
        `goal` at least the xyz centre of the goal ball."""
        if obstacles is not None and (self._obs is None or obstacles is not self._obs_src):
            self.set_obstacles(obstacles)
            self._obs_src = obstacles
        if self._obs is None:
            self.set_obstacles(np.zeros((0, 6), np.float32))
        x0 = _f32(np.asarray(initial, np.float32).ravel()[: self.S])
        if x0.size < self.S:
            raise KpaxError(f"initial state needs {self.S} values")
        g = _f32(np.asarray(goal, np.float32).ravel()[:3])
        r = _Result()
        self._check(lib().kpax_plan(self._h, _p(x0), _p(g), int(seed) & (2**64 - 1), ctypes.byref(r)), "plan")
        out = _plan_result(r, self.S, self.C)
        lib().kpax_result_free(ctypes.byref(r))
        return out

    def plan_dump(self, initial, goal, directory, seed=0, bench_itr=0):
        """plan() plus KPAX::planBench's per-iteration CSV dumps under `directory`
        (the reference's Data/ layout, read by its MATLAB visualisers)."""
        if self._obs is None:
            self.set_obstacles(np.zeros((0, 6), np.float32))
        x0 = _f32(np.asarray(initial, np.float32).ravel()[: self.S])
        g = _f32(np.asarray(goal, np.float32).ravel()[:3])
        r = _Result()
        self._check(lib().kpax_plan_dump(self._h, _p(x0), _p(g), int(seed) & (2**64 - 1), str(directory).encode(),
                                         int(bench_itr), ctypes.byref(r)), "plan_dump")
        out = _plan_result(r, self.S, self.C)
        lib().kpax_result_free(ctypes.byref(r))
        return out

    def set_environment(self, env, obstacles):
        """Obstacles of environment `env` for batched queries (environment 0 = set_obstacles)."""
        obs = _f32(obstacles).reshape(-1, 6)
        self._check(lib().kpax_set_environment(self._h, int(env), _p(obs), obs.shape[0]), "set_environment")
        if env == 0:
            self._obs = obs

    def _queries(self, queries):
        keep = []
        arr = (_Query * max(len(queries), 1))()
        for i, q in enumerate(queries):
            x0 = _f32(np.asarray(q.start, np.float32).ravel()[: self.S])
            g = _f32(np.asarray(q.goal, np.float32).ravel()[:3])
            if x0.shape[0] != self.S or g.shape[0] != 3:
                raise ValueError(f"query {i}: start needs {self.S} values and goal 3, got {x0.shape[0]} and "
                                 f"{g.shape[0]}")
            keep += [x0, g]
            arr[i] = _Query(x0.ctypes.data, g.ctypes.data, int(q.seed) & (2**64 - 1), int(q.budget), int(q.env))
        return arr, keep

    def plan_batch(self, queries):
        """Plan a list of Query on the handle's slots -> list of PlanResult (same order)."""
        n = len(queries)
        arr, keep = self._queries(queries)
        res = (_Result * max(n, 1))()
        self._check(lib().kpax_plan_batch(self._h, arr, n, res), "plan_batch")
        out = []
        for i in range(n):
            out.append(_plan_result(res[i], self.S, self.C))
            lib().kpax_result_free(ctypes.byref(res[i]))
        return out

    # ---- learned control proposal ---------------------------------------------
    def set_proposal(self, params, mix=0.5, min_budget=0):
        """Draw each child's control from its parent's learned Gaussian with
        probability `mix` (params None or mix 0: the plain uniform sampler);
        queries with a tree budget below `min_budget` sample uniformly."""
        p = None if params is None else _f32(params, (MLP_PARAMS,))
        self._check(lib().kpax_set_proposal(self._h, _p(p), float(mix if p is not None else 0.0)), "set_proposal")
        self._check(lib().kpax_set_proposal_min_budget(self._h, int(min_budget)), "set_proposal_min_budget")

    def mlp_features(self, env, states, goals):
        """Features (n, 80) of states (S, n) with goals (3, n) in environment `env`."""
        st = _f32(states).reshape(self.S, -1)
        n = st.shape[1]
        g = _f32(goals).reshape(3, n)
        out = np.zeros((n, MLP_FEATURES), np.float32)
        self._check(lib().kpax_mlp_features(self._h, int(env), _p(st), _p(g), n, _p(out)), "mlp_features")
        return out

    def mlp_infer(self, params, X):
        X = _f32(X).reshape(-1, MLP_FEATURES)
        out = np.zeros((X.shape[0], MLP_OUTPUTS), np.float32)
        self._check(lib().kpax_mlp_infer(self._h, _p(_f32(params, (MLP_PARAMS,))), _p(X), X.shape[0], _p(out)),
                    "mlp_infer")
        return out

    def propose_sample(self, mean_sigma, words):
        ms = _f32(mean_sigma).reshape(-1, 8)
        w = np.ascontiguousarray(words, np.uint32).reshape(-1, 4)
        out = np.zeros((ms.shape[0], 4), np.float32)
        self._check(lib().kpax_propose_sample(self._h, _p(ms), _p(w), ms.shape[0], _p(out)), "propose_sample")
        return out

    def launch(self):
        """Launch shape: expand_blocks, expand_chunk, scan_blocks, use_graph, order_children, refill_lanes, scan_tile,
        lds_grid, expand_threads, wave_priority."""
        l = _Launch()
        self._check(lib().kpax_get_launch(self._h, ctypes.byref(l)), "get_launch")
        return {k: getattr(l, k) for k, _ in _Launch._fields_}

    def set_launch(self, **kw):
        cur = self.launch()
        cur.update(kw)
        l = _Launch(*(int(cur[k]) for k, _ in _Launch._fields_))
        self._check(lib().kpax_set_launch(self._h, ctypes.byref(l)), "set_launch")

    def autotune(self, queries, repeats=2):
        """Pick the fastest launch shape on sample queries; returns it (and applies it)."""
        arr, keep = self._queries(queries)
        l = _Launch()
        self._check(lib().kpax_autotune(self._h, arr, len(queries), int(repeats), ctypes.byref(l)), "autotune")
        return {k: getattr(l, k) for k, _ in _Launch._fields_}

    def tree(self):
        """Tree of the last plan: states (S, n), controls (C, n), steps, parents, regions."""
        ts = ctypes.c_uint32()
        self._check(lib().kpax_get_tree(self._h, 0, None, None, None, None, None, ctypes.byref(ts)), "get_tree")
        n = ts.value
        st = np.zeros((self.S, n), np.float32)
        ct = np.zeros((self.C, n), np.float32)
        steps = np.zeros(n, np.int32)
        par = np.zeros(n, np.int32)
        reg = np.zeros(n, np.int32)
        self._check(lib().kpax_get_tree(self._h, n, _p(st), _p(ct), _p(steps), _p(par), _p(reg), ctypes.byref(ts)),
                    "get_tree")
        return {"states": st, "controls": ct, "steps": steps, "parents": par, "regions": reg}

    def regions(self):
        n = self.config.r1_count
        v, t, c = (np.zeros(n, np.int32) for _ in range(3))
        s = np.zeros(n, np.float32)
        self._check(lib().kpax_get_regions(self._h, _p(v), _p(t), _p(c), _p(s)), "get_regions")
        return {"valid": v, "total": t, "covered": c, "scores": s}

    def repeat_expand_noalloc(self, d_weights, n, d_out, out_cap, d_total):
        """kpax_repeat_expand_noalloc on device buffers (integer addresses);
        enqueued on the handle's stream, not waited for."""
        self._check(lib().kpax_repeat_expand_noalloc(self._h, ctypes.c_void_p(d_weights), int(n), ctypes.c_void_p(d_out),
                                                     int(out_cap), ctypes.c_void_p(d_total)), "repeat_expand_noalloc")

    def proposal_table(self):
        """Per-node Gaussians (8, n) of the last plan's tree under the learned proposal."""
        ts = ctypes.c_uint32()
        tree = self.tree()
        n = tree["parents"].shape[0]
        out = np.zeros((8, max(n, 1)), np.float32)
        self._check(lib().kpax_get_proposal(self._h, max(n, 1), _p(out), ctypes.byref(ts)), "get_proposal")
        return out[:, :ts.value]

    def lds_grid(self, env=0):
        """(dims, bytes) of environment env's LDS-staged grid; ((0, 0, 0), 0) when it has none."""
        dims = np.zeros(3, np.int32)
        b = ctypes.c_uint32()
        self._check(lib().kpax_get_lds_grid(self._h, int(env), _p(dims), ctypes.byref(b)), "get_lds_grid")
        return tuple(int(x) for x in dims), int(b.value)

    def grid(self):
        nc, nk = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(lib().kpax_get_grid(self._h, None, None, 0, ctypes.byref(nc), ctypes.byref(nk)), "get_grid")
        start = np.zeros(nc.value + 1, np.int32)
        items = np.zeros(max(nk.value, 1), np.int32)
        self._check(lib().kpax_get_grid(self._h, _p(start), _p(items), items.size, ctypes.byref(nc),
                                        ctypes.byref(nk)), "get_grid")
        return start, items[: nk.value]

    def set_profiling(self, on=True):
        self._check(lib().kpax_set_profiling(self._h, int(bool(on))), "set_profiling")

    def stats(self):
        s = _Stats()
        self._check(lib().kpax_get_stats(self._h, ctypes.byref(s)), "get_stats")
        return {"phase_ms": dict(zip(PHASES, list(s.phase_ms))),
                "phase_launches": dict(zip(PHASES, list(s.phase_launches))),
                "expansions": s.expansions, "integrated_steps": s.integrated_steps, "plans": s.plans,
                "iterations": s.iterations, "graph_launches": s.graph_launches}

    def reset_stats(self):
        self._check(lib().kpax_reset_stats(self._h), "reset_stats")

    # ---- hot-path operators -------------------------------------------------
    def propagate_batch(self, x0, u, steps):
        """Injected-control propagation: x0 (S, n), u (C, n), steps (n,) ->
        x1 (S, n), valid (n,) bool, fail_step (n,)."""
        steps = np.ascontiguousarray(steps, np.int32)
        n = steps.shape[0]
        x0 = _f32(x0, (self.S, n))
        u = _f32(u, (self.C, n))
        x1 = np.zeros((self.S, n), np.float32)
        valid = np.zeros(n, np.uint8)
        fail = np.zeros(n, np.int32)
        self._check(lib().kpax_propagate_batch(self._h, _p(x0), _p(u), _p(steps), n, _p(x1), _p(valid), _p(fail)),
                    "propagate_batch")
        return x1, valid.astype(bool), fail

    def collide_batch(self, bbmin, bbmax):
        bbmin = _f32(bbmin)
        n = bbmin.shape[1] if bbmin.ndim == 2 else 0
        bbmin = bbmin.reshape(3, n)
        bbmax = _f32(bbmax, (3, n))
        out = np.zeros(n, np.uint8)
        self._check(lib().kpax_collide_batch(self._h, _p(bbmin), _p(bbmax), n, _p(out)), "collide_batch")
        return out.astype(bool)

    def region_batch(self, states):
        st = _f32(states)
        n = st.shape[1]
        st = np.ascontiguousarray(st[:3])
        r1 = np.zeros(n, np.int32)
        r2 = np.zeros(n, np.int32)
        self._check(lib().kpax_region_batch(self._h, _p(st), n, _p(r1), _p(r2)), "region_batch")
        return r1, r2

    def region_histogram(self, states, valid, total=None, valid_count=None, covered=None):
        st = _f32(states)
        n = st.shape[1]
        st = np.ascontiguousarray(st[:3])
        v = np.ascontiguousarray(valid, np.uint8)
        R1 = self.config.r1_count
        total = np.zeros(R1, np.int32) if total is None else np.array(total, np.int32)
        valid_count = np.zeros(R1, np.int32) if valid_count is None else np.array(valid_count, np.int32)
        covered = np.zeros(R1, np.int32) if covered is None else np.array(covered, np.int32)
        self._check(lib().kpax_region_histogram(self._h, _p(st), _p(v), n, _p(total), _p(valid_count), _p(covered)),
                    "region_histogram")
        return total, valid_count, covered

    def scores(self, valid_count, total, covered):
        R1 = self.config.r1_count
        v = np.ascontiguousarray(valid_count, np.int32)
        t = np.ascontiguousarray(total, np.int32)
        c = np.ascontiguousarray(covered, np.int32)
        out = np.zeros(R1, np.float32)
        self._check(lib().kpax_scores(self._h, _p(v), _p(t), _p(c), _p(out)), "scores")
        return out

    def scan_compact(self, flags, cap=None):
        """Indices of the set flags, increasing (kpax_scan_compact)."""
        f = np.ascontiguousarray(np.asarray(flags) != 0, np.uint8)
        cap = int(f.sum()) if cap is None else int(cap)
        out = np.zeros(max(cap, 1), np.uint32)
        count = ctypes.c_uint32()
        self._check(lib().kpax_scan_compact(self._h, _p(f), f.shape[0], _p(out), cap, ctypes.byref(count)),
                    "scan_compact")
        return out[: min(cap, count.value)], count.value

    def repeat_expand(self, weights, cap=None):
        w = np.ascontiguousarray(weights, np.uint32)
        cap = int(w.sum(dtype=np.uint64)) if cap is None else int(cap)
        out = np.zeros(max(cap, 1), np.uint32)
        total = ctypes.c_uint32()
        self._check(lib().kpax_repeat_expand(self._h, _p(w), w.shape[0], _p(out), cap, ctypes.byref(total)),
                    "repeat_expand")
        return out[: min(cap, total.value)], total.value
