"""ctypes wrapper of liboracle_kpax.so — the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  Configs are passed as the
kpax_config ctypes structure (kino-pax-fork_amd/kpax.py: Config).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_kpax.so")
_L = None


def _sigs():
    P, u32, i32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
    f32 = ctypes.c_float
    return {
        "oracle_philox": (None, [P, P, P]),
        "oracle_sincos_f32": (None, [P, u32, P, P]),
        "oracle_sincos_shift": (None, [P, P, u32, P, P]),
        "oracle_propagate": (i32, [P, P, u32, P, P, P, u32, i32, P, P, P]),
        "oracle_collide": (None, [P, u32, P, P, u32, P]),
        "oracle_collide_grid": (None, [P, P, u32, P, P, u32, P]),
        "oracle_grid_counts": (None, [P, P, u32, P]),
        "oracle_regions": (None, [P, P, u32, P, P]),
        "oracle_histogram": (None, [P, P, P, u32, i32, P, P, P, P, P]),
        "oracle_scores": (None, [P, P, P, P, P, P]),
        "oracle_f_accept": (f32, [P, u32, u32]),
        "oracle_repeat_expand": (u64, [P, u32, P, u64]),
        "oracle_plan": (i32, [P, P, u32, P, P, u64, i32, P, P, P, P, P, P, P, P, P, P]),
        "oracle_plan_ex": (i32, [P, P, u32, P, P, u64, i32, P, P, P, P, P, P, P, P, P, P, P, P, P, f32]),
        "oracle_fan_out": (None, [P, u64, u32, P, P]),
        "oracle_admit_by_score": (i32, [u32, f32]),
        "oracle_readmit": (i32, [u32, f32, f32]),
        "oracle_new_node_repeat": (u32, [P, i32]),
        "oracle_quad_ode": (None, [P, P, P, P, i32]),
    }


def load(path):
    """A CDLL of an oracle build at `path` with its signatures set (the
    default build is lib(); tests load deliberately broken builds this way)."""
    L = ctypes.CDLL(path)
    for name, (res, args) in _sigs().items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with oracle/build.py")
        _L = load(LIB_PATH)
    return _L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _obs(o):
    o = np.ascontiguousarray(o if o is not None else np.zeros((0, 6)), np.float32).reshape(-1, 6)
    return o, o.shape[0]


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().oracle_philox(_p(c), _p(k), _p(o))
    return o


def sincos_f32(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().oracle_sincos_f32(_p(x), x.size, _p(s), _p(c))
    return s, c


def sincos_shift(a, d):
    """sin and cos of a + d as the quadcopter's RK4 stages compute them: from
    sincos_f32(a) by the angle-sum identities (|d| <= 1) or sincos_f32(a + d)."""
    a = np.ascontiguousarray(a, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    s = np.zeros_like(a)
    c = np.zeros_like(a)
    lib().oracle_sincos_shift(_p(a), _p(d), a.size, _p(s), _p(c))
    return s, c


def propagate(cfg, obstacles, x0, u, steps, double=False):
    """x0 (S, n), u (C, n), steps (n,) -> x1 (S, n), valid (n,), fail_step (n,)."""
    steps = np.ascontiguousarray(steps, np.int32)
    n = steps.shape[0]
    S = 12 if cfg.model == 3 else 6
    C = 4 if cfg.model == 3 else 3
    x0 = np.ascontiguousarray(x0, np.float32).reshape(S, n)
    u = np.ascontiguousarray(u, np.float32).reshape(C, n)
    x1 = np.zeros((S, n), np.float64 if double else np.float32)
    valid = np.zeros(n, np.uint8)
    fail = np.zeros(n, np.int32)
    o, no = _obs(obstacles)
    lib().oracle_propagate(ctypes.byref(cfg), _p(o), no, _p(x0), _p(u), _p(steps), n, int(double), _p(x1), _p(valid),
                           _p(fail))
    return x1, valid.astype(bool), fail


def collide(obstacles, bbmin, bbmax):
    bbmin = np.ascontiguousarray(bbmin, np.float32)
    n = bbmin.shape[1]
    bbmax = np.ascontiguousarray(bbmax, np.float32).reshape(3, n)
    o, no = _obs(obstacles)
    out = np.zeros(n, np.uint8)
    lib().oracle_collide(_p(o), no, _p(bbmin), _p(bbmax), n, _p(out))
    return out.astype(bool)


def collide_grid(cfg, obstacles, bbmin, bbmax):
    """Verdicts through the oracle planner's uniform grid (cfg.grid_dims)."""
    bbmin = np.ascontiguousarray(bbmin, np.float32)
    n = bbmin.shape[1]
    bbmax = np.ascontiguousarray(bbmax, np.float32).reshape(3, n)
    o, no = _obs(obstacles)
    out = np.zeros(n, np.uint8)
    lib().oracle_collide_grid(ctypes.byref(cfg), _p(o), no, _p(bbmin), _p(bbmax), n, _p(out))
    return out.astype(bool)


def grid_counts(cfg, obstacles):
    o, no = _obs(obstacles)
    G = int(cfg.grid_dims[0]) * int(cfg.grid_dims[1]) * int(cfg.grid_dims[2])
    counts = np.zeros(G, np.int32)
    lib().oracle_grid_counts(ctypes.byref(cfg), _p(o), no, _p(counts))
    return counts


def regions(cfg, states):
    st = np.ascontiguousarray(np.asarray(states, np.float32)[:3])
    n = st.shape[1]
    r1 = np.zeros(n, np.int32)
    r2 = np.zeros(n, np.int32)
    lib().oracle_regions(ctypes.byref(cfg), _p(st), n, _p(r1), _p(r2))
    return r1, r2


def histogram(cfg, states, valid, iteration=1, total=None, valid_count=None, covered=None, r2_first=None):
    st = np.ascontiguousarray(np.asarray(states, np.float32)[:3])
    n = st.shape[1]
    v = np.ascontiguousarray(valid, np.uint8)
    R1 = int(np.prod(list(cfg.r1_len)))
    R2 = R1 * int(np.prod(list(cfg.r2_len)))
    total = np.zeros(R1, np.int32) if total is None else np.array(total, np.int32)
    valid_count = np.zeros(R1, np.int32) if valid_count is None else np.array(valid_count, np.int32)
    covered = np.zeros(R1, np.int32) if covered is None else np.array(covered, np.int32)
    r2_first = np.full(R2, np.iinfo(np.int32).max, np.int32) if r2_first is None else r2_first
    is_new = np.zeros(n, np.uint8)
    lib().oracle_histogram(ctypes.byref(cfg), _p(st), _p(v), n, int(iteration), _p(total), _p(valid_count),
                           _p(covered), _p(r2_first), _p(is_new))
    return total, valid_count, covered, r2_first, is_new.astype(bool)


def scores(cfg, valid_count, total, covered):
    R1 = int(np.prod(list(cfg.r1_len)))
    v = np.ascontiguousarray(valid_count, np.int32)
    t = np.ascontiguousarray(total, np.int32)
    c = np.ascontiguousarray(covered, np.int32)
    out = np.zeros(R1, np.float32)
    tot = np.zeros(1, np.float32)
    lib().oracle_scores(ctypes.byref(cfg), _p(v), _p(t), _p(c), _p(out), _p(tot))
    return out, float(tot[0])


def f_accept(cfg, iteration, tree_plus_new):
    return lib().oracle_f_accept(ctypes.byref(cfg), int(iteration), int(tree_plus_new))


def repeat_expand(weights, cap=None):
    w = np.ascontiguousarray(weights, np.uint32)
    cap = int(w.sum(dtype=np.uint64)) if cap is None else int(cap)
    out = np.zeros(max(cap, 1), np.uint32)
    total = lib().oracle_repeat_expand(_p(w), w.shape[0], _p(out), cap)
    return out[: min(cap, total)], int(total)


def plan(cfg, obstacles, start, goal, seed, threads=0, want_tree=True, want_trace=False, L=None, proposal=None,
         mix=0.0):
    """Whole planner; returns a dict with status, iterations, tree arrays,
    counters.  want_trace adds "weights" (every node's repeat count in the last
    repeat list built) and "trace" (per iteration: W, fan, M, admitted, tree
    size after the append, fAccept, goal index or -1, next W).  L: an oracle
    build from load() instead of the default one.  proposal (8 x n floats:
    per tree node 4 means then 4 sigmas, n <= max_tree) with mix > 0: children
    draw their control from the parent's Gaussian with probability mix."""
    L = lib() if L is None else L
    o, no = _obs(obstacles)
    N = int(cfg.max_tree)
    S = 12 if cfg.model == 3 else 6
    C = 4 if cfg.model == 3 else 3
    R1 = int(np.prod(list(cfg.r1_len)))
    x0 = np.zeros(16, np.float32)
    x0[:S] = np.asarray(start, np.float32).ravel()[:S]
    g = np.zeros(4, np.float32)
    g[:3] = np.asarray(goal, np.float32).ravel()[:3]
    info = np.zeros(5, np.uint32)
    st = np.zeros((S, N), np.float32) if want_tree else None
    ct = np.zeros((C, N), np.float32) if want_tree else None
    steps = np.zeros(N, np.int32) if want_tree else None
    par = np.zeros(N, np.int32) if want_tree else None
    reg = np.zeros(N, np.int32) if want_tree else None
    v, t, c = (np.zeros(R1, np.int32) for _ in range(3))
    sc = np.zeros(R1, np.float32)
    wts = np.zeros(N, np.uint32) if want_trace else None
    tr = np.zeros((max(int(cfg.max_iter), 1), 8), np.uint32) if want_trace else None
    table = None
    if proposal is not None:
        proposal = np.asarray(proposal, np.float32)
        table = np.zeros((8, N), np.float32)
        table[:, :proposal.shape[1]] = proposal
    status = L.oracle_plan_ex(ctypes.byref(cfg), _p(o), no, _p(x0), _p(g), int(seed) & (2**64 - 1), int(threads),
                              _p(info), _p(st), _p(ct), _p(steps), _p(par), _p(reg), _p(v), _p(t), _p(c), _p(sc),
                              _p(wts), _p(tr), _p(table), float(mix if table is not None else 0.0))
    ts = int(info[1])
    out = {"status": int(status), "iterations": int(info[0]), "tree_size": ts,
           "goal_index": int(np.int32(info[2])), "expansions": int(info[3]) | (int(info[4]) << 32),
           "valid": v, "total": t, "covered": c, "scores": sc}
    if want_tree:
        out.update(states=st[:, :ts], controls=ct[:, :ts], steps=steps[:ts], parents=par[:ts], regions=reg[:ts])
    if want_trace:
        n = out["iterations"]
        rows = tr[:n]
        out["weights"] = wts[:ts]
        out["trace"] = [{"W": int(r[0]), "fan": int(r[1]), "M": int(r[2]), "admitted": int(r[3]), "ts": int(r[4]),
                         "f_accept": float(np.array([r[5]], np.uint32).view(np.float32)[0]),
                         "goal": int(np.int32(r[6])), "next_W": int(r[7])} for r in rows]
    return out


def fan_out(cfg, W, remaining, L=None):
    L = lib() if L is None else L
    fan = ctypes.c_uint32()
    M = ctypes.c_uint64()
    L.oracle_fan_out(ctypes.byref(cfg), int(W), int(remaining), ctypes.byref(fan), ctypes.byref(M))
    return fan.value, M.value


def admit_by_score(word, score, L=None):
    return bool((lib() if L is None else L).oracle_admit_by_score(int(word), float(score)))


def readmit(word, score, fa, L=None):
    return bool((lib() if L is None else L).oracle_readmit(int(word), float(score), float(fa)))


def quad_ode(cfg, x, u, use_float=False, L=None):
    """The quadcopter's state derivative at x (12) under control u (4): the
    float form the planner integrates, or double."""
    x = np.ascontiguousarray(x, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros(12, np.float64)
    (lib() if L is None else L).oracle_quad_ode(ctypes.byref(cfg), _p(x), _p(u), _p(out), int(use_float))
    return out


def new_node_repeat(cfg, region_valid, L=None):
    return int((lib() if L is None else L).oracle_new_node_repeat(ctypes.byref(cfg), int(region_valid)))


# ---- learned control proposal (the build's own definition; see mlp_oracle.py)
def _prop_lib():
    L = lib()
    if not getattr(L, "_prop_sigs", False):
        P, u32 = ctypes.c_void_p, ctypes.c_uint32
        for name, res, args in (("oracle_occupancy", None, [P, P, u32, P]),
                                ("oracle_mlp_features", None, [P, P, u32, P, P, u32, P]),
                                ("oracle_propose_sample", None, [P, P, P, u32, P]),
                                ("oracle_ln_f32", ctypes.c_float, [ctypes.c_float])):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._prop_sigs = True
    return L


def occupancy(cfg, obstacles):
    o, no = _obs(obstacles)
    bits = np.zeros(128, np.uint32)
    _prop_lib().oracle_occupancy(ctypes.byref(cfg), _p(o), no, _p(bits))
    return bits


def mlp_features(cfg, obstacles, states, goals):
    o, no = _obs(obstacles)
    st = np.ascontiguousarray(states, np.float32)
    g = np.ascontiguousarray(goals, np.float32)
    n = st.shape[1]
    out = np.zeros((n, 80), np.float32)
    _prop_lib().oracle_mlp_features(ctypes.byref(cfg), _p(o), no, _p(st), _p(g), n, _p(out))
    return out


def propose_sample(cfg, mean_sigma, words):
    ms = np.ascontiguousarray(mean_sigma, np.float32)
    w = np.ascontiguousarray(words, np.uint32)
    n = ms.shape[0]
    out = np.zeros((n, 4), np.float32)
    _prop_lib().oracle_propose_sample(ctypes.byref(cfg), _p(ms), _p(w), n, _p(out))
    return out


def ln_f32(x):
    return float(_prop_lib().oracle_ln_f32(float(x)))
