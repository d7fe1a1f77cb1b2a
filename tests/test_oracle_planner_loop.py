"""Pins the oracle's planner loop (oracle/kpax_oracle.cpp, Planner::plan) to
the reference's iteration rules with a second, separately written statement
of them in numpy.

The re-derivation below takes from the oracle only what other tests pin on
their own: Philox words (Random123 known answers), one edge's propagation
with injected controls (closed forms, special cases, FP64 agreement) and the
region of a state (lattice and boundary cases).  Everything the loop decides
is restated here from the reference:

  fan-out and M           KPAX.cu:323-333  (32 per repeat entry while it fits,
                                            else int(float(rem) / float(W)),
                                            min 32, 0 -> 1; M capped at rem)
  children's parents      KPAX.cu:366-379, 414-419 (child c of entry c / fan)
  control and step draws  statePropagatorSpatialHash.cu:74-77, 141-144, 236-241
  region counters         KPAX.cu:386-390 (total always, valid when valid)
  admission               KPAX.cu:391-396 (u < previous score || new R2 cell)
  scores                  Graph.cu:243-255, 297-304
  append and goal         KPAX.cu:458-502 (slot order, distance < radius)
  fAccept                 KPAX.cu:527-528 (itr * EPSILON * pow(1 - size/N, 5))
  re-admission            KPAX.cu:505-517 (u <= score + fAccept, repeat 1)
  repeat counts           KPAX.cu:474-479 (15 if valid < 10 else 1)
  repeat list             Planner.cu:52-80 (index i repeated weight[i] times)

and the trees, counters, scores, per-iteration records and repeat weights of
oracle.plan must equal it exactly.  The boundary of each comparison the
loop makes is pinned on hand inputs through the oracle's own predicates.

test_mutants_are_caught compiles the oracle with one plausible slip at a time
(swapped 15/1 repeat counts, < for <= in the re-admission, fAccept from the
tree size before the append, admission against this iteration's scores, the
fan-out rounded up or never lifted from 0, the last goal hit instead of the
first, ...) and shows that every one of them fails these checks.  (A clamp
taken at M >= remaining instead of M > remaining is no slip: at equality
int(remaining / W) is the fan-out 32 itself.)
"""
import math

import numpy as np
import pytest

import oracle_variants

F32 = np.float32
M32 = 0xFFFFFFFF


# ---- Philox4x32-10 in numpy (Salmon et al., SC'11), vectorised over counters
def philox(c0, c1, c2, c3, seed):
    c = [np.asarray(x, np.uint64) & M32 for x in np.broadcast_arrays(c0, c1, c2, c3)]
    k0, k1 = seed & M32, (seed >> 32) & M32
    for r in range(10):
        if r:
            k0, k1 = (k0 + 0x9E3779B9) & M32, (k1 + 0xBB67AE85) & M32
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ np.uint64(k0), p1 & np.uint64(M32),
             (p0 >> np.uint64(32)) ^ c[3] ^ np.uint64(k1), p0 & np.uint64(M32)]
    return [x.astype(np.uint32) for x in c]


def u01(w):
    """24 high bits of a word as a float in [0, 1)."""
    return (np.asarray(w, np.uint32) >> np.uint32(8)).astype(F32) * F32(2.0 ** -24)


def region_scores(cfg, valid, total, covered):
    """Graph.cu:243-255 (raw score) and :297-304 (normalisation), float32, with
    the total summed as 1024 strided partial sums then a halving tree."""
    L = [int(x) for x in cfg.r1_len]
    ext = [F32(cfg.ws_max[d]) - F32(cfg.ws_min[d]) for d in range(3)]
    size = [ext[d] / F32(L[d]) for d in range(3)]
    vol = size[0] * size[1] * size[2]
    r2_per = F32(int(np.prod(list(cfg.r2_len))))
    eps = F32(cfg.eps)
    raw = np.zeros(valid.shape[0], F32)
    m = valid > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        cov = covered[m].astype(F32) / r2_per
        ev = eps + valid[m].astype(F32)
        fv = ev / (ev + (total[m] - valid[m]).astype(F32)) * vol
        n = total[m].astype(F32)
        raw[m] = (fv * fv) * (fv * fv) / ((F32(1) + cov) * (F32(1) + n * n))
        padded = np.zeros(-(-raw.shape[0] // 1024) * 1024, F32)
        padded[:raw.shape[0]] = raw
        part = np.zeros(1024, F32)
        for k in range(padded.shape[0] // 1024):
            part = part + padded[k * 1024:(k + 1) * 1024]
        off = 512
        while off >= 1:
            part[:off] = part[:off] + part[off:2 * off]
            off //= 2
        tot = part[0]
        return np.where(valid == 0, F32(1), eps + raw / tot).astype(F32)


def rederive(oracle, cfg, obs, start, goal, seed, iters):
    """The reference's planner iterations restated; returns the same fields as
    oracle.plan(..., want_trace=True)."""
    N, S, C = int(cfg.max_tree), cfg.S, cfg.C
    R1 = int(np.prod(list(cfg.r1_len)))
    R2 = R1 * int(np.prod(list(cfg.r2_len)))
    st = np.zeros((S, N), F32)
    ct = np.zeros((C, N), F32)
    steps = np.zeros(N, np.int32)
    parent = np.full(N, -1, np.int32)
    reg = np.zeros(N, np.int32)
    valid, total, covered = (np.zeros(R1, np.int32) for _ in range(3))
    cell_seen = np.zeros(R2, bool)
    score = np.zeros(R1, F32)  # zero-filled before the first iteration (KPAX.cu:70)
    st[:, 0] = np.asarray(start, F32)[:S]
    reg[0] = oracle.regions(cfg, st[:3, :1])[0][0]
    ts = 1
    weights = np.zeros(N, np.uint32)
    weights[0] = cfg.root_repeat  # KPAX.cu:77
    cmin = np.array(cfg.ctrl_min, F32)[:C, None]
    crng = (np.array(cfg.ctrl_max, F32) - np.array(cfg.ctrl_min, F32))[:C, None]
    span = int(cfg.steps_max) - int(cfg.steps_min) + 1
    g = np.asarray(goal, F32)[:3]
    trace, status, goal_idx, expansions, t = [], 1, -1, 0, 0
    for t in range(1, iters + 1):
        W = int(weights[:ts].sum(dtype=np.uint64))
        rep = np.repeat(np.arange(ts, dtype=np.uint32), weights[:ts])[:N]
        remaining = N - ts
        if remaining == 0:
            status, t = 2, t - 1
            break
        fan = int(cfg.fanout)
        M = W * fan
        if M > remaining:
            fan = min(int(F32(remaining) / F32(W)), int(cfg.fanout))
            fan = fan if fan > 0 else 1
            M = min(W * fan, remaining)
        expansions += M
        ci = np.arange(M, dtype=np.uint64)
        par = rep[ci // fan]
        dc = philox(ci, t, 0, 0, seed)
        ds = philox(ci, t, 0, 1, seed)
        u = cmin + u01(np.stack(dc[:C])) * crng
        k = np.minimum((u01(ds[0]) * F32(span)).astype(np.int32), span - 1)
        nsteps = (int(cfg.steps_min) + k).astype(np.int32)
        x1, ok, _ = oracle.propagate(cfg, obs, st[:, par], u, nsteps)
        r1, r2 = oracle.regions(cfg, x1)
        np.add.at(total, r1, 1)
        np.add.at(valid, r1[ok], 1)
        by_score = ok & (u01(ds[1]) < score[r1])
        fresh = np.zeros(M, bool)
        for i in np.flatnonzero(ok):  # threads one at a time, in slot order
            if not cell_seen[r2[i]]:
                cell_seen[r2[i]] = True
                covered[r1[i]] += 1
                fresh[i] = True
        admitted = np.flatnonzero(by_score | fresh)
        score = region_scores(cfg, valid, total, covered)
        new = np.arange(ts, ts + admitted.shape[0])
        st[:, new] = x1[:, admitted]
        ct[:, new] = u[:, admitted]
        steps[new] = nsteps[admitted]
        parent[new] = par[admitted]
        reg[new] = r1[admitted]
        d = x1[:3, admitted] - g[:, None]
        dist = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        hits = np.flatnonzero(dist < F32(cfg.goal_radius))
        goal_idx = int(new[hits[0]]) if hits.shape[0] else -1
        ts_old, ts = ts, ts + admitted.shape[0]
        tas = F32(1) - F32(ts) / F32(N)
        fa = F32(float(F32(t) * F32(cfg.eps)) * math.pow(float(tas), 5))
        rec = {"W": W, "fan": fan, "M": M, "admitted": int(admitted.shape[0]), "ts": ts, "f_accept": float(fa),
               "goal": goal_idx, "next_W": 0, "goal_hits": int(hits.shape[0])}
        trace.append(rec)
        if goal_idx >= 0:
            status = 0
            break
        old = np.arange(ts_old)
        dr = philox(old, t, 0, 2, seed)
        weights[:ts_old] = (u01(dr[0]) <= score[reg[old]] + fa).astype(np.uint32)
        weights[ts_old:ts] = np.where(valid[reg[ts_old:ts]] < int(cfg.sparse_valid_below), int(cfg.sparse_repeat),
                                      int(cfg.dense_repeat))
        rec["next_W"] = int(weights[:ts].sum(dtype=np.uint64))
    else:
        status = 1
    return {"status": status, "iterations": t, "tree_size": ts, "goal_index": goal_idx, "expansions": expansions,
            "states": st[:, :ts], "controls": ct[:, :ts], "steps": steps[:ts], "parents": parent[:ts],
            "regions": reg[:ts], "valid": valid, "total": total, "covered": covered, "scores": score,
            "weights": weights[:ts], "trace": trace}


def cases(kpax):
    """(name, cfg, obstacles, start, goal, seed).  Small trees: the fan-out
    clamp binds from the second or third iteration on, and the last iterations
    run with a fan-out of 0 lifted to 1."""
    from workloads import workload
    out = []
    cfg, obs, start, goal = workload(1, seed=0, max_tree=2048, max_iter=5)
    out.append(("double integrator", cfg, obs, start, goal, 5))
    cfg, obs, start, goal = workload(2, seed=0, max_tree=3000, max_iter=6)
    out.append(("quadcopter", cfg, obs, start, goal, 11))
    cfg, obs, start, goal = workload(3, seed=0, max_tree=2048, max_iter=8)
    out.append(("dubins", cfg, obs, start, goal, 3))
    # goal ball around the start: several admitted children of one iteration
    # reach it, so the smallest tree index must be the one kept
    cfg, obs, start, goal = workload(1, seed=0, max_tree=4096, max_iter=5)
    out.append(("double integrator, goal hit by several children", cfg, obs, start, start[:3].copy(), 9))
    return out


def mismatches(oracle, cfg, obs, start, goal, seed, L=None):
    """Fields where oracle.plan (build L) differs from the re-derivation."""
    o = oracle.plan(cfg, obs, start, goal, seed=seed, threads=4, want_trace=True, L=L)
    r = rederive(oracle, cfg, obs, start, goal, seed, int(cfg.max_iter))
    bad = [k for k in ("status", "iterations", "tree_size", "goal_index", "expansions") if o[k] != r[k]]
    if o["tree_size"] == r["tree_size"]:
        for k in ("states", "controls"):
            if not np.array_equal(o[k].view(np.uint32), r[k].view(np.uint32)):
                bad.append(k)
        for k in ("steps", "parents", "regions", "weights"):
            if not np.array_equal(o[k], r[k]):
                bad.append(k)
    for k in ("valid", "total", "covered"):
        if not np.array_equal(o[k], r[k]):
            bad.append(k)
    if not np.array_equal(o["scores"].view(np.uint32), r["scores"].view(np.uint32)):
        bad.append("scores")
    keys = ("W", "fan", "M", "admitted", "ts", "f_accept", "goal", "next_W")
    if [[t[k] for k in keys] for t in o["trace"]] != [[t[k] for k in keys] for t in r["trace"]]:
        bad.append("trace")
    return bad, r


def test_philox_restatement_matches_oracle(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        ctr = rng.integers(0, 2**32, 4, dtype=np.uint64).astype(np.uint32)
        seed = int(rng.integers(0, 2**63))
        want = oracle.philox(ctr, [seed & M32, seed >> 32])
        got = philox(*[int(x) for x in ctr], seed)
        assert [int(x) for x in got] == [int(x) for x in want]
    # Random123 known answer (counter 0, key 0)
    assert [int(x) for x in philox(0, 0, 0, 0, 0)] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]


def test_scores_restatement_matches_oracle(oracle, kpax):
    rng = np.random.default_rng(2)
    for k in (1, 2):
        cfg = kpax.config_workload(k)
        R1 = cfg.r1_count
        total = rng.integers(0, 400, R1).astype(np.int32)
        valid = (total * rng.uniform(0, 1, R1)).astype(np.int32)
        valid[rng.uniform(0, 1, R1) < 0.3] = 0
        covered = np.minimum(valid, rng.integers(0, cfg.r2_per + 1, R1)).astype(np.int32)
        want, _ = oracle.scores(cfg, valid, total, covered)
        assert np.array_equal(region_scores(cfg, valid, total, covered).view(np.uint32), want.view(np.uint32))


def test_decision_boundaries(oracle, kpax):
    """Each comparison of the loop at its boundary, as the reference writes it."""
    bad = boundary_mismatches(oracle, kpax)
    assert not bad, bad


def boundary_mismatches(oracle, kpax, L=None):
    cfg = kpax.config_workload(1)
    half = 1 << 31  # u01 = 0.5 exactly
    step = 1 << 8   # one 2^-24 step of u01
    checks = [
        # KPAX.cu:393 curand_uniform < vertexScores: strict
        ("admission at u == score", oracle.admit_by_score(half, 0.5, L=L), False),
        ("admission just below the score", oracle.admit_by_score(half - step, 0.5, L=L), True),
        # KPAX.cu:512 curand_uniform <= vertexScores + fAccept: inclusive
        ("re-admission at u == score + fAccept", oracle.readmit(half, 0.25, 0.25, L=L), True),
        ("re-admission just above", oracle.readmit(half + step, 0.25, 0.25, L=L), False),
        # KPAX.cu:476-479: 15 while the region's valid count is below 10
        ("repeat count at valid 9", oracle.new_node_repeat(cfg, 9, L=L), 15),
        ("repeat count at valid 10", oracle.new_node_repeat(cfg, 10, L=L), 1),
        ("repeat count at valid 0", oracle.new_node_repeat(cfg, 0, L=L), 15),
        # KPAX.cu:323-333
        ("fan-out with room", oracle.fan_out(cfg, 5, 60000, L=L), (32, 160)),
        ("fan-out when 32 W fills the tree exactly", oracle.fan_out(cfg, 3, 96, L=L), (32, 96)),
        ("fan-out clamp", oracle.fan_out(cfg, 99, 1941, L=L), (19, 1881)),
        ("fan-out 0 lifted to 1", oracle.fan_out(cfg, 1845, 104, L=L), (1, 104)),
    ]
    return [f"{what}: {got} != {want}" for what, got, want in checks if got != want]


@pytest.mark.parametrize("case", range(4))
def test_planner_loop_rederived(oracle, kpax, case):
    name, cfg, obs, start, goal, seed = cases(kpax)[case]
    bad, r = mismatches(oracle, cfg, obs, start, goal, seed)
    assert not bad, f"{name}: oracle differs from the re-derivation in {bad}"
    fans = [t["fan"] for t in r["trace"]]
    if case == 3:
        assert r["status"] == 0 and r["trace"][-1]["goal_hits"] >= 2, r["trace"][-1]
    else:
        assert fans[0] == cfg.fanout and any(1 < f < cfg.fanout for f in fans), fans  # the clamp bound
        assert 1 in fans, fans  # a fan-out of 0 lifted to 1
        assert any(t["admitted"] < t["M"] for t in r["trace"])  # some children are turned away


# One plausible slip each: (what, text in oracle/kpax_oracle.cpp, replacement).
MUTANTS = [
    ("repeat counts 15 and 1 swapped", "? c.sparse_repeat : c.dense_repeat", "? c.dense_repeat : c.sparse_repeat"),
    ("re-admission with < instead of <=", "return u01(word) <= score + fa;", "return u01(word) < score + fa;"),
    ("admission with <= instead of <", "return u01(word) < score; }", "return u01(word) <= score; }"),
    ("fAccept from the tree size before the append", "const float fa = f_accept(c, t, ts);",
     "const float fa = f_accept(c, t, ts_old);"),
    ("repeat threshold <= 10", "return region_valid < c.sparse_valid_below", "return region_valid <= c.sparse_valid_below"),
    ("fan-out rounded up", "fan = (uint32_t)(int)((float)remaining / (float)W);",
     "fan = (uint32_t)(int)std::ceil((float)remaining / (float)W);"),
    ("fan-out of 0 not lifted to 1", "if(fan == 0) fan = 1;", ""),
    ("goal: the last hit instead of the first", "if(goal_idx < 0 && goal_distance(", "if(goal_distance("),
    ("invalid children cover sub-regions", "if(cvalid[ci] && r2_first[cr2[ci]] ==", "if(r2_first[cr2[ci]] =="),
    ("re-admission draws another node's stream", "draw(seed, j, t, query, PURPOSE_READMIT)",
     "draw(seed, j + 1, t, query, PURPOSE_READMIT)"),
    ("children's parents taken round-robin", "rep[cix / fan]", "rep[cix % rep.size()]"),
    ("valid counter counts every child", "if(fail == 0) __atomic_fetch_add(&valid[a1]", "__atomic_fetch_add(&valid[a1]"),
]
SCORES_CALL = "            scores(c, R, valid.data(), total.data(), covered.data(), score.data(), &score_total);\n"
ADMISSION_LOOP = "            for(uint32_t ci = 0; ci < M; ++ci)\n            {\n                bool fresh = false;\n"


def _mutant_sources():
    srcs = oracle_variants.sources(MUTANTS)
    src = open(oracle_variants.SRC).read()
    # admission against the scores of this iteration: the score update moved
    # in front of the admission loop
    assert src.count(SCORES_CALL) == 1 and src.count(ADMISSION_LOOP) == 1
    moved = src.replace(SCORES_CALL, "").replace(ADMISSION_LOOP, SCORES_CALL + ADMISSION_LOOP)
    srcs.append(("admission against this iteration's scores", moved))
    return srcs


def test_mutants_are_caught(oracle, kpax):
    """Every slip in MUTANTS, planted in a copy of the oracle built on the
    side, fails the re-derivation or the boundary checks."""
    survivors = []
    all_cases = cases(kpax)
    with oracle_variants.Variants(_mutant_sources()) as built:
        for what, so in built:
            L = oracle.load(so)
            caught = bool(boundary_mismatches(oracle, kpax, L))
            for name, cfg, obs, start, goal, seed in all_cases:
                if caught:
                    break
                caught = bool(mismatches(oracle, cfg, obs, start, goal, seed, L=L)[0])
            if not caught:
                survivors.append(what)
    assert not survivors, f"slips the checks do not catch: {survivors}"
