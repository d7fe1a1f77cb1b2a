"""Pins of the CPU oracle: known-answer vectors, closed forms, special cases,
convergence order and brute force.  Each check is chosen so a plausible
mistake in the restatement (a dropped term, a wrong coefficient or index, a
swapped operand) fails it.  CPU only."""
import numpy as np
import pytest
from scipy.integrate import quad as quadrature

from workloads import random_states, workload

def _cfg(kpax, k, **over):
    c = kpax.config_workload(k)
    for key, v in over.items():
        setattr(c, key, v)
    return c


def _wide(cfg):
    """No workspace or velocity limits in the way."""
    for d in range(3):
        cfg.ws_min[d] = -1e6
        cfg.ws_max[d] = 1e6
    cfg.vel_min, cfg.vel_max = -1e6, 1e6
    cfg.pitch_min, cfg.pitch_max = -1e6, 1e6
    return cfg


# --- Philox4x32-10: Random123 known-answer vectors (kat_vectors, philox4x32 10)
@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(oracle, ctr, key, want):
    assert list(oracle.philox(ctr, key)) == want


def test_philox_uniformity(oracle):
    # chi-square of 64 bins over 2^16 draws of word 0 for consecutive slots
    words = np.array([oracle.philox([i, 3, 0, 1], [123, 456])[0] for i in range(1 << 14)], np.uint64)
    hist = np.bincount((words >> np.uint64(26)).astype(np.int64), minlength=64)
    exp = words.size / 64
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5


def test_sincos_within_two_ulp(oracle):
    x = np.concatenate([np.linspace(-60, 60, 400001, dtype=np.float32),
                        np.array([0, -0.0, 1e-30, np.pi / 2, np.pi, 3 * np.pi / 2, 1000.0, -777.25], np.float32)])
    s, c = oracle.sincos_f32(x)
    xd = x.astype(np.float64)
    for got, want in ((s, np.sin(xd)), (c, np.cos(xd))):
        ulp = np.spacing(np.maximum(np.abs(want), 2.0 ** -24).astype(np.float32)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - want) / ulp
        assert err.max() <= 2.0, err.max()
    assert s[x.size - 8] == 0 and c[x.size - 8] == 1


def test_sincos_shift_matches_the_exact_shifted_angle(oracle, rng):
    """The RK4 stages' sin/cos of a + d: angle-sum identities for |d| <= 1
    (no rounding of the argument: within 4 ulp of 1 of the exact value),
    sincos_f32 of the float sum beyond (its argument rounding, half an ulp of
    a + d, plus 2 ulp).  A sign or coefficient slip in either series or in
    the identities is off by orders of magnitude."""
    a = rng.uniform(-8, 8, 300000).astype(np.float32)
    d = np.concatenate([rng.uniform(-1, 1, 200000), rng.uniform(-3, 3, 100000)]).astype(np.float32)
    a[:5] = [0, 0, 1.5707964, -3.1415927, 7.9]
    d[:5] = [0, 1.0, -1.0, 1e-20, -0.999]
    s, c = oracle.sincos_shift(a, d)
    exact = a.astype(np.float64) + d.astype(np.float64)
    small = np.abs(d) <= 1
    arg_err = 0.5 * np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    bound = np.where(small, 4 * 2.0 ** -24, arg_err + 2 * 2.0 ** -24)
    for got, want in ((s, np.sin(exact)), (c, np.cos(exact))):
        err = np.abs(got.astype(np.float64) - want)
        assert np.all(err <= bound), (err / bound).max()
    assert small.sum() > 200000 and (~small).sum() > 30000


# --- double integrator ------------------------------------------------------
def _di_closed_form(x0, u, n, dt, textbook):
    p, v, a = x0[:3], x0[3:6], u
    if textbook:
        return np.concatenate([p + n * v * dt + 0.5 * a * (n * dt) ** 2, v + n * a * dt])
    # reference update (statePropagatorSpatialHash.cu:92-97):
    # x += (5/6) v dt + (1/3) a dt^2, v += a dt
    x = p + (5.0 / 6.0) * dt * (n * v + a * dt * n * (n - 1) / 2) + n * a * dt * dt / 3.0
    return np.concatenate([x, v + n * a * dt])


@pytest.mark.parametrize("textbook", [0, 1])
def test_double_integrator_closed_form(kpax, oracle, rng, textbook):
    cfg = _wide(_cfg(kpax, 1, di_textbook=textbook))
    n = 300
    x0 = rng.uniform(-1, 1, (6, n)).astype(np.float32)
    u = rng.uniform(-1, 1, (3, n)).astype(np.float32)
    steps = rng.integers(1, 40, n).astype(np.int32)
    d, valid, _ = oracle.propagate(cfg, None, x0, u, steps, double=True)
    f, fvalid, _ = oracle.propagate(cfg, None, x0, u, steps)
    assert valid.all() and fvalid.all()
    dt = float(np.float32(cfg.dt))
    for i in range(n):
        want = _di_closed_form(x0[:, i].astype(np.float64), u[:, i].astype(np.float64), steps[i], dt, textbook)
        assert np.allclose(d[:, i], want, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(f - d) / np.maximum(1, np.abs(d))) < 1e-5


def test_double_integrator_reference_differs_from_textbook(kpax, oracle, rng):
    """The reference's mis-parenthesised update is kept by default (Appendix A Q1)."""
    x0 = np.zeros((6, 1), np.float32)
    x0[3:] = 0.5
    u = np.full((3, 1), 0.5, np.float32)
    steps = np.array([10], np.int32)
    ref, _, _ = oracle.propagate(_wide(_cfg(kpax, 1)), None, x0, u, steps, double=True)
    tb, _, _ = oracle.propagate(_wide(_cfg(kpax, 1, di_textbook=1)), None, x0, u, steps, double=True)
    dt = float(np.float32(0.01))
    # difference of the two closed forms: -(1/6) n v dt + (1/3 - 1/2) n a dt^2 - ... checked numerically
    want = _di_closed_form(x0[:, 0].astype(np.float64), u[:, 0].astype(np.float64), 10, dt, 0) - \
        _di_closed_form(x0[:, 0].astype(np.float64), u[:, 0].astype(np.float64), 10, dt, 1)
    assert np.allclose(ref[:, 0] - tb[:, 0], want, atol=1e-12)
    assert abs(want[0]) > 1e-4


# --- Dubins airplane ----------------------------------------------------------
def test_dubins_is_simpson_of_the_exact_integrand(kpax, oracle, rng):
    cfg = _wide(_cfg(kpax, 3))
    n = 60
    x0 = np.zeros((6, n), np.float32)
    x0[:3] = rng.uniform(0.2, 0.8, (3, n))
    x0[3] = rng.uniform(-3, 3, n)
    x0[4] = rng.uniform(-0.8, 0.8, n)
    x0[5] = rng.uniform(0.05, 0.3, n)
    u = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.78, 0.78, n), rng.uniform(-0.78, 0.78, n)]).astype(
        np.float32)
    steps = rng.integers(5, 21, n).astype(np.int32)
    d, valid, _ = oracle.propagate(cfg, None, x0, u, steps, double=True)
    assert valid.all()
    dt = float(np.float32(cfg.dt))
    for i in range(n):
        x, y, z, yaw, pitch, v = x0[:, i].astype(np.float64)
        a, yr, pr = u[:, i].astype(np.float64)
        T = steps[i] * dt
        # yaw, pitch and v are linear in time (statePropagatorSpatialHash.cu:176-178)
        assert abs(d[3, i] - (yaw + T * yr)) < 1e-12
        assert abs(d[4, i] - (pitch + T * pr)) < 1e-12
        assert abs(d[5, i] - (v + T * a)) < 1e-12
        fx = lambda s: (v + a * s) * np.cos(pitch + pr * s) * np.cos(yaw + yr * s)
        fy = lambda s: (v + a * s) * np.cos(pitch + pr * s) * np.sin(yaw + yr * s)
        fz = lambda s: (v + a * s) * np.sin(pitch + pr * s)
        for k, f, p0 in ((0, fx, x), (1, fy, y), (2, fz, z)):
            exact = p0 + quadrature(f, 0, T, epsabs=1e-14, epsrel=1e-14)[0]
            # composite Simpson error <= T h^4 max|f''''| / 2880 with h = dt: ~1e-12 here
            assert abs(d[k, i] - exact) < 1e-10, (i, k, d[k, i] - exact)


# --- quadcopter -----------------------------------------------------------------
def _quad_state(n=1):
    x0 = np.zeros((12, n), np.float32)
    x0[:3] = np.array([[5.0], [5.0], [2.5]], np.float32)
    return x0


def test_quad_hover_is_a_fixed_point(kpax, oracle):
    cfg = _wide(_cfg(kpax, 2))
    x0 = _quad_state()
    u = np.array([[9.81], [0], [0], [0]], np.float32)  # thrust = -GRAVITY * MASS
    for double in (False, True):
        x1, valid, _ = oracle.propagate(cfg, None, x0, u, np.array([20], np.int32), double=double)
        assert valid.all() and np.array_equal(np.asarray(x1, np.float64), x0.astype(np.float64))


def test_quad_free_fall_matches_drag_solution(kpax, oracle):
    """Zc = 0 from rest: w' = G - NU w |w| / m exactly (ode rows 8)."""
    cfg = _wide(_cfg(kpax, 2))
    x0 = _quad_state()
    u = np.zeros((4, 1), np.float32)
    n = 20
    x1, valid, _ = oracle.propagate(cfg, None, x0, u, np.array([n], np.int32), double=True)
    g, k = 9.81, float(np.float32(10e-3))
    g = float(np.float32(9.81))
    t = n * float(np.float32(cfg.dt))
    w_exact = -np.sqrt(g / k) * np.tanh(np.sqrt(g * k) * t)
    z_exact = 2.5 - np.log(np.cosh(np.sqrt(g * k) * t)) / k
    assert abs(x1[8, 0] - w_exact) < 1e-9
    assert abs(x1[2, 0] - np.float32(2.5) + 2.5 - z_exact) < 1e-9
    assert x1[0, 0] == 5.0 and x1[1, 0] == 5.0
    assert np.all(x1[[3, 4, 5, 6, 7, 9, 10, 11], 0] == 0)


def test_quad_pure_torque_spins_about_one_axis(kpax, oracle):
    """Lc only, from rest: p' = Lc/IX - MU p|p|/IX, phi' = p, nothing else moves but w (gravity)."""
    cfg = _wide(_cfg(kpax, 2))
    x0 = _quad_state()
    u = np.array([[9.81], [0.1], [0], [0]], np.float32)
    n = 10
    x1, valid, _ = oracle.propagate(cfg, None, x0, u, np.array([n], np.int32), double=True)
    I = float(np.float32(1.0) / np.float32(1.0 / np.float32(0.01)))  # 1 / (1/IX) as folded
    ixi = float(np.float32(1.0) / np.float32(0.01))
    L, mu, t = float(np.float32(0.1)), float(np.float32(2e-6)), n * float(np.float32(0.02))
    # p' = ixi (L - mu p^2) -> p = sqrt(L/mu) tanh(ixi sqrt(L mu) t)
    p_exact = np.sqrt(L / mu) * np.tanh(ixi * np.sqrt(L * mu) * t)
    assert abs(x1[9, 0] - p_exact) < 1e-9
    assert x1[10, 0] == 0 and x1[11, 0] == 0 and x1[4, 0] == 0 and x1[5, 0] == 0
    assert x1[3, 0] > 0  # roll integrates p
    del I


def test_quad_rk4_is_fourth_order(kpax, oracle, rng):
    """Halving dt divides the global error by ~16 (double path)."""
    base = _wide(_cfg(kpax, 2))
    x0 = _quad_state()
    x0[3:6, 0] = [0.3, -0.2, 0.5]
    x0[6:9, 0] = [1.0, -0.5, 0.3]
    x0[9:12, 0] = [2.0, -1.5, 1.0]
    u = np.array([[12.0], [0.05], [-0.04], [0.03]], np.float32)
    T = 0.4
    sols = []
    for dt in (0.02, 0.01, 0.005, 0.0025, 0.000625):
        cfg = base.copy()
        cfg.dt = dt
        n = int(round(T / float(np.float32(dt))))
        x1, valid, _ = oracle.propagate(cfg, None, x0, u, np.array([n], np.int32), double=True)
        sols.append(x1[:, 0])
    ref = sols[-1]
    e = [np.max(np.abs(s - ref)) for s in sols[:3]]
    assert 12 < e[0] / e[1] < 20 and 12 < e[1] / e[2] < 20, e


@pytest.mark.parametrize("k", [1, 2, 3])
def test_float_path_error_against_fp64_is_stated(kpax, oracle, rng, k):
    """The float planner path stays within 1e-5 (relative to max(1,|x|)) of the FP64 truth."""
    cfg = _cfg(kpax, k)
    x0, u, steps = random_states(cfg, 3000, rng)
    f, fv, ff = oracle.propagate(cfg, None, x0, u, steps)
    d, dv, df = oracle.propagate(cfg, None, x0, u, steps, double=True)
    both = fv & dv & (ff == df)
    assert both.mean() > 0.5
    err = np.abs(f[:, both] - d[:, both]) / np.maximum(1.0, np.abs(d[:, both]))
    print(f"workload {k}: float vs FP64 max rel err {err.max():.2e}")
    assert err.max() < 1e-5


# --- collision ------------------------------------------------------------------
def test_broad_phase_semantics(oracle):
    obs = np.array([[0.4, 0.4, 0.4, 0.6, 0.6, 0.6]], np.float32)
    cases = [  # (lo, hi, free)
        ([0.5, 0.5, 0.5], [0.5, 0.5, 0.5], False),   # zero-length segment inside
        ([0.1, 0.1, 0.1], [0.45, 0.45, 0.45], False),  # overlap
        ([0.6, 0.4, 0.4], [0.7, 0.5, 0.5], True),    # touching the max face: free (<=)
        ([0.3, 0.4, 0.4], [0.4, 0.5, 0.5], True),    # touching the min face: free
        ([0.3, 0.3, 0.3], [0.7, 0.7, 0.7], False),   # box contains the obstacle
        ([0.7, 0.0, 0.0], [0.8, 1.0, 1.0], True),    # separated on x only
        ([0.0, 0.0, 0.61], [1.0, 1.0, 0.9], True),   # separated on z only
    ]
    lo = np.array([c[0] for c in cases], np.float32).T.copy()
    hi = np.array([c[1] for c in cases], np.float32).T.copy()
    got = oracle.collide(obs, lo, hi)
    assert list(got) == [c[2] for c in cases]


def test_collide_vectorised_brute_force(oracle, rng):
    obs = np.concatenate([rng.uniform(0, 1, (300, 3)), np.zeros((300, 3))], 1).astype(np.float32)
    obs[:, 3:] = obs[:, :3] + rng.uniform(0, 0.1, (300, 3)).astype(np.float32)
    a = rng.uniform(0, 1, (3, 5000)).astype(np.float32)
    b = a + rng.uniform(-0.05, 0.05, (3, 5000)).astype(np.float32)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    got = oracle.collide(obs, lo, hi)
    sep = (hi.T[:, None, :] <= obs[None, :, :3]) | (obs[None, :, 3:] <= lo.T[:, None, :])
    want = sep.any(2).all(1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("k", [2, 3])
def test_grid_verdicts_equal_brute_force(kpax, oracle, rng, k):
    """The oracle planner's grid (the reference's spatial hash as CSR) gives
    the brute-force verdict for random segments, segments touching obstacle
    faces (free), zero-length segments and segments beyond the grid."""
    cfg, obs, _, _ = workload(k, seed=1)
    lo_ws = np.array(cfg.ws_min, np.float32)[:, None]
    ext = (np.array(cfg.ws_max, np.float32) - np.array(cfg.ws_min, np.float32))[:, None]
    n = 40000
    a = (lo_ws + ext * rng.uniform(-0.1, 1.1, (3, n))).astype(np.float32)
    b = (a + ext * rng.uniform(-0.02, 0.02, (3, n))).astype(np.float32)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    # boxes ending exactly on an obstacle's face, and points
    pick = rng.integers(0, obs.shape[0], 2000)
    hi[0, :2000] = obs[pick, 0]
    lo[0, :2000] = np.minimum(lo[0, :2000], hi[0, :2000])
    lo[:, 2000:4000] = hi[:, 2000:4000]
    got = oracle.collide_grid(cfg, obs, lo, hi)
    want = oracle.collide(obs, lo, hi)
    assert np.array_equal(got, want)
    assert 0.05 < want.mean() < 0.999  # both verdicts occur


def test_grid_counts_cover_every_obstacle_point(kpax, oracle, rng):
    cfg, obs, _, _ = workload(2, seed=1)
    counts = oracle.grid_counts(cfg, obs)
    D = np.array(cfg.grid_dims)
    cell = (np.array(cfg.ws_max, np.float32) - np.array(cfg.ws_min, np.float32)) / D.astype(np.float32)
    # every point inside an obstacle lies in a cell that lists at least one obstacle
    for o in obs[:100]:
        pts = o[:3] + (o[3:] - o[:3]) * rng.uniform(0, 1, (50, 3))
        g = np.clip(((pts - np.array(cfg.ws_min)) / cell).astype(np.int64), 0, D - 1)
        idx = g[:, 0] + D[0] * (g[:, 1] + D[1] * g[:, 2])
        assert np.all(counts[idx] >= 1)
    # total insertions = sum over obstacles of the product of their cell spans
    lo = np.clip(((obs[:, :3] - np.array(cfg.ws_min)) / cell).astype(np.int64), 0, D - 1)
    hi = np.clip(((obs[:, 3:] - np.array(cfg.ws_min)) / cell).astype(np.int64), 0, D - 1)
    assert counts.sum() == np.prod(hi - lo + 1, axis=1).sum()


# --- regions ----------------------------------------------------------------------
@pytest.mark.parametrize("k", [1, 2])
def test_region_lattice(kpax, oracle, k):
    cfg = _cfg(kpax, k)
    L = np.array(cfg.r1_len)
    L2 = np.array(cfg.r2_len)
    lo = np.array(cfg.ws_min, np.float64)
    size = (np.array(cfg.ws_max, np.float64) - lo) / L
    ii = np.stack(np.meshgrid(*[np.arange(l) for l in L], indexing="ij"), -1).reshape(-1, 3)
    jj = np.array([[1, 2, 3]]) % L2
    centre = lo + (ii + (jj + 0.5) / L2) * size
    r1, r2 = oracle.regions(cfg, centre.T.astype(np.float32))
    want1 = (ii[:, 0] * L[1] + ii[:, 1]) * L[2] + ii[:, 2]
    want2 = want1 * L2.prod() + (jj[0, 0] * L2[1] + jj[0, 1]) * L2[2] + jj[0, 2]
    assert np.array_equal(r1, want1) and np.array_equal(r2, want2)
    # exact lower boundaries belong to the upper cell; the far wall and beyond clamp
    edge = (lo + ii * size).T.astype(np.float32)
    assert np.array_equal(oracle.regions(cfg, edge)[0], want1)
    far = np.array(cfg.ws_max, np.float32)[:, None] + np.array([[0], [0], [0]], np.float32)
    assert oracle.regions(cfg, far)[0][0] == L.prod() - 1
    out = np.array([[-5.0], [1e9], [np.nan]], np.float32)
    assert oracle.regions(cfg, out)[0][0] == (0 * L[1] + (L[1] - 1)) * L[2] + 0


def test_histogram_counts_and_first_arrival(kpax, oracle):
    cfg = _cfg(kpax, 1)
    # four children in the same R2 cell, two in another; child 0 invalid
    st = np.array([[0.01, 0.011, 0.012, 0.013, 0.2, 0.21], [0.01] * 4 + [0.2, 0.21], [0.01] * 4 + [0.2, 0.21]],
                  np.float32)
    valid = np.array([0, 1, 1, 1, 1, 1], np.uint8)
    total, vc, cov, r2f, new = oracle.histogram(cfg, st, valid)
    assert total.sum() == 6 and vc.sum() == 5
    assert list(new) == [False, True, False, False, True, False]
    assert cov.sum() == 2


def test_scores_formula(kpax, oracle, rng):
    cfg = _cfg(kpax, 2)
    R1 = cfg.r1_count
    valid = rng.integers(0, 40, R1).astype(np.int32)
    valid[rng.uniform(0, 1, R1) < 0.6] = 0
    total = valid + rng.integers(0, 20, R1).astype(np.int32)
    covered = np.minimum(valid, rng.integers(0, 65, R1)).astype(np.int32)
    got, tot = oracle.scores(cfg, valid, total, covered)
    eps = 1e-2
    vol = np.prod((np.array(cfg.ws_max, np.float64) - np.array(cfg.ws_min)) / np.array(cfg.r1_len))
    fv = (eps + valid) / (eps + valid + (total - valid)) * vol
    raw = np.where(valid > 0, fv ** 4 / ((1 + covered / cfg.r2_per) * (1 + total.astype(np.float64) ** 2)), 0)
    want = np.where(valid == 0, 1.0, eps + raw / raw.sum())
    assert abs(tot - raw.sum()) <= 1e-5 * raw.sum()
    assert np.allclose(got, want, rtol=2e-6)
    assert abs((got[valid > 0] - eps).sum() - 1.0) < 1e-5


def test_repeat_expand_matches_numpy(oracle, rng):
    w = rng.integers(0, 16, 10000).astype(np.uint32)
    out, total = oracle.repeat_expand(w)
    assert total == w.sum() and np.array_equal(out, np.repeat(np.arange(10000), w))


def test_f_accept(kpax, oracle):
    cfg = _cfg(kpax, 2)
    for it, ts in [(1, 161), (7, 500000), (30, 1 << 20)]:
        tas = np.float32(1.0) - np.float32(ts) / np.float32(cfg.max_tree)
        want = float(np.float32(it) * np.float32(0.01)) * float(tas) ** 5
        assert abs(oracle.f_accept(cfg, it, ts) - want) <= 1e-6 * abs(want) + 1e-30


# --- whole planner ----------------------------------------------------------------
@pytest.mark.parametrize("k,kw", [(1, dict(max_tree=20000, max_iter=10)), (2, dict(max_tree=1 << 15, max_iter=8)),
                                  (3, dict(max_tree=1 << 15, max_iter=8))])
def test_oracle_planner_invariants(kpax, oracle, k, kw):
    cfg, obs, start, goal = workload(k, seed=2, **kw)
    o = oracle.plan(cfg, obs, start, goal, seed=5, threads=4)
    o2 = oracle.plan(cfg, obs, start, goal, seed=5, threads=1)
    assert o["tree_size"] == o2["tree_size"] and np.array_equal(o["states"], o2["states"])  # deterministic
    ts = o["tree_size"]
    par = o["parents"]
    assert par[0] == -1 and np.all(par[1:] < np.arange(1, ts)) and np.all(par[1:] >= 0)
    # every node is its parent propagated with its stored control and steps
    x1, valid, _ = oracle.propagate(cfg, obs, o["states"][:, par[1:]], o["controls"][:, 1:], o["steps"][1:])
    assert valid.all()
    assert np.array_equal(x1.view(np.uint32), o["states"][:, 1:].view(np.uint32))
    assert np.array_equal(oracle.regions(cfg, o["states"])[0], o["regions"])
    assert o["total"].sum() == o["expansions"]
    assert np.all(o["valid"] <= o["total"]) and np.all(o["covered"] <= cfg.r2_per)
    assert o["covered"].sum() <= o["valid"].sum()
    if o["status"] == kpax.OK:
        g = o["goal_index"]
        assert np.linalg.norm(o["states"][:3, g] - goal) < cfg.goal_radius


def test_oracle_quad_plan_path_verifies_in_fp64(kpax, oracle):
    cfg, obs, start, goal = workload(2, seed=0, max_tree=1 << 18)
    o = oracle.plan(cfg, obs, start, goal, seed=11, threads=8)
    assert o["status"] == kpax.OK
    path = [o["goal_index"]]
    while o["parents"][path[-1]] != -1:
        path.append(o["parents"][path[-1]])
    path = path[::-1]
    S, C = cfg.S, cfg.C
    rows = np.zeros((len(path), S + C + 1), np.float32)
    rows[:, :S] = o["states"][:, path].T
    rows[:, S:S + C] = o["controls"][:, path].T
    rows[:, -1] = np.float32(cfg.dt) * o["steps"][path]
    ok, dev = kpax.verify_path(cfg, obs, rows, o["steps"][path], list(goal))
    assert ok and dev < 1e-3


# ---- learned control proposal: oracle pins ---------------------------------
def test_ln_f32_matches_log_within_two_ulp(oracle, kpax):
    xs = np.concatenate([np.float32(2.0) ** -np.arange(0, 25, dtype=np.float32),
                         np.random.default_rng(0).uniform(2**-24, 1, 2000).astype(np.float32),
                         np.array([np.float32(0.70710677), np.float32(0.70710683), np.float32(1.0)])])
    for x in xs:
        got = np.float32(oracle.ln_f32(float(x)))
        want = np.float32(np.log(np.float64(x)))
        ulp = np.spacing(np.abs(want)) if want != 0 else np.float32(1e-45)
        assert abs(np.float64(got) - np.float64(want)) <= 2 * ulp, (x, got, want)


def test_box_muller_normals_are_standard_normal(oracle, kpax):
    cfg = kpax.config_workload(2)
    n = 200000
    rng = np.random.default_rng(1)
    w = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    sigma = 0.05
    ms = np.zeros((n, 8), np.float32)
    ms[:, 4:] = sigma
    u = oracle.propose_sample(cfg, ms, w)
    lo, hi = np.array(cfg.ctrl_min[:4]), np.array(cfg.ctrl_max[:4])
    z = (2 * (u - lo) / (hi - lo) - 1) / sigma
    assert np.all(np.abs(z.mean(0)) < 0.01) and np.all(np.abs(z.std(0) - 1) < 0.01)
    from scipy import stats
    for j in range(4):
        assert stats.kstest(z[:, j], "norm").pvalue > 1e-3
    # the two normals of a pair are uncorrelated
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.01


def test_proposal_clamps_to_the_control_box(oracle, kpax):
    cfg = kpax.config_workload(2)
    ms = np.zeros((4, 8), np.float32)
    ms[:, :4] = [[5, -5, 5, -5]] * 4
    ms[:, 4:] = 0.01
    w = np.arange(16, dtype=np.uint32).reshape(4, 4) * 977
    u = oracle.propose_sample(cfg, ms, w)
    assert np.array_equal(u, np.tile(np.array([cfg.ctrl_max[0], cfg.ctrl_min[1], cfg.ctrl_max[2], cfg.ctrl_min[3]],
                                              np.float32), (4, 1)))


def test_features_closed_forms_and_occupancy(oracle, kpax):
    cfg = kpax.config_workload(2)  # [0,10]^2 x [0,5]; lattice cells 0.625 x 0.625 x 0.3125
    # one obstacle filling lattice cell (3, 4, 5) exactly: only that cell is set
    cell = np.array([10 / 16, 10 / 16, 5 / 16], np.float32)
    lo = np.array([3, 4, 5], np.float32) * cell
    obs = np.concatenate([lo, lo + cell])[None].astype(np.float32)
    bits = oracle.occupancy(cfg, obs)
    idx = (3 * 16 + 4) * 16 + 5
    assert bits[idx >> 5] == 1 << (idx & 31) and np.count_nonzero(bits) == 1
    # a state at the workspace centre, at rest, goal at the centre too
    x = np.zeros((12, 2), np.float32)
    x[:3, 0] = [5, 5, 2.5]
    x[:3, 1] = [-3, 5, 2.5]  # outside: every window cell left of the workspace is occupied
    x[3:6, 0] = [np.pi, -np.pi / 2, 0]
    x[6:9, 0] = [3, -1.5, 0]
    x[9:12, 0] = [4, 0, -2]
    g = np.array([[5, 10], [5, 5], [2.5, 2.5]], np.float32)
    f = oracle.mlp_features(cfg, obs, x, g)
    assert np.array_equal(f[0, :3], [0, 0, 0]) and np.array_equal(f[0, 12:15], [0, 0, 0])
    assert np.allclose(f[0, 3:6], [1, -0.5, 0], atol=1e-7)
    assert np.allclose(f[0, 6:9], [1, -0.5, 0]) and np.allclose(f[0, 9:12], [1, 0, -0.5])
    assert np.allclose(f[1, 12:15], [13 / 5, 0, 0])
    occ0 = f[0, 15:79].reshape(4, 4, 4)
    # centre (8, 8, 8) in cells -> window cells 6..9 on every axis; cell (3,4,5) is outside it
    assert occ0.sum() == 0
    occ1 = f[1, 15:79].reshape(4, 4, 4)  # x = -3 -> cells -7..-4: all outside the workspace
    assert occ1.min() == 1
    assert f[0, 79] == 0


def test_mlp_oracle_gradients_match_finite_differences(oracle, kpax):
    import mlp_oracle
    rng = np.random.default_rng(2)
    p = kpax.mlp_init_params(4).astype(np.float64)
    p += rng.normal(0, 0.05, p.shape)
    X = rng.normal(0, 1, (7, 80))
    X[:, 79] = 0
    Y = rng.uniform(-1, 1, (7, 4))
    _, g = mlp_oracle.loss_and_grads(p, X, Y, exact=True)
    idx = np.concatenate([rng.choice(np.arange(mlp_oracle.W1, mlp_oracle.B1), 30, replace=False),
                          np.arange(mlp_oracle.B1, mlp_oracle.B1 + 5),
                          rng.choice(np.arange(mlp_oracle.W2, mlp_oracle.B2), 30, replace=False),
                          np.arange(mlp_oracle.B2, mlp_oracle.B2 + 5),
                          np.arange(mlp_oracle.W3, mlp_oracle.W3 + 40), np.arange(mlp_oracle.B3, mlp_oracle.B3 + 8)])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mlp_oracle.loss_and_grads(q, X, Y, exact=True)
        q[i] -= 2 * h
        lm, _ = mlp_oracle.loss_and_grads(q, X, Y, exact=True)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    import mlp_oracle
    # bf16 keeps 7 fraction bits: the spacing at 1 is 2^-7
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 3 * 2**-9, 1.0 + 2**-9, -2.5], np.float32)
    r = mlp_oracle.bf16(x)
    # halfway cases go to the even neighbour: 1 + 2^-8 -> 1, 1 + 3 2^-8 -> 1 + 2^-6
    assert list(r) == [1.0, 1.0, 1.0 + 2**-6, 1.0 + 2**-7, 1.0, -2.5]


# ---- goal-progress pruning (PruneKPAX.cu:466-501) ----------------------------
def test_pruning_with_certain_acceptance_is_the_plain_planner(oracle, kpax):
    cfg, obs, start, goal = workload(1, seed=0, max_tree=20000, max_iter=8)
    plain = oracle.plan(cfg, obs, start, goal, seed=3, threads=4)
    cfg.prune, cfg.prune_exploration, cfg.prune_goal_bias = 1, 1.0, 0.0
    pruned = oracle.plan(cfg, obs, start, goal, seed=3, threads=4)
    assert pruned["tree_size"] == plain["tree_size"] and np.array_equal(pruned["parents"], plain["parents"])


def test_pruning_keeps_goal_ward_children(oracle, kpax):
    # exploration 0, goal bias 1, a small progress scale: only children that move
    # toward the goal survive, so every kept edge ends nearer the goal than it began
    cfg, obs, start, goal = workload(1, seed=0, max_tree=20000, max_iter=6, n_obs=0)
    cfg.prune, cfg.prune_exploration, cfg.prune_goal_bias, cfg.prune_progress_scale = 1, 0.0, 1.0, 1e-4
    o = oracle.plan(cfg, obs, start, goal, seed=5, threads=4)
    st, par = o["states"], o["parents"]
    d = np.linalg.norm(st[:3].T - np.asarray(goal[:3], np.float32), axis=1)
    kept = np.arange(1, o["tree_size"])
    closer = d[kept] < d[par[kept]]
    # p = eps for children that move away; almost none of those survive
    assert closer.mean() > 0.97, closer.mean()
    plain = cfg.copy()
    plain.prune = 0
    assert o["tree_size"] < oracle.plan(plain, obs, start, goal, seed=5, threads=4)["tree_size"]
