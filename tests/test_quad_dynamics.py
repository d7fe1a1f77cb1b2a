"""Pins every term of the oracle's quadcopter ODE (ode, statePropagator.cu:
325-370) with the mechanics it states, at random attitudes, rates and
velocities, in both the float form the planner integrates and double.

  kinematics   d(xyz)/dt = R(psi, theta, phi) (u, v, w), R the Z-Y-X Euler
               rotation (scipy's Rotation, not a formula of this repository),
               so |d(xyz)/dt| = |(u, v, w)| for any attitude  (:352-356)
  Euler rates  the body rate they imply, skew(w) = R^T dR/dt by a central
               difference of scipy's R, is (p, q, r)           (:357-359)
  translation  the world-frame acceleration R (dv_b/dt + w x v_b) is gravity
               (0, 0, GRAVITY) plus R times thrust and drag, with drag
               -NU |v| v / m: the Coriolis terms and the gravity projection
               cancel exactly as rigid-body motion requires   (:361-364)
  rotation     Euler's equations I dw/dt + w x (I w) = torque - MU |w| w;
               torque-free with MU = 0, kinetic energy and |I w| have zero
               derivative and stay constant along RK4 trajectories, the
               drift falling ~16x per halving of dt            (:366-369)

A dropped term, a wrong sign, a swapped inertia or a transposed rotation
fails at least one of these by orders of magnitude more than the tolerance:
test_slips_are_caught plants such slips in copies of the oracle and checks.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_variants

F_TOL = 2e-5   # float form, relative to the magnitudes involved
D_TOL = 1e-11  # double form
# Both forms take the constants as the reference's macros fold them in float
# (1.0f / MASS, (IY - IZ) / IX, 1 / IX, config.h:300-308).  Unequal inertias
# and a mass whose folded constants are exact in float keep the double
# form's mechanics exact, so its checks can be tight.
MASS, INERTIA = 2.0, (0.25, 0.5, 1.0)


def rot(phi, theta, psi):
    return Rotation.from_euler("ZYX", [psi, theta, phi]).as_matrix()


def states(rng, n):
    x = np.zeros((n, 12))
    x[:, 0:3] = rng.uniform(0, 5, (n, 3))
    x[:, 3] = rng.uniform(-np.pi, np.pi, n)
    x[:, 4] = rng.uniform(-1.3, 1.3, n)  # away from the gimbal lock at +-pi/2
    x[:, 5] = rng.uniform(-np.pi, np.pi, n)
    x[:, 6:9] = rng.uniform(-3, 3, (n, 3))
    x[:, 9:12] = rng.uniform(-4, 4, (n, 3))
    return x


@pytest.fixture(scope="module")
def quad(kpax):
    return kpax.config_workload(2)


def _cfg(cfg, **kw):
    c = cfg.copy()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("use_float", [False, True])
def test_kinematics_rotate_body_velocity(oracle, quad, rng, use_float, L=None):
    tol = F_TOL if use_float else D_TOL
    for x in states(rng, 300):
        d = oracle.quad_ode(quad, x, [0.0, 0.0, 0.0, 0.0], use_float, L=L)
        want = rot(*x[3:6]) @ x[6:9]
        scale = np.linalg.norm(x[6:9])
        assert np.abs(d[0:3] - want).max() <= tol * (1 + scale)
        assert abs(np.linalg.norm(d[0:3]) - scale) <= tol * (1 + scale)


@pytest.mark.parametrize("use_float", [False, True])
def test_euler_rates_give_the_body_rates(oracle, quad, rng, use_float, L=None):
    tol = 5e-4 if use_float else 1e-7
    h = 1e-6
    for x in states(rng, 200):
        d = oracle.quad_ode(quad, x, [0.0, 0.0, 0.0, 0.0], use_float, L=L)
        a, rates = x[3:6], d[3:6]
        R = rot(*a)
        dR = (rot(*(a + h * rates)) - rot(*(a - h * rates))) / (2 * h)
        W = R.T @ dR
        w = np.array([W[2, 1], W[0, 2], W[1, 0]])
        assert np.abs(W + W.T).max() < 1e-6  # skew-symmetric
        assert np.abs(w - x[9:12]).max() <= tol * (1 + np.abs(x[9:12]).max())


@pytest.mark.parametrize("use_float", [False, True])
def test_translational_dynamics_in_the_world_frame(oracle, quad, rng, use_float, L=None):
    """R (dv_b/dt + w x v_b) = (0, 0, GRAVITY) + R (0, 0, Zc / m) - NU |v| R v_b / m."""
    tol = F_TOL if use_float else D_TOL
    for nu in (0.0, 0.3):
        cfg = _cfg(quad, nu=nu, mass=MASS)
        for x in states(rng, 200):
            zc = rng.uniform(0, 20)
            d = oracle.quad_ode(cfg, x, [zc, 0.0, 0.0, 0.0], use_float, L=L)
            vb, w = x[6:9], x[9:12]
            R = rot(*x[3:6])
            acc = R @ (d[6:9] + np.cross(w, vb))
            m = np.float64(np.float32(cfg.mass))
            want = (np.array([0.0, 0.0, np.float32(cfg.gravity)]) + R @ np.array([0.0, 0.0, zc / m])
                    - np.float32(nu) * np.linalg.norm(vb) * (R @ vb) / m)
            scale = 10 + np.linalg.norm(w) * np.linalg.norm(vb) + zc
            assert np.abs(acc - want).max() <= tol * scale, (acc, want)


@pytest.mark.parametrize("use_float", [False, True])
def test_euler_equations(oracle, quad, rng, use_float, L=None):
    """I dw/dt + w x (I w) = torque - MU |w| w, for unequal inertias."""
    tol = F_TOL if use_float else D_TOL
    cfg = _cfg(quad, mu=0.05)
    cfg.inertia[0], cfg.inertia[1], cfg.inertia[2] = INERTIA
    I = np.array([np.float32(v) for v in cfg.inertia], np.float64)
    for x in states(rng, 300):
        tq = rng.uniform(-0.1, 0.1, 3)
        d = oracle.quad_ode(cfg, x, [0.0, *tq], use_float, L=L)
        w = x[9:12]
        lhs = I * d[9:12] + np.cross(w, I * w)
        rhs = tq - np.float32(0.05) * np.linalg.norm(w) * w
        scale = 0.1 + I.max() * (np.abs(d[9:12]).max() + np.linalg.norm(w) ** 2)
        assert np.abs(lhs - rhs).max() <= tol * scale, (lhs, rhs)


@pytest.mark.parametrize("use_float", [False, True])
def test_torque_free_rotation_conserves_energy_and_momentum(oracle, quad, rng, use_float, L=None):
    tol = F_TOL if use_float else D_TOL
    cfg = _cfg(quad, mu=0.0)
    cfg.inertia[0], cfg.inertia[1], cfg.inertia[2] = INERTIA
    I = np.array([np.float32(v) for v in cfg.inertia], np.float64)
    for x in states(rng, 300):
        d = oracle.quad_ode(cfg, x, [0.0, 0.0, 0.0, 0.0], use_float, L=L)
        w, wd = x[9:12], d[9:12]
        e_rate = np.sum(I * w * wd)          # d/dt 1/2 sum I w^2
        l_rate = np.sum(I * I * w * wd)      # d/dt 1/2 |I w|^2
        scale = np.linalg.norm(w) ** 3
        assert abs(e_rate) <= tol * scale and abs(l_rate) <= tol * scale


def test_torque_free_trajectory_conserves_energy_to_rk4_order(oracle, quad, rng):
    """Along double-precision RK4 edges with MU = NU = 0 and no torque,
    rotational energy and |I w| stay constant; their drift is 4th order in dt."""
    cfg = _cfg(quad, mu=0.0, nu=0.0, steps_min=1, steps_max=100000)
    cfg.inertia[0], cfg.inertia[1], cfg.inertia[2] = INERTIA
    for d in range(3):
        cfg.ws_min[d], cfg.ws_max[d] = -1e6, 1e6
    cfg.vel_min, cfg.vel_max = -1e6, 1e6
    I = np.array([np.float32(v) for v in cfg.inertia], np.float64)
    n = 64
    x0 = states(rng, n).T.astype(np.float32)
    x0[4] = rng.uniform(-0.5, 0.5, n)  # stay clear of gimbal lock over the edge
    u = np.zeros((4, n), np.float32)
    drift = []
    for dt, steps in ((0.004, 200), (0.002, 400)):
        cfg.dt = dt
        x1, ok, _ = oracle.propagate(cfg, None, x0, u, np.full(n, steps, np.int32), double=True)
        assert ok.all()
        w0, w1 = x0[9:12].astype(np.float64), x1[9:12]
        e0, e1 = np.sum(I[:, None] * w0 ** 2, 0), np.sum(I[:, None] * w1 ** 2, 0)
        l0, l1 = np.sum((I[:, None] * w0) ** 2, 0), np.sum((I[:, None] * w1) ** 2, 0)
        drift.append((np.abs(e1 / e0 - 1).max(), np.abs(l1 / l0 - 1).max()))
    (e_a, l_a), (e_b, l_b) = drift
    assert e_a < 1e-5 and l_a < 1e-5, drift
    assert e_b < e_a / 8 and l_b < l_a / 8, drift  # ~16x for 4th order


# One slip each in the float and double ODE (oracle/kpax_oracle.cpp quad_ode).
SLIPS = [
    ("(IZ - IX) / IY with its sign flipped", "k.k1y   = (T)((iz - ix) / iy);", "k.k1y   = (T)((ix - iz) / iy);"),
    ("a rotation entry with the wrong sign", "fma(fma(sfst, ss, cf * cs), v,", "fma(fma(sfst, ss, -(cf * cs)), v,"),
    ("cos and sin exchanged in z-dot", "fma(-st, uu, fma(sf * ct, v, cf * ct * w))", "fma(-st, uu, fma(cf * ct, v, sf * ct * w))"),
    ("tan(theta) replaced by sec(theta)", "out[3]       = fma(qr, tt, p);", "out[3]       = fma(qr, sec, p);"),
    ("theta-dot with phi's sin and cos exchanged", "out[4]       = fma(q, cf, -(r * sf));",
     "out[4]       = fma(q, sf, -(r * cf));"),
    ("a Coriolis term with the wrong sign", "fma(r, v, -(q * w))", "fma(r, v, q * w)"),
    ("gravity projected with cos(phi) on v", "fma(gct, sf, fma(p, w, -(r * uu)))", "fma(gct, cf, fma(p, w, -(r * uu)))"),
    ("drag without the 1 / m", "const T mX   = k.mi * (k.negnu *", "const T mX   = (k.negnu *"),
    ("rotational drag on the wrong rate", "fma(k.iyi * LMN, q, k.k1y * p * r)", "fma(k.iyi * LMN, p, k.k1y * p * r)"),
]


def test_slips_are_caught(oracle, quad):
    checks = [test_kinematics_rotate_body_velocity, test_euler_rates_give_the_body_rates,
              test_translational_dynamics_in_the_world_frame, test_euler_equations,
              test_torque_free_rotation_conserves_energy_and_momentum]
    survivors = []
    with oracle_variants.Variants(oracle_variants.sources(SLIPS)) as built:
        for what, so in built:
            L = oracle.load(so)
            caught = {False: False, True: False}
            for use_float in (False, True):
                for check in checks:
                    try:
                        check(oracle, quad, np.random.default_rng(5), use_float, L=L)
                    except AssertionError:
                        caught[use_float] = True
                        break
            if not (caught[False] and caught[True]):
                survivors.append((what, caught))
    assert not survivors, f"slips the checks do not catch: {survivors}"
