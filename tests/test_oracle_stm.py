"""kete's two-body state transition matrix, restated in the oracle
(oracle.compute_state_transition: compute_state_transition, stm_ivp_eqn,
accel_grad and the RK4(5) Fehlberg integrator; propagation/
state_transition.rs:6-112, acceleration.rs:254-290, runge_kutta.rs:33-147).

Pins:
  - RK45: kete's known answers (runge_kutta.rs:158-197) and exactness on
    y' = 5 t^4 (the fifth-order weights integrate degree-5 solutions
    exactly), which a wrong Fehlberg coefficient breaks;
  - with its two arithmetic defects switched off (STM_FIX_GRAD |
    STM_FIX_ACCEL) the restated algorithm gives the closed-form two-body
    STM (central differences of the Kepler solution) and the Kepler final
    state: the integration, the scaling and the reshapes are kete's;
  - kete's own arithmetic (fix 0) reproduces the defects that make its STM
    unusable as a reference: the lower-left block comes out -1/2 times the
    true one (accel_grad), the trajectory barely bends (central_accel's
    day-unit GMS inside the tau = k t equation), and the returned matrix is
    the transpose of d x(t1) / d x(t0); and it cannot integrate backward
    (a negative step fails the step-size check).
"""
import math

import numpy as np
import pytest

import oracle as o

GMS_SQRT = 0.01720209894996


@pytest.mark.parametrize("step", [1, 2, 3])
def test_rk45_exp_known_answer(step):
    st, y, _ = o.rk45_test(0, 1.0, 0.0, step * 10.0, 1e-14)
    assert st == 0 and abs(y - math.exp(-step * 10.0)) < 1e-14


def test_rk45_line_known_answer():
    st, y, _ = o.rk45_test(1, 1.0, 0.0, 10.0, 0.001)
    assert st == 0 and abs(y - 21.0) < 1e-12


def test_rk45_fifth_order_exact():
    st, y, ne = o.rk45_test(2, 1.0, 0.0, 2.0, 1e-9)
    assert st == 0 and ne > 6
    assert abs(y - 33.0) < 1e-12 * 33.0


def kepler_stm(p, v, dt):
    """d x(dt) / d x(0) by 5-point central differences of the Kepler solution."""
    x = np.concatenate([p, v])
    J = np.zeros((6, 6))
    for k in range(6):
        h = 1e-5 if k < 3 else 1e-7
        ev = {}
        for c in (-2, -1, 1, 2):
            xp = x.copy()
            xp[k] += c * h
            st, a, b = o.analytic_2_body(dt, xp[:3], xp[3:])
            assert st == 0
            ev[c] = np.concatenate([a, b])
        J[:, k] = (ev[-2] - 8 * ev[-1] + 8 * ev[1] - ev[2]) / (12 * h)
    return J


ORBITS = [  # circular 1 AU, eccentric NEO-like, hyperbolic
    (np.array([1.0, 0.0, 0.0]), np.array([0.0, GMS_SQRT, 0.0])),
    (np.array([0.9, 0.2, 0.05]), np.array([-0.004, 0.021, 0.002])),
    (np.array([1.2, 0.0, 0.1]), np.array([0.0, 0.03, 0.001])),
]


@pytest.mark.parametrize("orbit", range(len(ORBITS)))
@pytest.mark.parametrize("dt", [10.0, 100.0, 365.25])
def test_fixed_arithmetic_is_the_two_body_stm(orbit, dt):
    p, v = ORBITS[orbit]
    st, po, vo, M, ne = o.compute_state_transition(p, v, 0.0, dt, 1.0,
                                                   o.STM_FIX_GRAD | o.STM_FIX_ACCEL)
    assert st == 0 and ne > 0
    st2, pk, vk = o.analytic_2_body(dt, p, v)
    assert np.abs(po - pk).max() < 1e-10 and np.abs(vo - vk).max() < 1e-12
    T = kepler_stm(p, v, dt)
    assert np.abs(M.T - T).max() < 1e-7 * np.abs(T).max()
    # kete returns the transpose: M itself is not d x(t1) / d x(t0)
    assert np.abs(M - T).max() > 1e-3 * np.abs(T).max()


def test_backward_arcs_fail_as_in_kete():
    """RK45Integrator::step rejects step_size < 1e-30 (runge_kutta.rs:34-36),
    which every negative step is: kete's STM cannot go back in time (its
    caller unwraps the error)."""
    p, v = ORBITS[0]
    st, *_ = o.compute_state_transition(p, v, 0.0, -60.0, 1.0, 3)
    assert st == o.CONVERGENCE


def test_ketes_arithmetic_defects():
    p, v = ORBITS[0]
    # accel_grad: over a short arc STM - I ~ A dt, and the lower-left block
    # (d v / d r) comes out -1/2 of the true one
    dt = 0.05
    st, _, _, M, _ = o.compute_state_transition(p, v, 0.0, dt, 1.0, o.STM_FIX_ACCEL)
    T = kepler_stm(p, v, dt)
    ratio = M.T[3:, :3][np.abs(T[3:, :3]) > 1e-3 * np.abs(T[3:, :3]).max()] / \
        T[3:, :3][np.abs(T[3:, :3]) > 1e-3 * np.abs(T[3:, :3]).max()]
    assert np.allclose(ratio, -0.5, rtol=2e-2)
    # central_accel in the tau equation: gravity k^2 too weak, a near
    # straight line: 100 d of a 1 AU circular orbit ends ~1.15 AU off
    st, po, _, M0, _ = o.compute_state_transition(p, v, 0.0, 100.0, 1.0, 0)
    _, pk, _ = o.analytic_2_body(100.0, p, v)
    line = p + v * 100.0
    assert np.linalg.norm(po - pk) > 1.0
    assert np.linalg.norm(po - line) < 1e-3 * np.linalg.norm(pk - line)
    # and kete's matrix is far from the true STM
    T = kepler_stm(p, v, 100.0)
    assert np.abs(M0.T - T).max() > 0.3 * np.abs(T).max()
