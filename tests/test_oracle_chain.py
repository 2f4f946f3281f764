"""The oracle's orbit-fit observable chain (oracle.chain_observe): the
checker of the partials kernel.

- restart=True is kete's chain -- successive propagate_n_body calls
  (propagation/mod.rs:70-118), one light-time two-body step
  (fov/fov_like.rs:89-110), RA/Dec (simult_states.rs:86-120) -- assembled
  here from the oracle's pinned primitives; values agree to the integrator's
  error level (the chain keeps barycentric equatorial coordinates between
  arcs, so the last bits of the step decisions can differ);
- a recorded step schedule replays bitwise, in both chain modes;
- with the schedule frozen the map is smooth: 5-point differences at steps
  h and 10 h agree to 1e-7 of each column's largest entry, where the
  adaptive map's scatter by 1e-5 (its step decisions jump with the input).
  Differences of the frozen map are the reference for the derivative of the
  computed map with its step sequence held fixed, which is what the partials
  kernel returns.
"""
import math

import numpy as np
import pytest

import oracle
from kete_amd import synthetic as syn

JD0 = syn.EPOCH_JD
C_INV = 0.005775518331436995
ROT_C, ROT_S = math.cos(0.40909280422232897), math.sin(0.40909280422232897)


def ecl_to_equ(v):
    return np.array([v[0], ROT_C * v[1] - ROT_S * v[2], ROT_S * v[1] + ROT_C * v[2]])


@pytest.fixture(scope="module")
def workload(oracle_spk):
    xyz = syn.n_body_workload(n=12, neo_fraction=1.0, seed_index=7)
    jd_obs = JD0 + np.sort(np.random.default_rng(1).uniform(0.0, 730.0, 16))
    earth = np.array([np.concatenate(oracle.spk_state(399, t, 10, 17)[1:]) for t in jd_obs]).T
    return xyz, jd_obs, earth[:3]


def kete_chain(xyz, jd_obs, obs):
    n, m = xyz.shape[1], len(jd_obs)
    radec = np.full((2, m, n), np.nan)
    cur, t = xyz.copy(), np.full(n, JD0)
    for j in range(m):
        r = oracle.batch_n_body(t, cur, jd_obs[j])
        assert np.all(r["status"] == 0)
        cur, t = r["out"], np.full(n, jd_obs[j])
        ob = ecl_to_equ(obs[:, j])
        for i in range(n):
            x, v = ecl_to_equ(cur[:3, i]), ecl_to_equ(cur[3:, i])
            st, p2, v2 = oracle.analytic_2_body(-np.linalg.norm(x - ob) * C_INV, x, v)
            assert st == 0
            radec[:, j, i] = oracle.ra_dec_with_rates(p2, v2, ob, np.zeros(3))[:2]
    return radec


def test_restart_chain_is_ketes(workload):
    xyz, jd_obs, obs = workload
    r = oracle.chain_observe(JD0, xyz, jd_obs, obs, restart=True)
    assert np.all(r["status"] == 0)
    ref = kete_chain(xyz, jd_obs, obs)
    d = r["radec"] - ref
    d[0] = (d[0] + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(d).max() < 1e-9


@pytest.mark.parametrize("restart", [True, False])
def test_schedule_replays_bitwise(workload, restart):
    xyz, jd_obs, obs = workload
    rec = oracle.chain_observe(JD0, xyz, jd_obs, obs, restart=restart, record=True)
    rep = oracle.chain_observe(JD0, xyz, jd_obs, obs, restart=restart,
                               schedule=rec["schedule"])
    assert np.all(rec["status"] == 0) and np.all(rep["status"] == 0)
    assert np.array_equal(rec["radec"], rep["radec"])
    assert np.array_equal(rec["state"], rep["state"])
    assert np.all(rec["schedule"][3] > len(jd_obs))
    # the continuous chain takes fewer steps than restarting at every epoch
    if not restart:
        rr = oracle.chain_observe(JD0, xyz, jd_obs, obs, restart=True, record=True)
        assert np.all(rec["schedule"][3] < rr["schedule"][3])


def five_point(xyz, jd_obs, obs, restart, schedule, steps):
    n, m = xyz.shape[1], len(jd_obs)
    J = np.zeros((2, 6, m, n))
    w = {-2: 1.0, -1: -8.0, 1: 8.0, 2: -1.0}
    for k in range(6):
        ev = {}
        for c in w:
            x = xyz.copy()
            x[k] += c * steps[k]
            ev[c] = oracle.chain_observe(JD0, x, jd_obs, obs, restart=restart,
                                         schedule=schedule)["radec"]
        ra = sum(w[c] * ((ev[c][0] - ev[-2][0] + np.pi) % (2 * np.pi) - np.pi) for c in w)
        J[0, k] = ra / (12 * steps[k])
        J[1, k] = sum(w[c] * ev[c][1] for c in w) / (12 * steps[k])
    return J


def rel(a, b):
    scale = np.abs(b).max(axis=(0, 2), keepdims=True)
    return (np.abs(a - b) / scale).max()


@pytest.mark.parametrize("restart", [True, False])
def test_frozen_schedule_map_is_smooth(workload, restart):
    xyz, jd_obs, obs = workload
    rec = oracle.chain_observe(JD0, xyz, jd_obs, obs, restart=restart, record=True)
    steps = np.array([1e-6] * 3 + [1e-8] * 3)
    j1 = five_point(xyz, jd_obs, obs, restart, rec["schedule"], steps)
    j10 = five_point(xyz, jd_obs, obs, restart, rec["schedule"], 10 * steps)
    assert rel(j10, j1) < 1e-7
    a1 = five_point(xyz, jd_obs, obs, restart, None, steps)
    a10 = five_point(xyz, jd_obs, obs, restart, None, 10 * steps)
    assert rel(a10, a1) > 1e-6          # the adaptive map is not smooth at this level
