"""The oracle's fov_spk_check (checks.rs:123-142, fov_like.rs:196-226): SPK
objects are found in fields aimed at them, with the state moved back by one
two-body light-time step; fields aimed away find nothing; ids the ephemeris
does not hold are skipped; hits come by field (epoch order), then id."""
import math

import numpy as np
import pytest

import oracle
from kete_amd import synthetic as syn

AU = 149597870.7
C_INV = 0.005775518331436995  # 1 / c in day/AU


@pytest.fixture
def with_asteroids(oracle_spk):
    assert oracle.spk_load(syn.ephemeris_path(), syn.asteroid_ephemeris_path()) == oracle.OK
    yield oracle
    assert oracle.spk_load(syn.ephemeris_path()) == oracle.OK


def aimed_fields(ids, jds, width_deg=2.0, away=False):
    fjd, obs, nrm = [], [], []
    w = math.radians(width_deg)
    for jd in jds:
        _, op, ov = oracle.spk_state(399, jd, 10, 17)
        for aid in ids:
            _, p, _ = oracle.spk_state(aid, jd, 10, 17)
            d = (p - op) * (-1.0 if away else 1.0)
            fjd.append(jd)
            obs.append(np.concatenate([op, ov]))
            nrm.append(oracle.rect_normals(d / np.linalg.norm(d), 0.0, w, w).reshape(12))
    return np.array(fjd), np.array(obs), np.array(nrm)


def test_spk_objects_found_with_light_time(with_asteroids):
    ids = [20000001, 20000002, 20000004, 20000010, 20000704]
    jds = syn.EPOCH_JD + np.array([-400.0, -10.0, 30.0, 200.0])
    fjd, obs, nrm = aimed_fields(ids, jds)
    hf, ho, hs = oracle.fov_spk_check(ids + [2099999999], fjd, obs, nrm)
    # each field sees the asteroid it is aimed at (and perhaps others)
    for f in range(len(fjd)):
        assert (f, f % len(ids)) in set(zip(hf.tolist(), ho.tolist()))
    # order: by field, then id index
    keys = list(zip(hf.tolist(), ho.tolist()))
    assert keys == sorted(keys)
    # the state is the SPK state moved by one two-body step of -|d|/c
    for (f, k), st in zip(keys, hs):
        jd = fjd[f]
        _, p, v = oracle.spk_state(ids[k], jd, 10, 17)
        dt = -np.linalg.norm(p - obs[f, :3]) * C_INV
        # propagate_two_body(state, jd + dt) steps by (jd + dt) - jd
        s2, p2, v2 = oracle.analytic_2_body((jd + dt) - jd, p, v)
        assert s2 == 0
        assert st[0] == jd + dt
        assert np.abs(st[1:4] - p2).max() < 1e-15 and np.abs(st[4:] - v2).max() < 1e-17
        assert abs(dt) > 1e-3  # the light time is a few minutes at 2-4 AU


def test_spk_objects_not_found_away_or_missing(with_asteroids):
    ids = [20000001, 20000004]
    fjd, obs, nrm = aimed_fields(ids, [syn.EPOCH_JD + 5.0], away=True)
    hf, _, _ = oracle.fov_spk_check(ids, fjd, obs, nrm)
    assert len(hf) == 0
    # outside the asteroid segments' span: no state, skipped
    fjd2, obs2, nrm2 = aimed_fields(ids, [syn.EPOCH_JD])
    fjd2 = fjd2 + 5000.0
    hf, _, _ = oracle.fov_spk_check(ids, fjd2, obs2, nrm2)
    assert len(hf) == 0
