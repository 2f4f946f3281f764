"""GPU parity of fov_state_check (kete's field-of-view path with light time)
and the RA/Dec + V-magnitude epilogue, against the oracle.

Visible sets must be identical except for objects within 1e-9 rad of a field
edge; light-time-corrected states agree to 1e-9 AU; RA/Dec to 1e-9 rad.
"""
import math

import numpy as np
import pytest

import kete_amd as ka
from kete_amd import synthetic as syn
from kete_amd.fov import fov_state_check, ra_dec_vmag

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload():
    xyz, fjd, obs, pts, h = syn.fov_workload(n=3000, n_fields=24, span_days=40.0, aimed=0.75,
                                             seed_index=4)
    import oracle
    w = math.radians(6.9)
    nrm = np.array([oracle.rect_normals(p, 0.0, w, w).reshape(12) for p in pts])
    return xyz, fjd, obs, nrm, h


def edge_distance(nrm, obs, state):
    """Angular distance (rad) of the object's direction from the nearest
    edge plane of the field."""
    rel = state[1:4] - obs[:3]
    u = rel / np.linalg.norm(rel)
    n = nrm.reshape(4, 3)
    return np.min(np.abs(n @ u) / np.linalg.norm(n, axis=1))


def test_fov_parity(ctx, oracle_spk, workload):
    xyz, fjd, obs, nrm, h = workload
    n = xyz.shape[1]
    batch = ka.StateArray(np.full(n, syn.EPOCH_JD), xyz)
    hits = fov_state_check(batch, None, dt_limit=3.0, fields=(fjd, obs, nrm))
    of, oo, os_ = oracle_spk.fov_state_check(syn.EPOCH_JD, xyz, fjd, obs.T.copy(), nrm,
                                             dt_limit=3.0)
    assert len(of) > 50, "workload should produce detections"
    g = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(hits.field, hits.obj))}
    r = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(of, oo))}
    for key in set(g) ^ set(r):
        f, i = key
        st = os_[r[key]] if key in r else hits.state[:, g[key]]
        assert edge_distance(nrm[f], obs[:, f], st) < 1e-9, key
    common = sorted(set(g) & set(r))
    gs = np.array([hits.state[:, g[k]] for k in common])
    rs = np.array([os_[r[k]] for k in common])
    assert np.abs(gs[:, 0] - rs[:, 0]).max() < 1e-9        # light-time epoch (days)
    assert np.abs(gs[:, 1:4] - rs[:, 1:4]).max() < 1e-9    # position (AU)
    # order: fields by epoch, objects ascending (as kete returns them)
    keys = list(zip(hits.field, hits.obj))
    order = sorted(range(len(keys)), key=lambda k: (fjd[keys[k][0]], keys[k][0], keys[k][1]))
    assert order == list(range(len(keys)))


def test_fov_list_overflow_reruns(ctx, workload, monkeypatch):
    """Candidate and n-body lists far too small for the workload: the groups
    run with truncated lists, the overflow is found once after the last
    group, and the call reruns with lists sized from the counts it saw.  The
    detections must be those of the default capacities, bit for bit."""
    xyz, fjd, obs, nrm, h = workload
    batch = ka.StateArray(np.full(xyz.shape[1], syn.EPOCH_JD), xyz)
    ref = fov_state_check(batch, None, dt_limit=3.0, fields=(fjd, obs, nrm))
    monkeypatch.setenv("KH_FOV_LIST_CAP", "8")
    small = fov_state_check(batch, None, dt_limit=3.0, fields=(fjd, obs, nrm))
    assert len(ref.field) > 50
    np.testing.assert_array_equal(small.field, ref.field)
    np.testing.assert_array_equal(small.obj, ref.obj)
    np.testing.assert_array_equal(small.state, ref.state)


@pytest.mark.parametrize("schedule", ["continuous", "kete"])
def test_fov_parity_long_span(ctx, oracle_spk, monkeypatch, schedule):
    """400 days of fields: many small steps and kete's big-step resets
    (checks.rs:63-85).  'continuous' (the default) integrates the states once
    through the small-step epochs; 'kete' restarts at every step as kete
    does.  Both must give kete's visible sets up to the edge tolerance."""
    if schedule == "kete":
        monkeypatch.setenv("KH_FOV_SCHEDULE", "kete")
    xyz, fjd, obs, pts, h = syn.fov_workload(n=800, n_fields=40, span_days=400.0, aimed=0.8,
                                             seed_index=14)
    import oracle
    w = math.radians(6.9)
    nrm = np.array([oracle.rect_normals(p, 0.0, w, w).reshape(12) for p in pts])
    batch = ka.StateArray(np.full(xyz.shape[1], syn.EPOCH_JD), xyz)
    hits = fov_state_check(batch, None, dt_limit=3.0, fields=(fjd, obs, nrm))
    of, oo, os_ = oracle_spk.fov_state_check(syn.EPOCH_JD, xyz, fjd, obs.T.copy(), nrm,
                                             dt_limit=3.0)
    assert len(of) > 20, "workload should produce detections"
    g = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(hits.field, hits.obj))}
    r = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(of, oo))}
    for key in set(g) ^ set(r):
        f, i = key
        st = os_[r[key]] if key in r else hits.state[:, g[key]]
        assert edge_distance(nrm[f], obs[:, f], st) < 1e-9, key
    common = sorted(set(g) & set(r))
    gs = np.array([hits.state[:, g[k]] for k in common])
    rs = np.array([os_[r[k]] for k in common])
    assert np.abs(gs[:, 1:4] - rs[:, 1:4]).max() < 1e-9


def test_radec_vmag_parity(ctx, oracle_spk, workload):
    xyz, fjd, obs, nrm, h = workload
    n = xyz.shape[1]
    batch = ka.StateArray(np.full(n, syn.EPOCH_JD), xyz)
    hits = fov_state_check(batch, None, dt_limit=3.0, fields=(fjd, obs, nrm))
    st = hits.state[1:7]
    ob = obs[:, hits.field]
    out = ra_dec_vmag(ctx, st, ob, h[hits.obj], 0.15)
    c, s = math.cos(0.40909280422232897), math.sin(0.40909280422232897)

    def equ(v):
        return np.array([v[0], c * v[1] - s * v[2], s * v[1] + c * v[2]])

    for k in range(0, len(hits), max(1, len(hits) // 200)):
        ref = oracle_spk.ra_dec_with_rates(equ(st[:3, k]), equ(st[3:, k]), equ(ob[:3, k]),
                                           equ(ob[3:, k]))
        assert np.abs(out[:2, k] - ref[:2]).max() < 1e-12
        assert np.abs(out[2:4, k] - ref[2:4]).max() < 1e-12
        v = oracle_spk.hg_apparent_mag(h[hits.obj[k]], 0.15, st[:3, k], ob[:3, k])
        assert abs(out[4, k] - v) < 1e-9


def test_fov_empty_and_far_fields(ctx, workload):
    xyz, fjd, obs, nrm, h = workload
    batch = ka.StateArray(np.full(10, syn.EPOCH_JD), xyz[:, :10])
    none = fov_state_check(batch, None, fields=(fjd[:0], obs[:, :0], nrm[:0]))
    assert len(none) == 0
    # a field 200 days after the states' epoch exercises the re-propagation
    far = fov_state_check(batch, None, fields=(fjd[:1] + 200.0, obs[:, :1], nrm[:1]))
    assert far.field.shape == far.obj.shape


def test_fov_group_wider_than_one_tile(ctx, oracle_spk):
    """1000 fields within 2*dt_limit of the first: one group of kete's
    schedule (checks.rs:34-57), checked in four LDS tiles of at most 256
    fields.  Fields up to 2.95 d from the states' epoch take the linear
    check, the later ones the two-body margin and the n-body re-check; the
    visible set must equal the oracle's (edge exemption as above)."""
    xyz, fjd, obs, pts, h = syn.fov_workload(n=2000, n_fields=1000, span_days=5.9, aimed=0.25,
                                             seed_index=6)
    w = math.radians(6.9)
    nrm = np.array([oracle_spk.rect_normals(p, 0.0, w, w).reshape(12) for p in pts])
    assert fjd.max() - fjd.min() < 6.0
    n = xyz.shape[1]
    batch = ka.StateArray(np.full(n, syn.EPOCH_JD), xyz)
    print("fov: 2000 states x 1000 fields in one group", flush=True)
    hits = fov_state_check(batch, None, dt_limit=3.0, fields=(fjd, obs, nrm))
    ms, launches, pairs = ctx.last_fov_prefilter()
    assert launches == 4 and pairs == n * 1000 and ms > 0.0
    of, oo, os_ = oracle_spk.fov_state_check(syn.EPOCH_JD, xyz, fjd, obs.T.copy(), nrm,
                                             dt_limit=3.0, cap=100000)
    assert len(of) > 2000
    assert np.any(fjd[of] - syn.EPOCH_JD > 3.0) and np.any(fjd[of] - syn.EPOCH_JD < 3.0)
    g = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(hits.field, hits.obj))}
    r = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(of, oo))}
    for key in set(g) ^ set(r):
        f, i = key
        st = os_[r[key]] if key in r else hits.state[:, g[key]]
        assert edge_distance(nrm[f], obs[:, f], st) < 1e-9, key
    common = sorted(set(g) & set(r))
    gs = np.array([hits.state[:, g[k]] for k in common])
    rs = np.array([os_[r[k]] for k in common])
    assert np.abs(gs[:, 1:4] - rs[:, 1:4]).max() < 1e-9
