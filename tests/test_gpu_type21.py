"""GPU parity of SPK Type 21 and of include_asteroids (kete's MASSES force
list) against the oracle, through the C ABI.

* kh_spk_states on Type 21 objects -- the Isis records kete ships
  (tests/golden/type21_isis_sample.npz) and the five synthetic perturbers --
  against the oracle's exact modified-difference evaluation: 1e-6 km in
  position (the device evaluates each record as the polynomial it is, so
  the two differ by rounding only), 1e-12 relative in velocity.
* propagate_n_body with include_asteroids against the oracle: 1e-9 AU and
  1e-11 AU/day after a year, the tolerance of the planets-only path, and
  the asteroids' effect itself well above that tolerance.
* status semantics: DAFLimits without asteroid segments, an impact with an
  asteroid, and the FOV driver with include_asteroids.
"""
import os

import numpy as np
import pytest

import kete_amd as ka
import oracle
from kete_amd import synthetic as syn
from kete_amd.spk import Segment, type21_array

pytestmark = pytest.mark.gpu

AU_KM = 149597870.7
JD0 = syn.EPOCH_JD
POS_TOL = 1e-9
VEL_TOL = 1e-11
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = np.load(os.path.join(HERE, "golden", "type21_isis_sample.npz"))
ASTEROIDS = [20000001, 20000002, 20000004, 20000010, 20000704]


@pytest.fixture
def oracle_ast(oracle_spk):
    assert oracle.spk_load(syn.ephemeris_path(), syn.asteroid_ephemeris_path()) == oracle.OK
    yield oracle
    assert oracle.spk_load(syn.ephemeris_path()) == oracle.OK


def isis_segments():
    segs = []
    for k in range(2):
        rec, ep, span = SAMPLE[f"s{k}_records"], SAMPLE[f"s{k}_epochs"], SAMPLE[f"s{k}_span_s"]
        segs.append(Segment(20000042, 10, 1, 21, float(span[0]), float(span[1]),
                            type21_array(rec, ep)))
    return segs


def compare_states(ctx, obj_id, jds, center, frame):
    g, gst = ctx.spk_state(obj_id, jds, center=center, frame=frame)
    for k, jd in enumerate(jds):
        st, p, v = oracle.spk_state(obj_id, jd, center, frame)
        assert gst[k] == st, (jd, gst[k], st)
        if st:
            assert np.all(np.isnan(g[:, k]))
            continue
        assert np.abs(g[:3, k] - p).max() * AU_KM < 1e-6, (obj_id, jd)
        assert np.abs(g[3:, k] - v).max() < 1e-12 * np.abs(v).max(), (obj_id, jd)


def test_isis_records_on_device(oracle_spk, tmp_path):
    path = str(tmp_path / "isis.bsp")
    from kete_amd.spk import write_spk
    write_spk(path, isis_segments(), name="ISIS SAMPLE")
    ctx = ka.Context(0)
    try:
        ctx.load_spk(syn.ephemeris_path(), path)
        assert oracle.spk_load(syn.ephemeris_path(), path) == oracle.OK
        rng = np.random.default_rng(11)
        jds = []
        for k in range(2):
            ep, span = SAMPLE[f"s{k}_epochs"], SAMPLE[f"s{k}_span_s"]
            js = np.concatenate([rng.uniform(span[0], span[1], 300), ep,
                                 np.nextafter(ep, np.inf), [span[0] - 3600.0, span[1] + 3600.0]])
            jds.append(js / 86400.0 + 2451545.0)
        jds = np.concatenate(jds)
        for center, frame in [(10, 1), (0, 1), (10, 17)]:
            compare_states(ctx, 20000042, jds, center, frame)
    finally:
        ctx.close()
        assert oracle.spk_load(syn.ephemeris_path()) == oracle.OK


def test_synthetic_asteroids_on_device(ctx, oracle_ast):
    rng = np.random.default_rng(12)
    jds = np.concatenate([rng.uniform(2456288.5, 2463712.5, 400),
                          [JD0, np.nextafter(JD0, 0), np.nextafter(JD0, np.inf),
                           2456288.5, 2463712.5, 2456288.0, 2463713.0]])
    for aid in ASTEROIDS:
        compare_states(ctx, aid, jds, 0, 1)
    compare_states(ctx, ASTEROIDS[0], jds, 10, 17)


@pytest.fixture(scope="module")
def belt():
    # main-belt and near-Earth orbits; the first objects start next to Ceres
    # and Vesta, where the asteroids' pull is largest
    xyz = syn.n_body_workload(n=600, neo_fraction=0.2, seed_index=7)
    for k, aid in enumerate([20000001, 20000004]):
        assert oracle.spk_load(syn.ephemeris_path(), syn.asteroid_ephemeris_path()) == 0
        _, p, v = oracle.spk_state(aid, JD0, 10, 17)
        assert oracle.spk_load(syn.ephemeris_path()) == 0
        off = np.array([0.02, -0.01, 0.005]) * (k + 1)
        xyz[:3, k] = p + off
        xyz[3:, k] = v + 1e-4
    return xyz


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_include_asteroids_parity(ctx, oracle_ast, belt, mode):
    n = belt.shape[1]
    batch = ka.StateArray(np.full(n, JD0), belt)
    jf = JD0 + 365.25
    g = ctx.n_body(batch, jf, mode=mode, include_asteroids=True)
    r = oracle_ast.batch_n_body(JD0, belt, jf, include_asteroids=True)
    assert np.array_equal(g["status"], r["status"])
    assert np.all(g["status"] == 0)
    dp = np.abs(g["out"][:3] - r["out"][:3]).max(axis=0)
    dv = np.abs(g["out"][3:] - r["out"][3:]).max(axis=0)
    assert dp.max() < POS_TOL, (dp.max(), int(dp.argmax()))
    assert dv.max() < VEL_TOL, dv.max()
    if mode == 0:
        # the asteroids matter: planets-only states differ far beyond the tolerance
        p = ctx.n_body(batch, jf, mode=mode)
        effect = np.abs(p["out"][:3] - g["out"][:3]).max(axis=0)
        assert effect[:2].min() > 100 * POS_TOL, effect[:2]


# Object 827773 of the 1e6-object headline workload (n_body_workload(n=1e6,
# seed_index=2)): a main-belt orbit that passes 0.018 AU from Ceres on day
# ~172.  kete's integration at epsilon 1e-6 steps over the encounter (a
# 67-day step from day 131) and lands 2.4e-7 AU from the converged answer;
# the convergence test of one step (radau.rs:313-319) is decided by the last
# bits of the force there, so a second implementation takes one more step
# and ends 2.8e-7 AU from kete -- and closer to the converged answer.
CERES_ENCOUNTER = np.array([-2.134568357213274, -2.147989090389715, -0.2926967667858866,
                            0.007870867058782137, -0.006058316429118384, 0.0015975507635635556])


def test_include_asteroids_ceres_encounter(ctx, oracle_ast):
    xyz = CERES_ENCOUNTER[:, None]
    jf = JD0 + 365.25
    g = ctx.n_body(ka.StateArray(np.array([JD0]), xyz), jf, include_asteroids=True)
    r = oracle_ast.batch_n_body(JD0, xyz, jf, include_asteroids=True)
    old = oracle_ast.radau_epsilon(1e-10)
    try:
        conv = oracle_ast.batch_n_body(JD0, xyz, jf, include_asteroids=True)
        oracle_ast.radau_epsilon(1e-11)
        conv2 = oracle_ast.batch_n_body(JD0, xyz, jf, include_asteroids=True)
    finally:
        oracle_ast.radau_epsilon(old)
    assert g["status"][0] == 0 and r["status"][0] == 0
    # the converged answer is converged: epsilon 1e-10 and 1e-11 agree
    assert np.abs(conv["out"][:3] - conv2["out"][:3]).max() < 1e-10
    err_gpu = np.abs(g["out"][:3, 0] - conv["out"][:3, 0]).max()
    err_kete = np.abs(r["out"][:3, 0] - conv["out"][:3, 0]).max()
    # kete's own error here is far above the 1e-9 AU bar, and the device's
    # result is no farther from the converged answer than kete's
    assert err_kete > 1e-8, err_kete
    assert err_gpu <= err_kete + POS_TOL, (err_gpu, err_kete)


def test_include_asteroids_backwards_across_segments(ctx, oracle_ast, belt):
    # backwards from the epoch: the earlier (backward-laid) segment of each asteroid
    xyz = belt[:, :200]
    n = xyz.shape[1]
    batch = ka.StateArray(np.full(n, JD0 + 40.0), xyz)
    jf = JD0 - 300.0
    g = ctx.n_body(batch, jf, include_asteroids=True)
    r = oracle_ast.batch_n_body(JD0 + 40.0, xyz, jf, include_asteroids=True)
    assert np.array_equal(g["status"], r["status"])
    ok = g["status"] == 0
    assert ok.all()
    assert np.abs(g["out"][:3] - r["out"][:3]).max() < POS_TOL


def test_include_asteroids_without_segments_is_daf_limits(oracle_spk, belt):
    ctx = ka.Context(0)
    try:
        ctx.load_spk(syn.ephemeris_path())
        batch = ka.StateArray(np.full(8, JD0), belt[:, :8])
        g = ctx.n_body(batch, JD0 + 10.0, include_asteroids=True)
        r = oracle_spk.batch_n_body(JD0, belt[:, :8], JD0 + 10.0, include_asteroids=True)
        assert np.all(r["status"] == oracle.DAF_LIMITS)
        assert np.array_equal(g["status"], r["status"])
        assert np.all(np.isnan(g["out"]))
        # planets only still works on that context
        assert np.all(ctx.n_body(batch, JD0 + 10.0)["status"] == 0)
    finally:
        ctx.close()


@pytest.mark.parametrize("mode", [0, 2])
def test_impact_with_an_asteroid(ctx, oracle_ast, mode):
    _, pv, vv = oracle_ast.spk_state(20000004, JD0, 10, 17)
    xyz = np.zeros((6, 3))
    xyz[:3, 0], xyz[3:, 0] = pv + 1e-6, vv          # inside Vesta's radius (3.5e-6 AU)
    xyz[:3, 1], xyz[3:, 1] = pv + 1e-3, vv + 1e-5   # a near miss
    xyz[:3, 2], xyz[3:, 2] = [2.5, 0.3, 0.1], [0.0, 0.0105, 0.0]
    batch = ka.StateArray(np.full(3, JD0), xyz)
    g = ctx.n_body(batch, JD0 + 20.0, mode=mode, include_asteroids=True)
    r = oracle_ast.batch_n_body(JD0, xyz, JD0 + 20.0, include_asteroids=True)
    assert np.array_equal(g["status"], r["status"])
    assert g["status"][0] == oracle.IMPACT and g["impact_id"][0] == 20000004
    assert r["impact_id"][0] == 20000004
    assert g["impact_jd"][0] == r["impact_jd"][0]
    # planets only: no impact
    p = ctx.n_body(batch, JD0 + 20.0, mode=mode)
    assert p["status"][0] == 0


def test_fov_include_asteroids(ctx, oracle_ast):
    """fov_state_check with include_asteroids (checks.rs:25-31): the same
    visible sets as the oracle, up to objects within 1e-9 rad of an edge."""
    import math

    from kete_amd.fov import fov_state_check
    xyz, fjd, obs, pts, _h = syn.fov_workload(n=2000, n_fields=16, span_days=40.0, aimed=0.75,
                                              seed_index=9)
    w = math.radians(6.9)
    nrm = np.array([oracle.rect_normals(p, 0.0, w, w).reshape(12) for p in pts])
    n = xyz.shape[1]
    hits = fov_state_check(ka.StateArray(np.full(n, JD0), xyz), None, include_asteroids=True,
                           ctx=ctx, fields=(fjd, obs, nrm))
    of, oo, os_ = oracle_ast.fov_state_check(JD0, xyz, fjd, obs.T.copy(), nrm,
                                             include_asteroids=True)
    assert len(of) > 30
    g = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(hits.field, hits.obj))}
    r = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(of, oo))}
    for key in set(g) ^ set(r):
        f, i = key
        st = os_[r[key]] if key in r else hits.state[:, g[key]]
        rel = st[1:4] - obs[:3, f]
        u = rel / np.linalg.norm(rel)
        nn = nrm[f].reshape(4, 3)
        assert np.min(np.abs(nn @ u) / np.linalg.norm(nn, axis=1)) < 1e-9, key
    common = sorted(set(g) & set(r))
    gs = np.array([hits.state[:, g[k]] for k in common])
    rs = np.array([os_[r[k]] for k in common])
    assert np.abs(gs[:, 1:4] - rs[:, 1:4]).max() < 1e-9


def test_fov_spk_check(ctx, oracle_ast):
    """fov_spk_check (checks.rs:123-142): the Type 21 asteroids and planet
    barycentres in fields aimed at them and at random, against the oracle:
    the same (field, object) pairs and states to 1e-9 AU."""
    import math

    from kete_amd.fov import fov_spk_check
    ids = [20000001, 20000002, 20000004, 20000010, 20000704, 4, 5, 6, 2099999999]
    rng = np.random.default_rng(21)
    w = math.radians(6.9)
    fjd, obs, nrm = [], [], []
    for jd in np.sort(rng.uniform(2456400.0, 2463600.0, 40)):
        _, op, ov = oracle_ast.spk_state(399, jd, 10, 17)
        target = ids[rng.integers(0, 8)]
        if rng.uniform() < 0.7:
            _, p, _ = oracle_ast.spk_state(target, jd, 10, 17)
            d = p - op
        else:
            d = rng.normal(size=3)
        fjd.append(jd)
        obs.append(np.concatenate([op, ov]))
        nrm.append(oracle_ast.rect_normals(d / np.linalg.norm(d), 0.3, w, w).reshape(12))
    fjd, obs, nrm = np.array(fjd), np.array(obs), np.array(nrm)
    # fields handed over out of epoch order: the driver sorts them as kete does
    perm = rng.permutation(len(fjd))
    hits = fov_spk_check(ids, None, ctx=ctx, fields=(fjd[perm], obs[perm].T.copy(), nrm[perm]))
    hf, ho, hs = oracle_ast.fov_spk_check(ids, fjd, obs, nrm)
    assert len(hf) >= 25
    got = [(int(perm[f]), int(o)) for f, o in zip(hits.field, hits.obj)]
    want = list(zip(hf.tolist(), ho.tolist()))
    assert got == want
    assert np.abs(hits.state[1:4].T - hs[:, 1:4]).max() < 1e-9
    assert np.abs(hits.state[0] - hs[:, 0]).max() < 1e-12
