"""The oracle's SPK Type 21 evaluation and kete's MASSES force list, pinned
against the Type 21 records of asteroid (42) Isis that kete ships
(docs/data/20000042.bsp, sampled in tests/golden/type21_isis_sample.npz) and
against what the mathematics fixes:

* the closed form of a modified-difference line: position and velocity are
  the reference state plus the double and single integral of the
  acceleration polynomial sum_j DT_j phi_j(tau), phi_j = prod_{i<j}
  (tau + G_{i-1}) / G_i, G_{-1} = 0 -- a wrong pass count, index or sign in
  the W recursion of spk_segments.rs:1045-1100 fails it;
* the reference epoch reproduces the stored state exactly;
* consecutive records of the real file agree at their common epoch, the
  velocity is the derivative of the position, and the acceleration is the
  Sun's gravity up to planetary perturbations;
* the record chosen for a time is the first whose final epoch is not before
  it (spk_segments.rs:1023-1030);
* a two-body orbit written as Type 21 records reads back as that orbit.
"""
import os

import numpy as np
import pytest

import oracle
from kete_amd import synthetic as syn
from kete_amd.spk import Segment, read_spk, type21_array, write_spk

AU_KM = 149597870.7
GMS = 0.00029591220828411956
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = np.load(os.path.join(HERE, "golden", "type21_isis_sample.npz"))


def _sample(k):
    ids = SAMPLE[f"s{k}_ids"]
    return (int(ids[4]), SAMPLE[f"s{k}_records"], SAMPLE[f"s{k}_epochs"],
            SAMPLE[f"s{k}_span_s"])


def test_fixture_file_layout():
    # docs/data/20000042.bsp: seven Type 21 segments of 42 Isis about the Sun,
    # equatorial, MAXDIM 20
    summ = SAMPLE["file_summary"]
    assert int(SAMPLE["n_segments_in_file"]) == 7
    assert (summ[:, 0] == 20000042).all() and (summ[:, 1] == 10).all()
    assert (summ[:, 2] == 1).all() and (summ[:, 3] == 21).all() and (summ[:, 5] == 20).all()
    assert list(summ[:, 4]) == [500, 500, 500, 500, 500, 370, 316]
    for k in range(2):
        n, rec, ep, span = _sample(k)
        assert rec.shape[1] == 4 * n + 11
        assert np.all(np.diff(ep) > 0) and span[0] < ep[0] and span[1] == ep[-1]
        kqmax1 = rec[:, 4 * n + 7]
        kq = rec[:, 4 * n + 8: 4 * n + 11]
        assert np.all(kq <= kqmax1[:, None] - 1) and np.all(kqmax1 <= n + 1)


def _closed_form(rec, n, dt):
    """Reference state + integrals of the acceleration polynomial, exact
    polynomial arithmetic in numpy (km, s)."""
    g = rec[1: n + 1]
    ref = rec[n + 1: n + 7]
    dd = rec[n + 7: 4 * n + 7].reshape(3, n)
    kq = rec[4 * n + 8: 4 * n + 11].astype(int)
    P = np.polynomial.Polynomial
    pos, vel = np.zeros(3), np.zeros(3)
    for c in range(3):
        phi, acc = P([1.0]), P([0.0])
        for j in range(kq[c]):
            acc = acc + dd[c, j] * phi
            if j + 1 < kq[c]:
                gm1 = g[j - 1] if j >= 1 else 0.0
                phi = phi * P([gm1 / g[j], 1.0 / g[j]])
        a1 = acc.integ()
        a2 = a1.integ()
        pos[c] = ref[2 * c] + dt * ref[2 * c + 1] + a2(dt)
        vel[c] = ref[2 * c + 1] + a1(dt)
    return pos / AU_KM, vel / AU_KM * 86400.0


@pytest.mark.parametrize("seed", range(6))
def test_record_closed_form(seed):
    rng = np.random.default_rng(seed)
    n = 20
    for _ in range(20):
        kqmax1 = int(rng.integers(2, 17))
        sign = rng.choice([-1.0, 1.0])
        g = np.zeros(n)
        g[: kqmax1 - 2] = sign * np.cumsum(rng.uniform(0.5, 2.0, kqmax1 - 2))
        rec = np.zeros(4 * n + 11)
        rec[0] = rng.uniform(-1e3, 1e3)
        rec[1: n + 1] = g
        rec[n + 1: n + 7] = rng.normal(size=6) * [3.0, 0.1, 3.0, 0.1, 3.0, 0.1]
        kq = rng.integers(1, kqmax1, size=3)
        for c in range(3):
            rec[n + 7 + c * n: n + 7 + c * n + kq[c]] = rng.normal(size=kq[c]) * 1e-2
        rec[4 * n + 7] = kqmax1
        rec[4 * n + 8: 4 * n + 11] = kq
        for dt in rng.uniform(-1.5, 1.5, 4):
            st, p, v = oracle.spk21_record_eval(rec, n, rec[0] + dt)
            assert st == oracle.OK
            pe, ve = _closed_form(rec, n, dt)
            np.testing.assert_allclose(p, pe, rtol=1e-12, atol=1e-13 / AU_KM)
            np.testing.assert_allclose(v, ve, rtol=1e-12, atol=1e-13 * 86400 / AU_KM)


def test_record_reference_epoch_is_the_stored_state():
    for k in range(2):
        n, rec, _ep, _span = _sample(k)
        for r in rec:
            st, p, v = oracle.spk21_record_eval(r, n, r[0])
            assert st == oracle.OK
            ref = r[n + 1: n + 7]
            assert np.array_equal(p, ref[0::2] / AU_KM)
            assert np.array_equal(v, ref[1::2] / AU_KM * 86400.0)


def test_fixture_records_continuous_and_physical():
    for k in range(2):
        n, rec, ep, span = _sample(k)
        starts = np.concatenate([[span[0]], ep[:-1]])
        for i in range(len(rec)):
            if i + 1 < len(rec):
                _, p1, v1 = oracle.spk21_record_eval(rec[i], n, ep[i])
                _, p2, v2 = oracle.spk21_record_eval(rec[i + 1], n, ep[i])
                assert np.abs(p1 - p2).max() * AU_KM < 2e-5          # km
                assert np.abs(v1 - v2).max() < 1e-11 * np.abs(v1).max()
            js, h = 0.5 * (starts[i] + ep[i]), 30.0
            _, pa, _ = oracle.spk21_record_eval(rec[i], n, js + h)
            _, pb, _ = oracle.spk21_record_eval(rec[i], n, js - h)
            _, pm, vm = oracle.spk21_record_eval(rec[i], n, js)
            vfd = (pa - pb) / (2 * h) * 86400.0
            assert np.abs(vfd - vm).max() < 1e-9 * np.abs(vm).max()
            acc = (pa - 2 * pm + pb) / (h * h) * 86400.0 ** 2
            a_sun = -GMS * pm / np.linalg.norm(pm) ** 3
            assert np.abs(acc - a_sun).max() < 1e-3 * np.abs(a_sun).max()


def test_record_rejects_malformed():
    n, rec, _ep, _span = _sample(0)
    bad = rec[3].copy()
    bad[1] = 0.0  # zero step size (kete: "invalid contents")
    assert oracle.spk21_record_eval(bad, n, bad[0] + 1.0)[0] == oracle.IO_ERROR
    bad = rec[3].copy()
    bad[4 * n + 8] = bad[4 * n + 7]  # order beyond KQMAX1-1
    assert oracle.spk21_record_eval(bad, n, bad[0] + 1.0)[0] == oracle.IO_ERROR


def _write_sample(path):
    segs = []
    for k in range(2):
        n, rec, ep, span = _sample(k)
        segs.append(Segment(20000042, 10, 1, 21, float(span[0]), float(span[1]),
                            type21_array(rec, ep)))
    write_spk(path, segs, name="ISIS SAMPLE")
    return segs


def test_fixture_daf_round_trip_and_record_selection(tmp_path, oracle_spk):
    path = str(tmp_path / "isis.bsp")
    _write_sample(path)
    back = read_spk(path)
    assert [s.seg_type for s in back] == [21, 21]
    rng = np.random.default_rng(5)
    try:
        assert oracle.spk_load(path) == oracle.OK
        for k, s in enumerate(back):
            n, rec, ep, span = _sample(k)
            assert s.n_rec == len(rec) and s.n_coef == n
            assert np.array_equal(s.records(), rec) and np.array_equal(s.epochs(), ep)
            js_all = np.concatenate([rng.uniform(span[0], span[1], 40), ep[:-1],
                                     np.nextafter(ep[:-1], np.inf)])
            for js in js_all:
                jd = js / 86400.0 + 2451545.0
                jsr = (jd - 2451545.0) * 86400.0   # the time kete evaluates at
                if not (span[0] <= jsr <= span[1]):
                    continue
                i = min(int(np.searchsorted(ep, jsr, side="left")), len(ep) - 1)
                st, p, v = oracle.spk_state(20000042, jd, 10, 1)
                assert st == oracle.OK
                _, pe, ve = oracle.spk21_record_eval(rec[i], n, jsr)
                assert np.array_equal(p, pe) and np.array_equal(v, ve)
        # outside every segment: DAFLimits
        st, _, _ = oracle.spk_state(20000042, back[0].jd_start - 1.0, 10, 1)
        assert st == oracle.DAF_LIMITS
    finally:
        assert oracle.spk_load(syn.ephemeris_path()) == oracle.OK


@pytest.fixture(scope="module")
def asteroid_spk(tmp_path_factory):
    return syn.asteroid_ephemeris_path(str(tmp_path_factory.mktemp("ast") / "ast.bsp"))


def test_synthetic_asteroids_follow_their_orbits(asteroid_spk, oracle_spk):
    try:
        assert oracle.spk_load(syn.ephemeris_path(), asteroid_spk) == oracle.OK
        rng = np.random.default_rng(2)
        for aid, (a, e, inc, node, argp, m0) in syn._ASTEROID_ELEMENTS.items():
            jds = np.concatenate([rng.uniform(2456288.5, 2463712.5, 60), [syn.EPOCH_JD]])
            for jd in jds:
                st, p, v = oracle.spk_state(aid, jd, 10, 1)
                assert st == oracle.OK
                mm = np.radians(m0) + np.sqrt(GMS / a ** 3) * (jd - syn.EPOCH_JD)
                pk, vk = syn.elements_to_pos_vel(a, e, np.radians(inc), np.radians(node),
                                                 np.radians(argp), mm, GMS)
                assert np.abs(p - syn._ecl_to_equ(pk)).max() * AU_KM < 0.01
                assert np.abs(v - syn._ecl_to_equ(vk)).max() < 1e-9 * np.abs(vk).max()
    finally:
        assert oracle.spk_load(syn.ephemeris_path()) == oracle.OK


# kete's MASSES beyond PLANETS (constants/gravity.rs:103-134): id, GM/GMS, radius AU
ASTEROIDS = [(20000001, 4.7191659919436e-10, 6.276827e-6),
             (20000002, 1.0259143964958e-10, 3.415824e-6),
             (20000004, 1.3026452498655e-10, 3.512082e-6),
             (20000010, 4.3953391300849e-11, 2.894426e-6),
             (20000704, 1.911e-11, 2.219283e-6)]


def test_masses_adds_five_newtonian_asteroids(asteroid_spk, oracle_spk):
    try:
        assert oracle.spk_load(syn.ephemeris_path(), asteroid_spk) == oracle.OK
        rng = np.random.default_rng(3)
        for _ in range(20):
            t = rng.uniform(2456400.0, 2463600.0)
            x = rng.normal(size=3) * 2.0
            v = rng.normal(size=3) * 0.01
            st, a_pl, _ = oracle.spk_accel(t, x, v)
            st2, a_all, _ = oracle.spk_accel(t, x, v, include_asteroids=True)
            assert st == st2 == oracle.OK
            extra = np.zeros(3)
            for aid, mass, _r in ASTEROIDS:
                _, pa, _ = oracle.spk_state(aid, t, 0, 1)
                d = x - pa
                extra -= d * (mass * GMS * np.linalg.norm(d) ** -3)
            # the sum is rounded at the scale of the full acceleration
            atol = 64 * np.finfo(float).eps * np.abs(a_pl).max()
            np.testing.assert_allclose(a_all - a_pl, extra, rtol=1e-9, atol=atol)
        # an exact evaluation inside an asteroid's radius is an impact with it
        t = syn.EPOCH_JD + 10.0
        _, pc, vc = oracle.spk_state(20000004, t, 0, 1)
        st, _, iid = oracle.spk_accel(t, pc + 1e-6, vc, exact=True, include_asteroids=True)
        assert st == oracle.IMPACT and iid == 20000004
        st, _, _ = oracle.spk_accel(t, pc + 1e-6, vc, exact=True)
        assert st == oracle.OK
    finally:
        assert oracle.spk_load(syn.ephemeris_path()) == oracle.OK


def test_masses_without_asteroid_ephemeris_is_daf_limits(oracle_spk):
    st, _, _ = oracle.spk_accel(syn.EPOCH_JD, [1.0, 0.2, 0.0], [0.0, 0.017, 0.0],
                                include_asteroids=True)
    assert st == oracle.DAF_LIMITS
