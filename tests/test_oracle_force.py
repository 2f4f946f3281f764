"""Pins of the oracle's force terms and state changes.

- GravParams::add_acceleration (constants/gravity.rs:171-253) through the
  orbital effects each term causes, integrated with kete's Radau around one
  body at rest (the oracle's one-body test hook):
    * Newton alone (Mars, which has no special terms) keeps a Kepler orbit:
      the state after whole periods is the initial one;
    * the Sun's GR term advances a Mercury-like perihelion by
      6 pi GM / (c^2 a (1 - e^2)) per orbit (the 1PN result its equation
      gives; the Sun's J2 adds 6e-5 of that);
    * Jupiter's J2 term, applied in the ecliptic frame, turns the node of an
      orbit inclined to the ecliptic at -3/2 n J2 (R/p)^2 cos i;
    * the Earth's J2 term, applied in the equatorial frame, turns the node
      about the equatorial pole at the same rate.
  A dropped factor, a wrong sign, or J2 in the wrong frame fails one of these.
- Frames and centers (state.rs:357-460): equatorial <-> ecliptic round trips
  to 10 eps, the ecliptic pole's equatorial direction (0, -sin e, cos e),
  and Sun -> SSB -> Sun through the ephemeris to 10 eps.
"""
import numpy as np
import pytest

import oracle

GMS = 0.00029591220828411956           # constants/gravity.rs:9
C_AU_DAY = 173.14463267424034           # constants/universal.rs:11
EPS_OBL = 0.40909280422232897           # frames/definitions.rs obliquity
# constants/gravity.rs:23,27,40,65-66,83-84,91-92
SUN_J2, EARTH_J2, JUPITER_J2 = 2.2e-7, 0.00108262998905, 0.014696572
GM_EARTH, R_EARTH = 3.00348961546514e-06 * GMS, 4.33036684926559e-05
GM_JUP, R_JUP = 0.0009547919099414246 * GMS, 0.00047789450254521576
GM_MARS = 3.22715608291416e-07 * GMS


def ecl_to_equ(v):
    c, s = np.cos(EPS_OBL), np.sin(EPS_OBL)
    return np.array([v[0], c * v[1] - s * v[2], s * v[1] + c * v[2]])


def equ_to_ecl(v):
    c, s = np.cos(EPS_OBL), np.sin(EPS_OBL)
    return np.array([v[0], c * v[1] + s * v[2], -s * v[1] + c * v[2]])


def orbit(gm, a, e, incl, node, phase_peri=True):
    """Perihelion state of an orbit with node `node` and inclination `incl`
    (argument of perihelion 0), in the reference plane's frame."""
    r = a * (1 - e)
    v = np.sqrt(gm * (1 + e) / r)
    p = np.array([r, 0.0, 0.0])
    u = np.array([0.0, v, 0.0])
    ci, si, cn, sn = np.cos(incl), np.sin(incl), np.cos(node), np.sin(node)
    m = np.array([[cn, -sn, 0], [sn, cn, 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, ci, -si],
                                                                     [0, si, ci]])
    return m @ p, m @ u


def node_angle(p, v):
    h = np.cross(p, v)
    return np.arctan2(h[0], -h[1])


def peri_angle(p, v, gm):
    ev = np.cross(v, np.cross(p, v)) / gm - p / np.linalg.norm(p)
    return np.arctan2(ev[1], ev[0])


def test_newton_term_keeps_kepler_orbits():
    gm, a, e = GM_MARS, 2e-4, 0.3
    p0, v0 = orbit(gm, a, e, 0.4, 1.0)
    period = 2 * np.pi * np.sqrt(a ** 3 / gm)
    st, p, v = oracle.radau_one_body(4, 0.0, 5 * period, p0, v0)
    assert st == 0
    assert np.abs(p - p0).max() < 1e-9 * a and np.abs(v - v0).max() < 1e-9 * np.linalg.norm(v0)


def test_sun_gr_perihelion_advance():
    a, e, n_orb = 0.387098, 0.20563, 10
    p0, v0 = orbit(GMS, a, e, 0.0, 0.0)           # in the ecliptic plane
    period = 2 * np.pi * np.sqrt(a ** 3 / GMS)
    st, p, v = oracle.radau_one_body(10, 0.0, n_orb * period, ecl_to_equ(p0), ecl_to_equ(v0))
    assert st == 0
    dw = peri_angle(equ_to_ecl(p), equ_to_ecl(v), GMS) - peri_angle(p0, v0, GMS)
    expect = n_orb * 6 * np.pi * GMS / (C_AU_DAY ** 2 * a * (1 - e * e))
    assert dw > 0 and abs(dw / expect - 1) < 2e-3, (dw, expect)


@pytest.mark.parametrize("body", ["jupiter", "earth"])
def test_j2_nodal_precession(body):
    if body == "jupiter":   # J2 about the ecliptic pole
        naif, gm, rad, j2, to_equ, to_ref = 5, GM_JUP, R_JUP, JUPITER_J2, ecl_to_equ, equ_to_ecl
    else:                   # J2 about the equatorial pole
        naif, gm, rad, j2 = 399, GM_EARTH, R_EARTH, EARTH_J2
        to_equ = to_ref = lambda x: x
    a, e, incl = 12 * rad, 0.01, 0.5
    p0, v0 = orbit(gm, a, e, incl, 0.3)
    n = np.sqrt(gm / a ** 3)
    t1 = 400 * 2 * np.pi / n
    st, p, v = oracle.radau_one_body(naif, 0.0, t1, to_equ(p0), to_equ(v0))
    assert st == 0
    dnode = node_angle(to_ref(p), to_ref(v)) - node_angle(p0, v0)
    dnode = (dnode + np.pi) % (2 * np.pi) - np.pi
    pp = a * (1 - e * e)
    expect = -1.5 * n * j2 * (rad / pp) ** 2 * np.cos(incl) * t1
    assert abs(dnode / expect - 1) < 1e-2, (dnode, expect)


def test_frame_round_trips_and_pole():
    rng = np.random.default_rng(3)
    for _ in range(50):
        p, v = rng.normal(size=3), rng.normal(size=3)
        for f0, f1 in ((17, 1), (1, 17)):
            st, p1, v1 = oracle.change_frame(0.0, p, v, f0, 0, f1)
            st2, p2, v2 = oracle.change_frame(0.0, p1, v1, f1, 0, f0)
            assert st == st2 == 0
            assert np.abs(p2 - p).max() < 10 * np.finfo(float).eps * 4
            assert np.abs(v2 - v).max() < 10 * np.finfo(float).eps * 4
    st, pole, _ = oracle.change_frame(0.0, [0.0, 0.0, 1.0], [0.0, 0.0, 0.0], 17, 0, 1)
    assert st == 0 and np.allclose(pole, [0.0, -np.sin(EPS_OBL), np.cos(EPS_OBL)], atol=1e-16)
    assert oracle.change_frame(0.0, [1, 0, 0], [1, 0, 0], 17, 0, 10000)[0] == oracle.UNKNOWN_FRAME


def test_center_round_trip(oracle_spk):
    jd = 2460100.5
    p, v = np.array([1.5, -0.3, 0.2]), np.array([0.001, 0.012, -0.002])
    st, p1, v1 = oracle.change_center(jd, p, v, 17, 10, 0)
    assert st == 0 and np.abs(p1 - p).max() > 1e-4
    st, sp, sv = oracle.spk_state(10, jd, 0, 17)
    assert np.allclose(p1, p + sp, atol=1e-15) and np.allclose(v1, v + sv, atol=1e-17)
    st, p2, v2 = oracle.change_center(jd, p1, v1, 17, 0, 10)
    assert st == 0
    assert np.abs(p2 - p).max() < 10 * np.finfo(float).eps * 2
    assert np.abs(v2 - v).max() < 10 * np.finfo(float).eps * 0.02


def test_dust_beta_one_moves_linearly(oracle_spk):
    """kete's test_a_terms (src/tests/test_propagation.py:61-70): a dust grain
    with beta = 1 feels no net solar gravity, so over 10 days it moves on
    its initial velocity to 1e-5 AU (the planets and the Poynting-Robertson
    term are all that is left); without the model it does not."""
    from kete_amd import synthetic as syn
    x0 = np.array([[10.0], [0.0], [0.0], [0.0], [0.01], [0.0]])
    dust = [oracle.NonGrav(kind=2, beta=1.0)]
    r = oracle.batch_n_body(syn.EPOCH_JD, x0, syn.EPOCH_JD + 10.0, nongravs=dust)
    assert r["status"][0] == 0
    lin = x0[:3, 0] + x0[3:, 0] * 10.0
    assert np.allclose(r["out"][:3, 0], lin, atol=1e-5)
    r0 = oracle.batch_n_body(syn.EPOCH_JD, x0, syn.EPOCH_JD + 10.0)
    assert np.abs(r0["out"][:3, 0] - lin).max() > 1e-4
