"""Pins for the oracle (the CPU restatement of kete's path).

Each test is one of kete's own known-answer tests, restated with its numbers
(file:line given), or a closed form / invariant chosen so that a plausible
mistake in the restatement (a dropped term, a wrong sign, a transposed
operand) fails it.
"""
import math

import numpy as np
import pytest

import oracle as o

GMS = 0.00029591220828411956
GMS_SQRT = 0.01720209894996
TAU = 2 * math.pi
C_AU_PER_DAY = 173.14463267424034


def test_constants():
    assert o.const("GMS") == GMS
    assert o.const("GMS_SQRT") == GMS_SQRT
    assert o.const("C_AU_PER_DAY") == C_AU_PER_DAY


def test_newton_square_root():
    # fitting/newton.rs:72-78
    st, root = o.newton_square_root(0.0, 1e-10)
    assert st == 0 and abs(root - 1.0) < 1e-12


@pytest.mark.parametrize("r", [0.2, 0.5, 1.0, 2.0])
def test_kepler_circular(r):
    # propagation/kepler.rs:416-430
    pos = [0.0, r, 0.0]
    vel = [-GMS_SQRT / math.sqrt(r), 0.0, 0.0]
    year = TAU / GMS_SQRT * r ** 1.5
    for dt in (year, -year):
        st, p, v = o.analytic_2_body(dt, pos, vel)
        assert st == 0
        assert np.linalg.norm(p - pos) < 1e-8
        assert np.linalg.norm(v - vel) < 1e-8


def test_kepler_parabolic():
    # propagation/kepler.rs:445-459
    pos = np.array([0.0, 2.0, 0.0])
    vel = np.array([GMS_SQRT, 0.0, 0.0])
    year = -TAU / GMS_SQRT
    st, p, v = o.analytic_2_body(year, pos, vel)
    pos_exp = np.array([-4.448805955479905, -0.4739843046525608, 0.0])
    vel_exp = -np.array([-0.00768983428326951, -0.008552645144187791, 0.0])
    assert st == 0
    assert np.linalg.norm(p - pos_exp) < 1e-8
    assert np.linalg.norm(v - vel_exp) < 1e-8
    st, p2, v2 = o.analytic_2_body(-year, pos_exp, vel_exp)
    assert np.linalg.norm(p2 - pos) < 1e-8 and np.linalg.norm(v2 - vel) < 1e-8


def test_kepler_hyperbolic():
    # propagation/kepler.rs:462-476
    pos = np.array([0.0, 3.0, 0.0])
    vel = np.array([-GMS_SQRT, 0.0, 0.0])
    year = TAU / GMS_SQRT
    st, p, v = o.analytic_2_body(year, pos, vel)
    pos_exp = np.array([-5.556785268950049, 1.6076633958058089, 0.0])
    vel_exp = np.array([-0.013061655543084886, -0.005508140023183166, 0.0])
    assert st == 0
    assert np.linalg.norm(p - pos_exp) < 1e-8
    assert np.linalg.norm(v - vel_exp) < 1e-8
    st, p2, v2 = o.analytic_2_body(-year, pos_exp, vel_exp)
    assert np.linalg.norm(p2 - pos) < 1e-8 and np.linalg.norm(v2 - vel) < 1e-8


def test_eccentric_anomaly_hyperbolic_converges():
    # propagation/kepler.rs:433-442
    for m in range(-100, 100):
        st, _ = o.eccentric_anomaly(2.0, float(m), 0.1)
        assert st == 0


def test_eccentric_anomaly_solves_kepler_equation():
    for e in (0.0001, 0.1, 0.5, 0.9):
        for m in np.linspace(0.01, 6.2, 25):
            st, E = o.eccentric_anomaly(e, m, 0.3)
            assert st == 0
            assert abs(E - e * math.sin(E) - m) < 1e-10


def _energy_h(p, v):
    return 0.5 * v @ v - GMS / np.linalg.norm(p), np.cross(p, v)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_kepler_invariants_and_reversal(seed):
    """Energy, angular momentum, time reversal, and the Lagrange-coefficient
    Wronskian det[[f, g], [fd, gd]] = 1 over elliptic and hyperbolic orbits."""
    rng = np.random.default_rng(seed)
    for _ in range(40):
        r = rng.uniform(0.3, 5.0)
        p = rng.normal(size=3)
        p *= r / np.linalg.norm(p)
        vc = math.sqrt(GMS / r)
        v = rng.normal(size=3)
        v *= rng.uniform(0.3, 1.6) * vc / np.linalg.norm(v)
        dt = rng.uniform(-800, 800)
        st, p1, v1 = o.analytic_2_body(dt, p, v)
        assert st == 0
        e0, h0 = _energy_h(p, v)
        e1, h1 = _energy_h(p1, v1)
        assert abs(e1 - e0) < 1e-12 * max(1.0, abs(e0)) + 1e-15
        assert np.linalg.norm(h1 - h0) < 1e-12
        st, p2, v2 = o.analytic_2_body(-dt, p1, v1)
        assert np.linalg.norm(p2 - p) < 1e-9 * max(1, np.linalg.norm(p))
        # f, g from the two basis directions: p1 = f p + g v, v1 = fd p + gd v
        M = np.stack([p, v], axis=1)
        fg, *_ = np.linalg.lstsq(M, p1, rcond=None)
        fdgd, *_ = np.linalg.lstsq(M, v1, rcond=None)
        assert abs(fg[0] * fdgd[1] - fg[1] * fdgd[0] - 1.0) < 1e-8


def test_radau_central_known_answer():
    # propagation/radau.rs:356-374
    st, p, v, cnt = o.radau_central(0.0, 1000.0, [0.46937657, -0.8829981, 0.0],
                                    [0.01518942, 0.00807426, 0.0])
    assert st == 0
    assert abs(p[0] + 0.916350120888658) < 1e-8
    assert abs(p[1] + 0.4003771936559588) < 1e-8
    assert p[2] == 0.0
    assert abs(v[0] - 0.006887328686018099) < 1e-8
    assert abs(v[1] + 0.01576315407302832) < 1e-8
    assert v[2] == 0.0
    assert cnt.n_acc > 0 and cnt.n_sweep >= cnt.n_acc


def test_radau_matches_kepler():
    """propagate_two_body_radau (propagation/mod.rs:35-46) is kete's own
    Radau-vs-Kepler check; agreement is at the 1e-7 AU level over 300 d."""
    rng = np.random.default_rng(7)
    for _ in range(10):
        a = rng.uniform(0.8, 4.0)
        p = np.array([a, 0.0, 0.0])
        v = np.array([0.0, math.sqrt(GMS / a) * rng.uniform(0.9, 1.1), 0.001])
        st, pr, vr, _ = o.radau_central(0.0, 300.0, p, v)
        st2, pk, vk = o.analytic_2_body(300.0, p, v)
        assert st == 0 and st2 == 0
        assert np.linalg.norm(pr - pk) < 1e-7


def test_radau_epsilon_setting():
    """A tighter Radau tolerance (the labelled epsilon-1e-9 run) takes more
    steps, about (1e3)^(1/7) = 2.7x (the step scales as epsilon^(1/7)), and
    stays on the Kepler solution; kete's 1e-6 is restored afterwards."""
    p, v = np.array([1.2, 0.0, 0.0]), np.array([0.0, 0.019, 0.001])
    st2, pk, _ = o.analytic_2_body(300.0, p, v)
    _, p6, _, c6 = o.radau_central(0.0, 300.0, p, v)
    old = o.radau_epsilon(1e-9)
    try:
        assert old == 1e-6
        _, p9, _, c9 = o.radau_central(0.0, 300.0, p, v)
    finally:
        o.radau_epsilon(old)
    assert o.radau_epsilon() == 1e-6
    assert 2.0 < c9.n_acc / c6.n_acc < 3.5
    assert np.linalg.norm(p9 - pk) < 1e-12 and np.linalg.norm(p6 - pk) < 1e-12
    _, p6b, _, _ = o.radau_central(0.0, 300.0, p, v)
    assert np.array_equal(p6b, p6)


def test_chebyshev_closed_forms():
    """T_k(cos t) = cos kt and dT_k/dx = k U_{k-1} = k sin kt / sin t."""
    n = 14
    for k in range(n):
        c = np.zeros(n)
        c[k] = 1.0
        for th in np.linspace(0.1, 3.0, 7):
            x = math.cos(th)
            st, val, der = o.chebyshev_both(x, c, 2 * c, -c)
            assert st == 0
            assert abs(val[0] - math.cos(k * th)) < 1e-13
            assert abs(val[1] - 2 * math.cos(k * th)) < 1e-13
            assert abs(val[2] + math.cos(k * th)) < 1e-13
            d = k * math.sin(k * th) / math.sin(th)
            assert abs(der[0] - d) < 1e-11 * max(1, abs(d))
    st, _, _ = o.chebyshev_both(0.3, [1.0], [1.0], [1.0])
    assert st == 5  # IOError for fewer than 2 coefficients (interpolation.rs:33-38)


def test_chebyshev_reproduces_polynomial():
    rng = np.random.default_rng(3)
    c = rng.normal(size=14)
    xs = np.linspace(-1, 1, 17)
    ref = np.polynomial.chebyshev.chebval(xs, c)
    dref = np.polynomial.chebyshev.chebval(xs, np.polynomial.chebyshev.chebder(c))
    for x, r, dr in zip(xs, ref, dref):
        st, val, der = o.chebyshev_both(x, c, c, c)
        assert abs(val[0] - r) < 1e-12
        assert abs(der[0] - dr) < 1e-10


def _rect_pointing_and_widths(nrm):
    # SphericalPolygon::pointing, lat_width, lon_width (fov/patches.rs:187-198,258-281)
    acc = np.zeros(3)
    for i in range(4):
        c = np.cross(nrm[i], nrm[(i + 1) % 4])
        acc += c / np.linalg.norm(c)
    p = acc / np.linalg.norm(acc)

    def angle(a, b):
        return math.acos(np.clip(a @ b / np.linalg.norm(a) / np.linalg.norm(b), -1, 1))

    return p, 2 * (math.pi / 2 - angle(p, nrm[1])), 2 * (math.pi / 2 - angle(p, nrm[0]))


def test_rectangular_patch():
    # fov/patches.rs:366-385
    rot = math.radians(45)
    inside = [1.0, 0.01, 0.01]
    outside = [1.0, 0.1, 0.0]
    ji = [1.0, math.sin(0.05) * 0.99, math.sin(0.05) * 0.99]
    jo = [1.0, math.sin(0.05) * 1.01, math.sin(0.05) * 1.01]
    fov = o.rect_normals([1.0, 0, 0], 0.0, 0.1, 0.1)
    fov_rot = o.rect_normals([1.0, 0, 0], rot, 0.1, 0.1)
    assert o.rect_contains(fov, inside)[0]
    assert o.rect_contains(fov, ji)[0]
    assert not o.rect_contains(fov, outside)[0]
    assert not o.rect_contains(fov, jo)[0]
    assert o.rect_contains(fov_rot, inside)[0]
    assert not o.rect_contains(fov_rot, ji)[0]
    p, _, _ = _rect_pointing_and_widths(fov_rot)
    assert np.linalg.norm(p - [1, 0, 0]) < 1e-10


def test_rectangular_patch_latlon():
    # fov/patches.rs:387-396
    rot = math.radians(45)
    for r in (0.0, rot):
        nrm = o.rect_normals([1.0, 0, 0], r, 0.1, 0.2)
        _, lat_w, lon_w = _rect_pointing_and_widths(nrm)
        assert abs(lat_w - 0.2) < 1e-10
        assert abs(lon_w - 0.1) < 1e-10


def test_rect_contains_distance_and_nan():
    nrm = o.rect_normals([0.0, 1.0, 0.0], 0.0, 0.2, 0.2)
    inside, d = o.rect_contains(nrm, [0.0, 1.0, 0.0])
    assert inside
    # outside along +x by angle a: the signed distance is -sin(a - 0.1) of the
    # edge normal, capped by |v|
    inside, d = o.rect_contains(nrm, [math.sin(0.3), math.cos(0.3), 0.0])
    assert not inside and abs(d - math.sin(0.2)) < 1e-12
    inside, d = o.rect_contains(nrm, [float("nan"), 1.0, 0.0])
    assert not inside and math.isnan(d)


@pytest.mark.parametrize("g", np.arange(-0.5, 1, 0.1))
def test_hg_zero_phase(g):
    # tests/flux/test_common.py:84-90
    assert abs(o.hg_phase_curve_correction(g, 0.0) - 1.0) < 1e-12


def test_hg_apparent_mag_geometry():
    """At zero phase V = H + 5 log10(r * delta); beyond 160 degrees it is NaN
    (flux/reflected.rs:255-268)."""
    v = o.hg_apparent_mag(15.0, 0.15, [2.0, 0, 0], [1.0, 0, 0])
    assert abs(v - (15.0 + 5 * math.log10(2.0 * 1.0))) < 1e-12
    assert math.isnan(o.hg_apparent_mag(15.0, 0.15, [1.0, 0, 0], [2.0, 0.1, 0]))


def test_ra_dec_closed_form():
    # an object straight along +y of the observer, equatorial: RA 90 deg, Dec 0
    out = o.ra_dec_with_rates([0, 2.0, 0], [0.01, 0, 0.001], [0, 1.0, 0], [0, 0, 0])
    assert abs(out[0] - math.pi / 2) < 1e-14
    assert abs(out[1]) < 1e-14
    # d(RA)/dt = (x vy - y vx)/(x^2+y^2) = -0.01; scaled by cos(dec)=1
    assert abs(out[2] - (-0.01)) < 1e-14
    # d(Dec)/dt = vz / rho at dec 0
    assert abs(out[3] - 0.001) < 1e-14
    out = o.ra_dec_with_rates([0, 0, 5.0], [0, 0, 0], [0, 0, 0], [0, 0, 0])
    assert abs(out[1] - math.pi / 2) < 1e-14


def test_two_body_light_delay_identity(oracle_spk):
    """src/tests/test_propagation.py:89-105: propagate_two_body with an
    observer `au` AU away equals propagation to jd - au/c (synthetic Venus)."""
    jd = 2460000.5
    st, pos, vel = o.spk_state(2, jd, 10, o.FRAME_ECLIPTIC)
    assert st == 0
    xyz = np.concatenate([pos, vel]).reshape(6, 1)
    for au in range(0, 5):
        obs = pos + [au, 0.0, 0.0]
        delay = au / C_AU_PER_DAY
        should, _ = o.batch_two_body(jd, xyz, [jd - delay])
        calc, _ = o.batch_two_body(jd, xyz, [jd], observer_pos=obs)
        assert np.allclose(calc[:, 0, 0], should[:, 0, 0])


def test_spice_seconds_jd():
    # spice/mod.rs:40-68
    from kete_amd.spk import jd_to_spice_jd, spice_jds_to_jd
    assert spice_jds_to_jd(0.0) == 2451545.0
    assert spice_jds_to_jd(86400.0) == 2451546.0
    assert jd_to_spice_jd(2451545.0) == 0.0
    assert jd_to_spice_jd(2451546.0) == 86400.0
    assert abs(jd_to_spice_jd(spice_jds_to_jd(1.0)) - 1.0) < 1e-5


def test_elements_to_state_round_trip():
    """elements.rs:190-264: the state lies on the conic with the given a, e,
    and its energy matches -GM/2a."""
    for e in (0.0, 0.1, 0.5, 0.9):
        a = 2.5
        st, p, v = o.elements_to_state(2460000.5, e, 0.2, 1.0, 2.0, a * (1 - e), 2460000.5 - 50.0)
        assert st == 0
        eps = 0.5 * v @ v - GMS / np.linalg.norm(p)
        assert abs(eps + GMS / (2 * a)) < 1e-14
        h = np.cross(p, v)
        ecc = np.cross(v, h) / GMS - p / np.linalg.norm(p)
        assert abs(np.linalg.norm(ecc) - e) < 1e-10
