"""Pins of the oracle's MOID (kete moid, src/kete_core/src/propagation/kepler.rs:325-404;
moid_py src/kete/rust/propagation.rs:27-39).

- kete's known answers (src/tests/test_propagation.py:108-121): the Earth's
  MOID with itself is 0; Ceres against the Earth 1.58 and against Mercury
  2.18 (to 1e-2; the planets here come from the synthetic DE-style ephemeris,
  kete's test reads DE440).
- The perihelion time of CometElements::from_state (elements.rs:62-163) on
  kete's round-trip grid (elements.rs:411-472): elements -> state -> the
  restated perihelion time -> state reproduces the state to 1e-7, and the
  two-body state at that time is at perihelion (|r| = q, r.v = 0).
- Closed forms: coplanar circles (|r1 - r2|), a circle inside a coplanar
  ellipse (q - r), inclined circles of radii r1 < r2 (r2 - r1, reached on the
  line of nodes), equal inclined circles (0).
- Brute force: for random pairs the MOID is never below the minimum distance
  over a dense grid of eccentric anomalies refined by a local search, and
  matches it to 1e-7 AU for nearly all pairs (kete's search is local, from
  the best point of a 100 x 100 time grid).

The iteration sequence of the Nelder-Mead restatement (argmin 0.10, not in
the kete tree) is not pinned beyond the value it converges to.
"""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle
from kete_amd import synthetic as syn

GMS_SQRT = 0.01720209894996


def circ(r, incl=0.0, node=0.0, phase=0.0):
    """Sun-centred ecliptic state on a circular orbit of radius r."""
    v = GMS_SQRT / np.sqrt(r)
    p = np.array([r * np.cos(phase), r * np.sin(phase), 0.0])
    u = np.array([-v * np.sin(phase), v * np.cos(phase), 0.0])
    ci, si, cn, sn = np.cos(incl), np.sin(incl), np.cos(node), np.sin(node)
    rx = np.array([[1, 0, 0], [0, ci, -si], [0, si, ci]])
    rz = np.array([[cn, -sn, 0], [sn, cn, 0], [0, 0, 1]])
    m = rz @ rx
    return np.r_[m @ p, m @ u]


def moid_pair(a, b, jd_a=0.0, jd_b=0.0):
    out, st, g, it = oracle.moid_batch(jd_a, a[:, None], jd_b=jd_b, xyz_b=b[:, None])
    assert st[0] == 0
    return out[0], g[0], it[0]


def test_kete_known_answers(oracle_spk):
    jd = 2461161.5
    ceres = np.array([2.908844981531, -0.007973772593, -0.535807560572,
                      -0.00028988685, 0.009639190713, 0.000360810151])

    def planet(i):
        st, p, v = oracle.spk_state(i, jd, 10, oracle.FRAME_ECLIPTIC)
        assert st == 0
        return np.r_[p, v]

    e, m = planet(399), planet(1)
    assert np.isclose(moid_pair(e, e, jd, jd)[0], 0.0)
    assert np.isclose(moid_pair(ceres, e, 2462583.0, jd)[0], 1.58, atol=1e-2)
    assert np.isclose(moid_pair(ceres, m, 2462583.0, jd)[0], 2.18, atol=1e-2)
    # moid_py's default second state: the Earth at the first state's epoch
    out, st, _, _ = oracle.moid_batch(2462583.0, ceres[:, None])
    assert st[0] == 0 and np.isclose(out[0], 1.58, atol=1e-2)


@pytest.mark.parametrize("ecc", [0.001, 0.1, 0.5, 1.0, 2.0])
def test_perihelion_time_roundtrip(ecc):
    n = 0
    for epoch in (-10.0, 0.0, 10.0):
        for incl in (-2.0, 0.1, 0.0, 2.0):
            for lon in (-0.5, 0.0, 4.0):
                for peri_time in (-100.0, 0.0, 100.0):
                    for arg in (-1.0, 0.0, 1.0):
                        for q in (0.3, 0.5):
                            st, p, v = oracle.elements_to_state(epoch, ecc, incl, lon, arg, q,
                                                                peri_time)
                            assert st == 0
                            pt, period = oracle.comet_elements_peri(epoch, p, v)
                            st, p2, v2 = oracle.elements_to_state(epoch, ecc, incl, lon, arg,
                                                                  q, pt)
                            assert st == 0
                            assert np.abs(p2 - p).max() < 1e-7 and np.abs(v2 - v).max() < 1e-7
                            st, pp, vp = oracle.analytic_2_body(pt - epoch, p, v)
                            assert st == 0
                            assert abs(np.linalg.norm(pp) - q) < 1e-9
                            assert abs(np.dot(pp, vp)) < 1e-9
                            a = q / (1.0 - ecc) if abs(ecc - 1.0) > 1e-3 else np.nan
                            if ecc < 1.0:
                                assert np.isclose(period, 2 * np.pi * a ** 1.5 / GMS_SQRT,
                                                  rtol=1e-12)
                            else:  # NaN semi-major axis (parabolic) or negative
                                assert not np.isfinite(period)
                            n += 1
    assert n == 648


def test_closed_forms():
    m, g, _ = moid_pair(circ(1.0, phase=0.3), circ(2.5, phase=2.0))
    assert abs(m - 1.5) < 1e-9
    # coplanar ellipse with q = 1.2, Q = 3 around a unit circle: q - 1
    st, p, v = oracle.elements_to_state(0.0, (3.0 - 1.2) / 4.2, 0.0, 0.0, 0.7, 1.2, 40.0)
    assert st == 0
    m, _, _ = moid_pair(circ(1.0, phase=1.0), np.r_[p, v])
    assert abs(m - 0.2) < 1e-9
    # inclined circles: the line of nodes gives r2 - r1
    for incl in (0.2, 0.5, 1.2):
        m, _, _ = moid_pair(circ(1.0, phase=0.4), circ(2.0, incl=incl, node=0.9, phase=2.5))
        assert abs(m - 1.0) < 1e-9
    # equal inclined circles cross on the line of nodes
    m, _, _ = moid_pair(circ(1.3, phase=0.1), circ(1.3, incl=0.4, node=2.0, phase=1.0))
    assert m < 1e-6


def orbit_curve(x):
    """(P, Q, a, e) of the conic through state x (Sun-centred, AU, AU/day)."""
    r, v = x[:3], x[3:] / GMS_SQRT
    h = np.cross(r, v)
    ev = np.cross(v, h) - r / np.linalg.norm(r)
    e = np.linalg.norm(ev)
    a = 1.0 / (2.0 / np.linalg.norm(r) - v @ v)
    P = ev / e
    Q = np.cross(h / np.linalg.norm(h), P)
    return P, Q, a, e


def ell_pos(c, E):
    P, Q, a, e = c
    E = np.asarray(E)
    return (a * (np.cos(E) - e))[..., None] * P + (a * np.sqrt(1 - e * e) * np.sin(E))[..., None] * Q


def brute_moid(xa, xb):
    ca, cb = orbit_curve(xa), orbit_curve(xb)
    E = np.linspace(0, 2 * np.pi, 721)[:-1]
    pa, pb = ell_pos(ca, E), ell_pos(cb, E)
    d = np.linalg.norm(pa[:, None, :] - pb[None, :, :], axis=-1)
    f = lambda z: np.linalg.norm(ell_pos(ca, z[0]) - ell_pos(cb, z[1]))
    best = np.inf
    for k in np.argsort(d, axis=None)[:6]:
        i, j = np.unravel_index(k, d.shape)
        r = minimize(f, [E[i], E[j]], method="Nelder-Mead",
                     options=dict(xatol=1e-12, fatol=1e-14, maxiter=4000))
        best = min(best, r.fun)
    return best


def test_brute_force_pairs(oracle_spk):
    xyz = syn.n_body_workload(n=40, neo_fraction=0.3, seed_index=21)
    st, p, v = oracle.spk_state(399, syn.EPOCH_JD, 10, oracle.FRAME_ECLIPTIC)
    earth = np.r_[p, v]
    out, stat, g, it = oracle.moid_batch(syn.EPOCH_JD, xyz)
    assert np.all(stat == 0)
    close = 0
    for i in range(xyz.shape[1]):
        bm = brute_moid(xyz[:, i], earth)
        assert out[i] >= bm - 1e-9
        close += abs(out[i] - bm) < 1e-7
    assert close >= 36, close
