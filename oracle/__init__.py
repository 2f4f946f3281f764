"""ctypes binding of the oracle: a CPU restatement of kete's orbit-propagation path.

TEST INFRASTRUCTURE ONLY.  This package is the parity checker and the CPU
baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it; the product (``kete_amd``)
never does.

The C source (``kete_oracle.c``) cites the kete Rust it restates line by line.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkete_oracle.so")
# the same restatement built with FMA contraction (-ffp-contract=fast): the
# rounding yardstick of the long-arc parity tests (use with use_library)
FMA_LIB_PATH = os.path.join(HERE, "libkete_oracle_fma.so")

OK, CONVERGENCE, VALUE_ERROR, DAF_LIMITS, UNKNOWN_FRAME, IO_ERROR, IMPACT = range(7)
FRAME_EQUATORIAL, FRAME_ECLIPTIC = 1, 17

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -O3 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class Counters(C.Structure):
    _fields_ = [("n_force", C.c_int64), ("n_sweep", C.c_int64),
                ("n_acc", C.c_int64), ("n_fail", C.c_int64)]


class KoState(C.Structure):
    _fields_ = [("jd", C.c_double), ("pos", C.c_double * 3), ("vel", C.c_double * 3),
                ("frame", C.c_int32), ("center", C.c_int64)]


class NonGrav(C.Structure):
    _fields_ = [("kind", C.c_int32)] + [(k, C.c_double) for k in
                                        ("a1", "a2", "a3", "alpha", "r_0", "m", "n", "k", "dt", "beta")]


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ptrs(cols):
    arr = (C.POINTER(C.c_double) * len(cols))()
    for i, c in enumerate(cols):
        arr[i] = _d(c)
    arr._keep = list(cols)  # the pointers do not hold their arrays
    return arr


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


class use_library:
    """Context manager: route every oracle call through another build of the
    oracle (a copy compiled with a planted mistake, for the tests that show
    the pins catch it), then back to the default library."""

    def __init__(self, path: str):
        self.path = path

    def __enter__(self):
        global _lib
        lib()
        self._saved = _lib
        _lib = _configure(C.CDLL(self.path))
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self._saved
        return False


def _configure(L):
    D, P, I64, I32 = C.c_double, C.POINTER(C.c_double), C.c_int64, C.c_int32
    PP = C.POINTER(P)
    L.ko_const.restype = D
    L.ko_set_radau_epsilon.restype = D
    L.ko_set_radau_epsilon.argtypes = [D]
    L.ko_rk45_test.argtypes = [C.c_int, D, D, D, D, P, C.POINTER(I64)]
    L.ko_compute_state_transition.argtypes = [P, P, D, D, D, C.c_int, P, P, P, C.POINTER(I64)]
    L.ko_const.argtypes = [C.c_char_p]
    L.ko_newton_square_root_test.argtypes = [D, D, P]
    L.ko_analytic_2_body.argtypes = [D, P, P, P, P]
    L.ko_compute_eccentric_anomaly.argtypes = [D, D, D, P]
    L.ko_elements_to_state.argtypes = [D, D, D, D, D, D, D, P, P]
    L.ko_spk_load_file.argtypes = [C.c_char_p]
    L.ko_spk_get_state.argtypes = [I64, D, I64, I32, P, P]
    L.ko_chebyshev_evaluate_both.argtypes = [D, P, P, P, C.c_int, P, P]
    L.ko_spk_accel.argtypes = [D, P, P, C.POINTER(NonGrav), C.c_int, P, C.POINTER(I64)]
    L.ko_spk_accel_ext.argtypes = [D, P, P, C.POINTER(NonGrav), C.c_int, C.c_int, P,
                                   C.POINTER(I64)]
    L.ko_spk21_record_eval.argtypes = [P, I64, D, P, P]
    L.ko_radau_central.argtypes = [D, D, P, P, P, P, C.POINTER(Counters)]
    L.ko_batch_two_body.argtypes = [I64, P, PP, I64, I32, I64, P, P, PP,
                                    C.POINTER(I32), C.c_int]
    L.ko_batch_n_body.argtypes = [I64, P, PP, I64, I32, P, C.c_int, C.POINTER(NonGrav),
                                  PP, C.POINTER(I32), C.POINTER(I64), P,
                                  C.POINTER(Counters), C.c_int]
    L.ko_batch_n_body_ext.argtypes = [I64, P, PP, I64, I32, P, C.c_int, C.c_int,
                                      C.POINTER(NonGrav), PP, C.POINTER(I32),
                                      C.POINTER(I64), P, C.POINTER(Counters), C.c_int]
    L.ko_rect_normals.argtypes = [P, D, D, D, P]
    L.ko_rect_contains.argtypes = [P, P, P]
    L.ko_fov_check_visible.restype = I64
    L.ko_fov_check_visible.argtypes = [P, D, P, P, I32, I64, P, PP, D,
                                       C.POINTER(I64), PP, C.c_int]
    L.ko_fov_state_check.restype = I64
    L.ko_fov_state_check.argtypes = [I64, D, PP, I32, I64, P, P, P, D, I64,
                                     C.POINTER(I64), C.POINTER(I64), P, C.c_int]
    L.ko_fov_state_check_ext.restype = I64
    L.ko_fov_state_check_ext.argtypes = [I64, D, PP, I32, I64, P, P, P, D, C.c_int, I64,
                                         C.POINTER(I64), C.POINTER(I64), P, C.c_int]
    L.ko_propagate_n_body_long.argtypes = [I64, D, PP, I32, D, P, C.POINTER(NonGrav), I64,
                                           PP, P, C.c_int]
    L.ko_fov_spk_check.restype = I64
    L.ko_fov_spk_check.argtypes = [I64, C.POINTER(I64), I32, I64, P, P, P, I64,
                                   C.POINTER(I64), C.POINTER(I64), P]
    L.ko_radau_one_body.argtypes = [I64, D, D, P, P, P, P]
    L.ko_change_frame.argtypes = [C.POINTER(KoState), I32]
    L.ko_change_center.argtypes = [C.POINTER(KoState), I64]
    L.ko_comet_elements_peri.restype = None
    L.ko_comet_elements_peri.argtypes = [D, P, P, P, P]
    L.ko_moid_batch.restype = None
    L.ko_moid_batch.argtypes = [I64, P, PP, I32, P, PP, P, P, C.POINTER(I32),
                                C.POINTER(I32), C.c_int]
    L.ko_ra_dec_with_rates.argtypes = [P, P, P, P, P]
    L.ko_hg_apparent_mag.restype = D
    L.ko_hg_apparent_mag.argtypes = [D, D, P, P]
    L.ko_hg_phase_curve_correction.restype = D
    L.ko_hg_phase_curve_correction.argtypes = [D, D]
    L.ko_chain_observe_batch.restype = C.c_int
    L.ko_chain_observe_batch.argtypes = [I64, P, PP, I32, C.POINTER(NonGrav), I32, P, P,
                                         C.c_int, C.c_int, I64, P, C.POINTER(I32),
                                         C.POINTER(I32), C.POINTER(I64), P, P, C.POINTER(I32),
                                         C.c_int]
    L.ko_fov_trace_set.restype = None
    L.ko_fov_trace_set.argtypes = [P, I64]
    L.ko_fov_trace_len.restype = I64
    L.ko_fov_trace_len.argtypes = []
    return L


def const(name: str) -> float:
    return lib().ko_const(name.encode())


def _v3(x):
    return np.ascontiguousarray(x, dtype=np.float64).reshape(3)


def newton_square_root(start=0.0, atol=1e-10):
    out = np.zeros(1)
    st = lib().ko_newton_square_root_test(start, atol, _d(out))
    return st, out[0]


def analytic_2_body(dt, pos, vel):
    p, v = _v3(pos), _v3(vel)
    po, vo = np.zeros(3), np.zeros(3)
    st = lib().ko_analytic_2_body(dt, _d(p), _d(v), _d(po), _d(vo))
    return st, po, vo


def eccentric_anomaly(ecc, mean_anom, peri_dist):
    out = np.zeros(1)
    st = lib().ko_compute_eccentric_anomaly(ecc, mean_anom, peri_dist, _d(out))
    return st, out[0]


def elements_to_state(epoch, ecc, incl, lon_asc, peri_arg, peri_dist, peri_time):
    p, v = np.zeros(3), np.zeros(3)
    st = lib().ko_elements_to_state(epoch, ecc, incl, lon_asc, peri_arg, peri_dist,
                                    peri_time, _d(p), _d(v))
    return st, p, v


def spk_load(*paths: str) -> int:
    """Replace the loaded SPK collection by the files given, in load order."""
    lib().ko_spk_reset()
    for path in paths:
        st = lib().ko_spk_load_file(path.encode())
        if st:
            return st
    return OK


def spk21_record_eval(record, n_coef, js):
    """One Type 21 record at js (TDB seconds past J2000) -> (status, pos AU, vel AU/day)."""
    rec = np.ascontiguousarray(record, dtype=np.float64)
    assert len(rec) == 4 * n_coef + 11
    p, v = np.zeros(3), np.zeros(3)
    st = lib().ko_spk21_record_eval(_d(rec), n_coef, js, _d(p), _d(v))
    return st, p, v


def rk45_test(kind, y0, t0, t1, tol):
    """kete's RK45 on y' = -y (kind 0) or y' = 2 (kind 1) -> (status, y1, evaluations)."""
    y1, ne = C.c_double(), C.c_int64()
    st = lib().ko_rk45_test(int(kind), float(y0), float(t0), float(t1), float(tol),
                            C.byref(y1), C.byref(ne))
    return st, y1.value, ne.value


STM_FIX_GRAD, STM_FIX_ACCEL = 1, 2


def compute_state_transition(pos, vel, jd0, jd1, central_mass=1.0, fix=0):
    """kete's compute_state_transition (state_transition.rs:52-112) ->
    (status, final pos, final vel, the 6x6 matrix kete returns, evaluations).
    kete's matrix is the transpose of d x(jd1) / d x(jd0).  fix: bits
    STM_FIX_GRAD / STM_FIX_ACCEL replace the two defects of its arithmetic
    (the cross-check's diagnosis; 0 is kete's own)."""
    p, v = _v3(pos), _v3(vel)
    po, vo, m = np.zeros(3), np.zeros(3), np.zeros(36)
    ne = C.c_int64()
    st = lib().ko_compute_state_transition(_d(p), _d(v), float(jd0), float(jd1),
                                           float(central_mass), int(fix), _d(po), _d(vo),
                                           _d(m), C.byref(ne))
    return st, po, vo, m.reshape(6, 6), ne.value


def radau_epsilon(eps=0.0):
    """Set the oracle's Radau tolerance (kete's 1e-6 unless changed);
    returns the previous value (eps <= 0 only reads it)."""
    return lib().ko_set_radau_epsilon(float(eps))


def spk_state(naif_id, jd, center=0, frame=FRAME_EQUATORIAL):
    p, v = np.zeros(3), np.zeros(3)
    st = lib().ko_spk_get_state(naif_id, jd, center, frame, _d(p), _d(v))
    return st, p, v


def chebyshev_both(x, cx, cy, cz):
    cx, cy, cz = (np.ascontiguousarray(c, dtype=np.float64) for c in (cx, cy, cz))
    val, der = np.zeros(3), np.zeros(3)
    st = lib().ko_chebyshev_evaluate_both(x, _d(cx), _d(cy), _d(cz), len(cx), _d(val), _d(der))
    return st, val, der


def spk_accel(t, pos, vel, exact=False, nongrav=None, include_asteroids=False):
    p, v = _v3(pos), _v3(vel)
    a = np.zeros(3)
    iid = C.c_int64(0)
    ng = C.byref(nongrav) if nongrav is not None else None
    st = lib().ko_spk_accel_ext(t, _d(p), _d(v), ng, int(exact), int(include_asteroids), _d(a),
                                C.byref(iid))
    return st, a, iid.value


def radau_central(t0, t1, pos, vel):
    p, v = _v3(pos), _v3(vel)
    po, vo = np.zeros(3), np.zeros(3)
    cnt = Counters()
    st = lib().ko_radau_central(t0, t1, _d(p), _d(v), _d(po), _d(vo), C.byref(cnt))
    return st, po, vo, cnt


def _soa(xyz):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    assert xyz.ndim == 2 and xyz.shape[0] == 6
    return xyz


def batch_two_body(jd0, xyz, targets, center=10, frame=FRAME_ECLIPTIC,
                   observer_pos=None, n_threads=None):
    """propagate_two_body over n objects x m target epochs -> (out[6,m,n], status[m,n])."""
    xyz = _soa(xyz)
    n = xyz.shape[1]
    jd0 = np.ascontiguousarray(np.broadcast_to(jd0, (n,)), dtype=np.float64)
    targets = np.ascontiguousarray(np.atleast_1d(targets), dtype=np.float64)
    m = len(targets)
    out = np.empty((6, m, n))
    status = np.empty((m, n), dtype=np.int32)
    obs = None if observer_pos is None else _d(_v3(observer_pos).copy())
    obs_keep = None if observer_pos is None else _v3(observer_pos).copy()
    if obs_keep is not None:
        obs = _d(obs_keep)
    lib().ko_batch_two_body(n, _d(jd0), _ptrs(list(xyz)), center, frame, m, _d(targets), obs,
                            _ptrs([out[k] for k in range(6)]),
                            status.ctypes.data_as(C.POINTER(C.c_int32)),
                            n_threads or os.cpu_count())
    return out, status


def batch_n_body(jd0, xyz, jd_final, center=10, frame=FRAME_ECLIPTIC, nongravs=None,
                 n_threads=None, include_asteroids=False):
    """propagate_n_body over n objects -> dict(out[6,n], status, impact_id, impact_jd, counters)."""
    xyz = _soa(xyz)
    n = xyz.shape[1]
    jd0 = np.ascontiguousarray(np.broadcast_to(jd0, (n,)), dtype=np.float64)
    jf = np.ascontiguousarray(np.atleast_1d(jd_final), dtype=np.float64)
    stride = 0 if len(jf) == 1 else 1
    if stride:
        assert len(jf) == n
    out = np.empty((6, n))
    status = np.empty(n, dtype=np.int32)
    iid = np.empty(n, dtype=np.int64)
    ijd = np.empty(n)
    cnt = (Counters * max(n, 1))()
    ng = None
    if nongravs is not None:
        ng = (NonGrav * n)(*nongravs)
    lib().ko_batch_n_body_ext(n, _d(jd0), _ptrs(list(xyz)), center, frame, _d(jf), stride,
                              int(include_asteroids), ng,
                              _ptrs([out[k] for k in range(6)]),
                              status.ctypes.data_as(C.POINTER(C.c_int32)),
                              iid.ctypes.data_as(C.POINTER(C.c_int64)), _d(ijd), cnt,
                              n_threads or os.cpu_count())
    counters = np.array([(c.n_force, c.n_sweep, c.n_acc, c.n_fail) for c in cnt[:n]],
                        dtype=np.int64).reshape(n, 4)
    return dict(out=out, status=status, impact_id=iid, impact_jd=ijd, counters=counters)


def rect_normals(pointing, rotation, lon_width, lat_width):
    p = _v3(pointing)
    out = np.zeros(12)
    lib().ko_rect_normals(_d(p), rotation, lon_width, lat_width, _d(out))
    return out.reshape(4, 3)


def rect_contains(normals, v):
    nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(12)
    vv = _v3(v)
    dist = np.zeros(1)
    inside = lib().ko_rect_contains(_d(nrm), _d(vv), _d(dist))
    return bool(inside), dist[0]


def fov_check_visible(normals, obs_jd, obs_pos, obs_vel, jd, xyz, dt_limit=3.0,
                      frame=FRAME_ECLIPTIC, n_threads=None):
    xyz = _soa(xyz)
    n = xyz.shape[1]
    jd = np.ascontiguousarray(np.broadcast_to(jd, (n,)), dtype=np.float64)
    nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(12)
    idx = np.empty(max(n, 1), dtype=np.int64)
    vs = np.empty((7, max(n, 1)))
    nv = lib().ko_fov_check_visible(_d(nrm), obs_jd, _d(_v3(obs_pos)), _d(_v3(obs_vel)),
                                    frame, n, _d(jd), _ptrs(list(xyz)), dt_limit,
                                    idx.ctypes.data_as(C.POINTER(C.c_int64)),
                                    _ptrs([vs[k] for k in range(7)]),
                                    n_threads or os.cpu_count())
    return idx[:nv].copy(), vs[:, :nv].copy()


def fov_state_check(jd0, xyz, field_jd, obs_xyz, normals, dt_limit=3.0,
                    frame=FRAME_ECLIPTIC, cap=None, n_threads=None, include_asteroids=False,
                    trace=False):
    """fov_checks_py -> (field index, object index, state[n_hits, 7]); with
    trace=True also the re-propagation schedule as a list of (kind, epoch):
    ('group', mean epoch), ('big', target), ('reset', epoch), ('small', target)."""
    if trace:
        buf = np.zeros(2 * 4096)
        lib().ko_fov_trace_set(_d(buf), 4096)
        try:
            res = fov_state_check(jd0, xyz, field_jd, obs_xyz, normals, dt_limit, frame, cap,
                                  n_threads, include_asteroids)
            nt = lib().ko_fov_trace_len()
        finally:
            lib().ko_fov_trace_set(None, 0)
        kinds = ("group", "big", "reset", "small")
        return res + ([(kinds[int(buf[2 * k])], float(buf[2 * k + 1])) for k in range(nt)],)
    xyz = _soa(xyz)
    n = xyz.shape[1]
    field_jd = np.ascontiguousarray(field_jd, dtype=np.float64)
    nf = len(field_jd)
    obs_xyz = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(nf, 6)
    nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(nf, 12)
    cap = cap or max(1024, n * 4)
    hf = np.empty(cap, dtype=np.int64)
    ho = np.empty(cap, dtype=np.int64)
    hs = np.empty((cap, 7))
    nh = lib().ko_fov_state_check_ext(n, jd0, _ptrs(list(xyz)), frame, nf, _d(field_jd),
                                      _d(obs_xyz), _d(nrm), dt_limit, int(include_asteroids),
                                      cap, hf.ctypes.data_as(C.POINTER(C.c_int64)),
                                      ho.ctypes.data_as(C.POINTER(C.c_int64)), _d(hs),
                                      n_threads or os.cpu_count())
    if nh < 0:
        raise RuntimeError("hit capacity exceeded")
    return hf[:nh].copy(), ho[:nh].copy(), hs[:nh].copy()


def propagate_n_body_long(jd0, xyz, jd_final, frame=FRAME_ECLIPTIC, planet_xyz=None,
                          nongravs=None, batch_size=10, n_threads=None):
    """-> (status, out[6, n] Sun-centred equatorial, planets[9, 6] of the last batch)"""
    xyz = _soa(xyz)
    n = xyz.shape[1]
    out = np.empty((6, n))
    planets = np.empty((9, 6))
    pl = None
    if planet_xyz is not None:
        pl_arr = np.ascontiguousarray(planet_xyz, dtype=np.float64).reshape(9, 6)
        pl = _d(pl_arr)
    ng = (NonGrav * n)(*nongravs) if nongravs is not None else None
    st = lib().ko_propagate_n_body_long(n, jd0, _ptrs(list(xyz)), frame, jd_final, pl, ng,
                                        batch_size, _ptrs([out[k] for k in range(6)]),
                                        _d(planets), n_threads or os.cpu_count())
    return st, out, planets


def fov_spk_check(ids, field_jd, obs_xyz, normals, frame=FRAME_ECLIPTIC, cap=None):
    """fov_spk_check over fields in epoch order -> (field, id index, state[n, 7])."""
    ids = np.ascontiguousarray(ids, dtype=np.int64)
    field_jd = np.ascontiguousarray(field_jd, dtype=np.float64)
    nf = len(field_jd)
    obs_xyz = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(nf, 6)
    nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(nf, 12)
    cap = cap or max(1024, len(ids) * nf)
    hf = np.empty(cap, dtype=np.int64)
    ho = np.empty(cap, dtype=np.int64)
    hs = np.empty((cap, 7))
    nh = lib().ko_fov_spk_check(len(ids), ids.ctypes.data_as(C.POINTER(C.c_int64)), frame, nf,
                                _d(field_jd), _d(obs_xyz), _d(nrm), cap,
                                hf.ctypes.data_as(C.POINTER(C.c_int64)),
                                ho.ctypes.data_as(C.POINTER(C.c_int64)), _d(hs))
    if nh < 0:
        raise RuntimeError("hit capacity exceeded")
    return hf[:nh].copy(), ho[:nh].copy(), hs[:nh].copy()


def radau_one_body(naif_id, t0, t1, pos, vel):
    """Test hook: kete's Radau with one PLANETS body at rest at the origin
    (its GravParams::add_acceleration as the whole force), equatorial."""
    p, v = _v3(pos), _v3(vel)
    po, vo = np.zeros(3), np.zeros(3)
    st = lib().ko_radau_one_body(naif_id, t0, t1, _d(p), _d(v), _d(po), _d(vo))
    return st, po, vo


def change_frame(jd, pos, vel, frame, center, new_frame):
    """State::try_change_frame_mut -> (status, pos, vel)."""
    s = KoState(jd, (C.c_double * 3)(*pos), (C.c_double * 3)(*vel), frame, center)
    st = lib().ko_change_frame(C.byref(s), new_frame)
    return st, np.array(s.pos[:]), np.array(s.vel[:])


def change_center(jd, pos, vel, frame, center, new_center):
    """State::try_change_center through the loaded SPK -> (status, pos, vel)."""
    s = KoState(jd, (C.c_double * 3)(*pos), (C.c_double * 3)(*vel), frame, center)
    st = lib().ko_change_center(C.byref(s), new_center)
    return st, np.array(s.pos[:]), np.array(s.vel[:])


def comet_elements_peri(epoch, pos, vel):
    """-> (peri_time, orbital_period) of CometElements::from_state."""
    pt, per = C.c_double(), C.c_double()
    p, v = _v3(pos), _v3(vel)
    lib().ko_comet_elements_peri(epoch, _d(p), _d(v), C.byref(pt), C.byref(per))
    return pt.value, per.value


def moid_batch(jd_a, xyz_a, frame=FRAME_ECLIPTIC, jd_b=None, xyz_b=None, n_threads=None):
    """kete moid per object -> (moid[n], status[n], grid_min[n], iters[n]); with
    xyz_b None the second orbit is the Earth's at jd_a (moid_py's default)."""
    xyz_a = _soa(xyz_a)
    n = xyz_a.shape[1]
    jd_a = np.ascontiguousarray(np.broadcast_to(np.asarray(jd_a, dtype=np.float64), (n,)))
    if jd_b is not None:
        jd_b = np.ascontiguousarray(np.broadcast_to(np.asarray(jd_b, dtype=np.float64), (n,)))
    xb = _ptrs(list(_soa(xyz_b))) if xyz_b is not None else None
    out, gmin = np.empty(n), np.empty(n)
    st, it = np.empty(n, dtype=np.int32), np.empty(n, dtype=np.int32)
    lib().ko_moid_batch(n, _d(jd_a), _ptrs(list(xyz_a)), frame,
                        _d(jd_b) if jd_b is not None else None, xb, _d(out), _d(gmin),
                        st.ctypes.data_as(C.POINTER(C.c_int32)),
                        it.ctypes.data_as(C.POINTER(C.c_int32)), n_threads or os.cpu_count())
    return out, st, gmin, it


def ra_dec_with_rates(pos, vel, obs_pos, obs_vel):
    out = np.zeros(4)
    lib().ko_ra_dec_with_rates(_d(_v3(pos)), _d(_v3(vel)), _d(_v3(obs_pos)),
                               _d(_v3(obs_vel)), _d(out))
    return out


def hg_apparent_mag(h, g, sun2obj, sun2obs):
    return lib().ko_hg_apparent_mag(h, g, _d(_v3(sun2obj)), _d(_v3(sun2obs)))


def hg_phase_curve_correction(g, phase):
    return lib().ko_hg_phase_curve_correction(g, phase)


def chain_observe(jd0, xyz, jd_obs, obs, frame=FRAME_ECLIPTIC, nongravs=None, restart=True,
                  schedule=None, record=False, cap=20000, n_threads=None):
    """The orbit-fit observable chain for n objects (xyz [6, n], Sun-centred
    in `frame` at jd0) at m epochs (jd_obs [m] or [m, n]; observer obs [3, m]
    or [3, m, n], Sun-centred in `frame`): N-body arc to each epoch (restart:
    a new integrator per epoch, else one integration through them), one
    light-time two-body step, RA/Dec.  record=True also returns each
    object's step schedule; schedule=<that> replays it (frozen steps).
    -> dict(radec [2, m, n], state [6, m, n], status [n], schedule)."""
    xyz = _soa(xyz)
    n = xyz.shape[1]
    jd_obs = np.asarray(jd_obs, dtype=np.float64)
    m = jd_obs.shape[0]
    jo = np.ascontiguousarray(np.broadcast_to(jd_obs.reshape(m, -1), (m, n)))
    ob = np.asarray(obs, dtype=np.float64).reshape(3, m, -1)
    ob = np.ascontiguousarray(np.broadcast_to(ob, (3, m, n)).transpose(1, 0, 2))  # [m, 3, n]
    jd0 = np.ascontiguousarray(np.broadcast_to(jd0, (n,)), dtype=np.float64)
    radec = np.empty((m, 2, n))
    state = np.empty((m, 6, n))
    status = np.empty(n, dtype=np.int32)
    ng = (NonGrav * n)(*nongravs) if nongravs is not None else None
    mode = 2 if schedule is not None else (1 if record else 0)
    if schedule is not None:
        sh, ssw, sacc, sn = schedule
        cap = sh.shape[1]
    elif record:
        sh, ssw, sacc = np.empty((n, cap)), np.empty((n, cap), np.int32), np.empty((n, cap),
                                                                                 np.int32)
        sn = np.zeros(n, dtype=np.int64)
    else:
        sh = ssw = sacc = sn = None
    i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None
    lib().ko_chain_observe_batch(n, _d(jd0), _ptrs(list(xyz)), frame, ng, m, _d(jo), _d(ob),
                                 int(bool(restart)), mode, cap,
                                 _d(sh) if sh is not None else None, i32(ssw), i32(sacc),
                                 sn.ctypes.data_as(C.POINTER(C.c_int64)) if sn is not None
                                 else None, _d(radec), _d(state), i32(status),
                                 n_threads or os.cpu_count())
    return dict(radec=radec.transpose(1, 0, 2), state=state.transpose(1, 0, 2), status=status,
                schedule=(sh, ssw, sacc, sn) if mode else None)

