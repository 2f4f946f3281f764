"""kete_amd -- kete's batched orbit-propagation path on one AMD MI355X.

The Python surface mirrors kete's (``kete.propagate_n_body``,
``kete.propagate_two_body``, ``kete.fov_state_check``, ``kete.State``,
``kete.RectangleFOV``; ``src/kete/propagation.py:8-14``, ``src/kete/fov.py``)
so that calling code can switch by changing the import.  Work is done by the
HIP kernels of ``libkete_hip.so`` through its C ABI (``include/kete_hip.h``);
batches can also be passed as structure-of-arrays (:class:`StateArray`) to skip
per-object Python objects.
"""
from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .spk import FRAME_ECLIPTIC, FRAME_EQUATORIAL, Segment, Type2Table, read_spk

__all__ = [
    "State", "StateArray", "NonGravModel", "Context", "propagate_n_body", "propagate_two_body",
    "spk_state", "load_spk", "get_context", "Frames", "KeteHipError", "propagate_n_body_long",
    "moid",
]

KeteHipError = _lib.KeteHipError

STATUS_NAMES = {0: "ok", 1: "Convergence", 2: "ValueError", 3: "DAFLimits", 4: "UnknownFrame",
                5: "IOError", 6: "Impact"}


class Frames:
    """kete's frame names with their SPK numbers."""
    Equatorial = FRAME_EQUATORIAL
    Ecliptic = FRAME_ECLIPTIC


@dataclass
class State:
    """One object's state (``state.rs:76-95``): position (AU), velocity
    (AU/day), TDB JD, frame, and center NAIF id."""
    desig: object
    jd: float
    pos: np.ndarray
    vel: np.ndarray
    frame: int = FRAME_ECLIPTIC
    center_id: int = 10

    def __post_init__(self):
        self.pos = np.asarray(self.pos, dtype=np.float64).reshape(3)
        self.vel = np.asarray(self.vel, dtype=np.float64).reshape(3)

    @property
    def is_finite(self) -> bool:
        return bool(np.all(np.isfinite(self.pos)) and np.all(np.isfinite(self.vel))
                    and np.isfinite(self.jd))


@dataclass
class StateArray:
    """A batch of states as structure-of-arrays: ``xyz[0:3]`` positions,
    ``xyz[3:6]`` velocities, one frame and center for the batch."""
    jd: np.ndarray
    xyz: np.ndarray
    frame: int = FRAME_ECLIPTIC
    center_id: int = 10
    desigs: list | None = field(default=None, repr=False)

    def __post_init__(self):
        self.xyz = np.ascontiguousarray(self.xyz, dtype=np.float64)
        if self.xyz.ndim != 2 or self.xyz.shape[0] != 6:
            raise ValueError("xyz must have shape (6, n)")
        self.jd = np.ascontiguousarray(np.broadcast_to(np.asarray(self.jd, dtype=np.float64),
                                                       (self.xyz.shape[1],)))

    def __len__(self):
        return self.xyz.shape[1]

    @classmethod
    def from_states(cls, states):
        states = list(states)
        if not states:
            raise ValueError("empty state list")
        frame, center = states[0].frame, states[0].center_id
        if any(s.frame != frame or s.center_id != center for s in states):
            raise ValueError("all states in a batch must share frame and center")
        xyz = np.array([np.concatenate([s.pos, s.vel]) for s in states]).T
        return cls(np.array([s.jd for s in states]), xyz, frame, center,
                   [s.desig for s in states])

    def to_states(self):
        d = self.desigs or [None] * len(self)
        return [State(d[i], float(self.jd[i]), self.xyz[:3, i].copy(), self.xyz[3:, i].copy(),
                      self.frame, self.center_id) for i in range(len(self))]


class NonGravModel:
    """Non-gravitational models (``propagation/nongrav.rs:12-61``; Python
    constructors ``src/kete/rust/nongrav.rs:72-161``)."""

    def __init__(self, kind, a1=0.0, a2=0.0, a3=0.0, alpha=1.0, r_0=1.0, m=2.0, n=1.0, k=0.0,
                 dt=0.0, beta=0.0):
        self.params = dict(kind=kind, a1=a1, a2=a2, a3=a3, alpha=alpha, r_0=r_0, m=m, n=n, k=k,
                           dt=dt, beta=beta)

    @classmethod
    def new_dust(cls, beta):
        return cls(2, beta=beta)

    @classmethod
    def new_comet(cls, a1=0.0, a2=0.0, a3=0.0, alpha=0.1112620426, r_0=2.808, m=2.15, n=5.093,
                  k=4.6142, dt=0.0):
        return cls(1, a1, a2, a3, alpha, r_0, m, n, k, dt)

    @classmethod
    def new_asteroid(cls, a1=0.0, a2=0.0, a3=0.0, alpha=1.0, r_0=1.0, m=2.0, n=1.0, k=0.0,
                     dt=0.0):
        return cls(1, a1, a2, a3, alpha, r_0, m, n, k, dt)


def nongrav_array(models, n):
    arr = np.zeros(n, dtype=_lib.NONGRAV_DTYPE)
    for i, m in enumerate(models):
        if m is None:
            continue
        for k, v in m.params.items():
            arr[k][i] = v
    return arr


class Context:
    """A device, its stream and the uploaded ephemeris (kete's process-global
    ``LOADED_SPK``, ``spice/spk.rs:368-377``, becomes per-context state)."""

    def __init__(self, device: int = 0):
        self.device = device
        self.L = _lib.load()
        h = C.c_void_p()
        _lib.check(self.L.kh_ctx_create(device, C.byref(h)))
        self.h = h
        self.table: Type2Table | None = None

    def close(self):
        if self.h:
            self.L.kh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- ephemeris ---------------------------------------------------------
    def load_spk(self, *paths_or_segments):
        """Replace the ephemeris by the segments of the files (or Segment
        lists) given, in load order: the Type 2 planet segments and the
        Type 21 small-body segments (kete ``spk.load``, ``spice/spk.rs:307-329``)."""
        segs = []
        for item in paths_or_segments:
            segs.extend(read_spk(item) if isinstance(item, str) else list(item))
        table = Type2Table(segs)
        arr = (_lib.KhSpk2Seg * len(table.segments))()
        keep = []
        for k, s in enumerate(table.segments):
            rec = np.ascontiguousarray(s.records())
            keep.append(rec)
            arr[k] = _lib.KhSpk2Seg(s.obj_id, s.center_id, s.frame, s.rsize, s.n_rec, s.jd_start,
                                    s.jd_end, s.intlen, _lib.dptr(rec.reshape(-1)))
        _lib.check(self.L.kh_spk_upload_type2(self.h, arr, len(table.segments)))
        t21 = [s for s in segs if s.seg_type == 21]
        arr21 = (_lib.KhSpk21Seg * max(len(t21), 1))()
        for k, s in enumerate(t21):
            rec = np.ascontiguousarray(s.records())
            ep = np.ascontiguousarray(s.epochs())
            keep += [rec, ep]
            arr21[k] = _lib.KhSpk21Seg(s.obj_id, s.center_id, s.frame, s.n_coef, s.n_rec,
                                       s.jd_start, s.jd_end, _lib.dptr(rec.reshape(-1)),
                                       _lib.dptr(ep))
        _lib.check(self.L.kh_spk_upload_type21(self.h, arr21, len(t21)))
        self.table = table
        self.type21 = t21
        return table

    def spk_state(self, obj_id, jd, center=10, frame=FRAME_ECLIPTIC):
        jd = np.ascontiguousarray(np.atleast_1d(jd), dtype=np.float64)
        out = np.empty((6, len(jd)))
        st = np.empty(len(jd), dtype=np.int32)
        _lib.check(self.L.kh_spk_states(self.h, obj_id, center, frame, len(jd), _lib.dptr(jd),
                                        _lib.ptr_array(list(out)), _lib.iptr(st, C.c_int32)))
        return out, st

    # -- propagation ---------------------------------------------------------
    def two_body(self, batch: StateArray, targets, observer_pos=None):
        """-> (out[6, m, n], status[m, n])"""
        targets = np.ascontiguousarray(np.atleast_1d(targets), dtype=np.float64)
        n, m = len(batch), len(targets)
        out = np.empty((6, m, n))
        st = np.empty((m, n), dtype=np.int32)
        obs = None
        if observer_pos is not None:
            obs_arr = np.ascontiguousarray(observer_pos, dtype=np.float64).reshape(3)
            obs = _lib.dptr(obs_arr)
        _lib.check(self.L.kh_propagate_two_body(
            self.h, n, _lib.dptr(batch.jd), _lib.ptr_array(list(batch.xyz)), batch.center_id,
            batch.frame, m, _lib.dptr(targets), obs,
            _lib.ptr_array([out[k] for k in range(6)]), _lib.iptr(st, C.c_int32)))
        return out, st

    def n_body(self, batch: StateArray, jd_final, non_gravs=None, epsilon=1e-6, mode=-1,
               counters=False, include_asteroids=False):
        """-> dict(out[6, n], status[n], impact_id[n], impact_jd[n], counters[n, 4])"""
        if self.table is None:
            raise KeteHipError("no ephemeris loaded")
        n = len(batch)
        jf = np.ascontiguousarray(np.atleast_1d(jd_final), dtype=np.float64)
        stride = 0 if len(jf) == 1 else 1
        if stride and len(jf) != n:
            raise ValueError("jd_final must have 1 or n entries")
        out = np.empty((6, n))
        st = np.empty(n, dtype=np.int32)
        iid = np.empty(n, dtype=np.int64)
        ijd = np.empty(n)
        cnt = np.zeros((n, 4), dtype=np.int32) if counters else None
        ng = None
        if non_gravs is not None:
            ng_arr = non_gravs if isinstance(non_gravs, np.ndarray) else nongrav_array(non_gravs, n)
            ng = ng_arr.ctypes.data_as(C.POINTER(_lib.KhNonGrav))
        _lib.check(self.L.kh_propagate_n_body(
            self.h, n, _lib.dptr(batch.jd), _lib.ptr_array(list(batch.xyz)), batch.center_id,
            batch.frame, _lib.dptr(jf), stride, int(bool(include_asteroids)), ng, epsilon, mode,
            _lib.ptr_array([out[k] for k in range(6)]), _lib.iptr(st, C.c_int32),
            _lib.iptr(iid, C.c_int64), _lib.dptr(ijd),
            cnt.ctypes.data_as(C.POINTER(_lib.KhCounters)) if cnt is not None else None))
        return dict(out=out, status=st, impact_id=iid, impact_jd=ijd, counters=cnt)

    def autotune(self, batch: StateArray, jd_final, include_asteroids=False, sample=0,
                 apply=True):
        """Pick the N-body launch shape for batches like this one
        (kh_autotune_n_body): the batch goes to the device once, every
        (n/sample)-th object is timed over the candidate shapes, and with
        ``apply`` later ``mode=-1`` calls on this context use the fastest.
        -> dict(mode, block, window, latency_k, ms, default_ms, n_tried, n_sample)"""
        import torch
        dev = torch.device("cuda", self.device)
        n = len(batch)
        jf = np.ascontiguousarray(np.atleast_1d(jd_final), dtype=np.float64)
        stride = 0 if len(jf) == 1 else 1
        t_jd0 = torch.from_numpy(np.array(batch.jd, dtype=np.float64)).to(dev)
        t_xyz = torch.from_numpy(np.array(batch.xyz, dtype=np.float64)).to(dev)
        t_jf = torch.from_numpy(jf).to(dev)
        torch.cuda.synchronize(dev)
        res = _lib.KhTuneResult()
        _lib.check(self.L.kh_autotune_n_body(
            self.h, n, _lib.raw(t_jd0.data_ptr()), _lib.ptr_array([t_xyz[k].data_ptr()
                                                                  for k in range(6)]),
            batch.center_id, batch.frame, _lib.raw(t_jf.data_ptr()), stride,
            int(bool(include_asteroids)), int(sample), int(bool(apply)), C.byref(res)))
        return {k: getattr(res, k) for k, _ in _lib.KhTuneResult._fields_}

    def autotune_clear(self):
        _lib.check(self.L.kh_autotune_clear(self.h))

    def radec_partials(self, jd0, xyz, jd_obs, obs, frame=FRAME_ECLIPTIC, non_gravs=None,
                       force=0, restart=False, epsilon=1e-6, states=False, counters=False):
        """RA/Dec at each observation epoch and their partials with respect to
        (x, y, z, vx, vy, vz, A1, A2, A3).

        jd0: [n] or scalar; xyz: [6, n] Sun-centred in `frame`; jd_obs: [m]
        or [m, n]; obs: observer position w.r.t. the Sun, [3, m] or [3, m, n].
        -> dict(radec[2, m, n], jac[2, 9, m, n], status[n], and with
        states=True state[6, m, n], dstate[6, 9, m, n]; counters[n, 4])"""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = xyz.shape[1]
        jd0 = np.ascontiguousarray(np.broadcast_to(np.asarray(jd0, dtype=np.float64), (n,)))
        jd_obs = np.asarray(jd_obs, dtype=np.float64)
        m = jd_obs.shape[0]
        jo = np.ascontiguousarray(np.broadcast_to(jd_obs.reshape(m, -1), (m, n)))
        ob = np.asarray(obs, dtype=np.float64).reshape(3, m, -1)
        ob = np.ascontiguousarray(np.broadcast_to(ob, (3, m, n)))
        radec = np.empty((2, m, n))
        jac = np.empty((2, 9, m, n))
        st = np.empty(n, dtype=np.int32)
        state = np.empty((6, m, n)) if states else None
        dstate = np.empty((6, 9, m, n)) if states else None
        cnt = np.zeros((n, 4), dtype=np.int32) if counters else None
        ng = None
        if non_gravs is not None:
            ng_arr = non_gravs if isinstance(non_gravs, np.ndarray) else nongrav_array(non_gravs, n)
            ng = ng_arr.ctypes.data_as(C.POINTER(_lib.KhNonGrav))
        _lib.check(self.L.kh_radec_partials(
            self.h, n, _lib.dptr(jd0), _lib.ptr_array(list(xyz)), frame, ng, m, _lib.dptr(jo),
            _lib.ptr_array(list(ob)), force, int(restart), epsilon, _lib.ptr_array(list(radec)),
            _lib.ptr_array([jac[0].reshape(-1), jac[1].reshape(-1)]),
            _lib.ptr_array(list(state)) if states else None,
            _lib.ptr_array([dstate[k].reshape(-1) for k in range(6)]) if states else None,
            _lib.iptr(st, C.c_int32),
            cnt.ctypes.data_as(C.POINTER(_lib.KhCounters)) if cnt is not None else None))
        out = dict(radec=radec, jac=jac, status=st, counters=cnt)
        if states:
            out.update(state=state, dstate=dstate)
        return out

    def last_kernel_ms(self) -> float:
        return self.L.kh_last_kernel_ms(self.h)

    def last_fov_prefilter(self):
        """(milliseconds, launches, state-field pairs) of the timed prefilter
        launches of the last fov_state_check on this context."""
        ms, nl, pairs = C.c_double(), C.c_int64(), C.c_int64()
        _lib.check(self.L.kh_fov_last_prefilter(self.h, C.byref(ms), C.byref(nl),
                                                C.byref(pairs)))
        return ms.value, nl.value, pairs.value

    def last_launch(self):
        b, g, m = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self.L.kh_last_launch(self.h, C.byref(b), C.byref(g), C.byref(m)))
        return b.value, g.value, m.value


_CTX: Context | None = None


def get_context() -> Context:
    """The default context, with the synthetic ephemeris loaded on first use:
    the planets (Type 2) and, as kete's core kernels hold them, the five
    perturbing asteroids (Type 21)."""
    global _CTX
    if _CTX is None:
        ctx = Context(0)
        from .synthetic import asteroid_ephemeris_path, ephemeris_path
        ctx.load_spk(ephemeris_path(), asteroid_ephemeris_path())
        _CTX = ctx
    return _CTX


def load_spk(*paths: str) -> Type2Table:
    """Replace the default context's ephemeris (kete ``spice.spk.load``)."""
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX.load_spk(*paths)


def spk_state(obj_id, jd, center=10, frame=FRAME_ECLIPTIC):
    out, st = get_context().spk_state(obj_id, jd, center, frame)
    return out, st


def _as_batch(states):
    if isinstance(states, StateArray):
        return states, True
    if isinstance(states, State):
        states = [states]
    return StateArray.from_states(states), False


def _per_frame_group(states, fn):
    """Apply fn(StateArray) -> StateArray to a state list whose members may
    differ in frame or center (kete handles each state on its own,
    src/kete/rust/propagation.rs:105-140): one batch per (frame, center),
    results back in input order.  fn also gets the members' indices."""
    if isinstance(states, StateArray):
        return fn(states, np.arange(len(states)))
    if isinstance(states, State):
        states = [states]
    states = list(states)
    if not states:
        raise ValueError("empty state list")
    keys = [(s.frame, s.center_id) for s in states]
    if len(set(keys)) == 1:
        return fn(StateArray.from_states(states), np.arange(len(states))).to_states()
    out = [None] * len(states)
    for key in dict.fromkeys(keys):
        idx = np.array([i for i, k in enumerate(keys) if k == key])
        res = fn(StateArray.from_states([states[i] for i in idx]), idx).to_states()
        for j, i in enumerate(idx):
            out[i] = res[j]
    return out


def propagate_two_body(states, jd, observer_pos=None):
    """kete ``propagate_two_body`` (``src/kete/rust/kepler.rs:68-108``): two-body
    motion about the Sun to ``jd``, with an optional one-step light delay as
    seen from ``observer_pos``.  Failed objects come back as NaN states."""
    def run(batch, idx):
        out, _st = get_context().two_body(batch, [jd], observer_pos)
        return StateArray(np.full(len(batch), float(jd)), out[:, 0, :], batch.frame,
                          batch.center_id, batch.desigs)
    return _per_frame_group(states, run)


# largest propagate_n_body_long batch: the batch and its nine planets in one
# wavefront of 64 lanes (kete accepts any size)
LONG_MAX_BATCH = 55


def _planet_array(planet_states, jd0):
    """The nine planet states of propagate_n_body_long as [9, 6] Sun-centred
    equatorial rows.  kete passes them to the integrator as given
    (propagation/mod.rs:186-190), i.e. it assumes equatorial input; states
    in the ecliptic frame are rotated here instead of being mixed in."""
    pb, _ = _as_batch(planet_states)
    if len(pb) != 9:
        raise ValueError("Input planet states must contain the correct number of states.")
    if pb.jd[0] != jd0:
        raise ValueError("Planet states JD must match JD of input state.")
    pl = np.empty((9, 6))
    c, sn = np.cos(0.40909280422232897), np.sin(0.40909280422232897)
    for k in range(9):
        p, v = pb.xyz[:3, k], pb.xyz[3:, k]
        if pb.frame == FRAME_ECLIPTIC:
            p = np.array([p[0], c * p[1] - sn * p[2], sn * p[1] + c * p[2]])
            v = np.array([v[0], c * v[1] - sn * v[2], sn * v[1] + c * v[2]])
        pl[k, :3], pl[k, 3:] = p, v
    return pl


def propagate_n_body_long(states, jd_final, planet_states=None, non_gravs=None, batch_size=10):
    """kete ``propagate_n_body_long`` (``src/kete/rust/propagation.rs:202-246``):
    the states (one epoch, Sun-centred) integrated in batches of
    ``batch_size`` together with the Sun, Mercury, Venus, the Earth-Moon
    barycentre and Mars to Neptune, with no ephemeris during the arc.
    Returns ``(states, planets)`` at ``jd_final``, Sun-centred in the
    equatorial frame as kete returns them; ``planets`` are the last batch's.
    ``planet_states``: nine Sun-centred states at the input epoch (any
    frame), else the loaded ephemeris' states.  Errors raise ValueError, as
    kete does for the whole call."""
    batch, soa = _as_batch(states)
    n = len(batch)
    if n == 0:
        raise ValueError("State vector is empty, propagation cannot continue")
    jd0 = float(batch.jd[0])
    if np.any(batch.jd != jd0):
        raise ValueError("All input states must have the same JD")
    if non_gravs is not None and len(non_gravs) != n:
        raise ValueError("Number of non-grav models doesnt match the number of provided objects.")
    if not 1 <= int(batch_size) <= LONG_MAX_BATCH:
        raise ValueError(f"batch_size must be 1 .. {LONG_MAX_BATCH} (a batch and its nine "
                         "planets share one wavefront)")
    pl = None if planet_states is None else _planet_array(planet_states, jd0)
    ctx = get_context()
    ng = None
    if non_gravs is not None:
        ng_arr = non_gravs if isinstance(non_gravs, np.ndarray) else nongrav_array(non_gravs, n)
        ng = ng_arr.ctypes.data_as(C.POINTER(_lib.KhNonGrav))
    out = np.empty((6, n))
    planets = np.empty((9, 6))
    st = C.c_int32(0)
    _lib.check(ctx.L.kh_propagate_n_body_long(
        ctx.h, n, jd0, _lib.ptr_array(list(batch.xyz)), batch.center_id, batch.frame,
        float(jd_final), _lib.dptr(pl.reshape(-1)) if pl is not None else None, ng,
        int(batch_size), _lib.ptr_array([out[k] for k in range(6)]),
        _lib.dptr(planets.reshape(-1)), C.byref(st)))
    if st.value != 0:
        raise ValueError(f"propagate_n_body_long: {STATUS_NAMES.get(st.value, st.value)}")
    res = StateArray(np.full(n, float(jd_final)), out, FRAME_EQUATORIAL, 10, batch.desigs)
    pls = StateArray(np.full(9, float(jd_final)), np.ascontiguousarray(planets.T),
                     FRAME_EQUATORIAL, 10)
    return (res, pls) if soa else (res.to_states(), pls.to_states())


_OBLIQUITY = 0.40909280422232897


def _rotate_frame(xyz, frame_from, frame_to):
    """Rotate [6, n] states between the equatorial (1) and ecliptic (17) frames."""
    if frame_from == frame_to:
        return xyz
    c, sn = np.cos(_OBLIQUITY), np.sin(_OBLIQUITY)
    sn = sn if frame_from == FRAME_ECLIPTIC else -sn  # ecliptic -> equatorial, or back
    out = np.empty_like(xyz)
    for o in (0, 3):
        out[o] = xyz[o]
        out[o + 1] = c * xyz[o + 1] - sn * xyz[o + 2]
        out[o + 2] = sn * xyz[o + 1] + c * xyz[o + 2]
    return out


def moid(state_a, state_b=None):
    """kete ``moid`` (``src/kete/rust/propagation.rs:27-39``,
    ``propagation/kepler.rs:325-404``): the minimum orbit intersection
    distance (AU) between the orbit of ``state_a`` and that of ``state_b``,
    by default the Earth's at ``state_a``'s epoch.  ``state_a`` may be one
    State (a float is returned, errors raise ValueError as kete's do) or a
    batch (State list or StateArray: an array with NaN where a pair fails);
    ``state_b`` one State for every pair, or a batch of the same length."""
    batch, soa = _as_batch(state_a)
    single = isinstance(state_a, State)
    n = len(batch)
    ctx = get_context()
    jd_b = xb = None
    if state_b is not None:
        bb, _ = _as_batch(state_b)
        if len(bb) == 1 and n != 1:
            bb = StateArray(np.full(n, bb.jd[0]), np.repeat(bb.xyz, n, axis=1), bb.frame,
                            bb.center_id)
        if len(bb) != n:
            raise ValueError("state_b must be one state or one per state_a")
        jd_b = np.ascontiguousarray(bb.jd)
        xb = np.ascontiguousarray(_rotate_frame(bb.xyz, bb.frame, batch.frame))
    out = np.empty(n)
    st = np.empty(n, dtype=np.int32)
    if n:
        _lib.check(ctx.L.kh_moid(
            ctx.h, n, _lib.dptr(batch.jd), _lib.ptr_array(list(batch.xyz)), batch.frame,
            _lib.dptr(jd_b) if jd_b is not None else None,
            _lib.ptr_array(list(xb)) if xb is not None else None, _lib.dptr(out),
            _lib.iptr(st, C.c_int32)))
    if single:
        if st[0] != 0:
            raise ValueError(f"moid: {STATUS_NAMES.get(int(st[0]), int(st[0]))}")
        return float(out[0])
    return out


def propagate_n_body(states, jd, include_asteroids=False, non_gravs=None, suppress_errors=True,
                     suppress_impact_errors=False):
    """kete ``propagate_n_body`` (``src/kete/rust/propagation.rs:70-151``):
    Radau-15 integration with the planets, Moon and solar/Jovian GR+J2 read from
    the loaded SPK, and with ``include_asteroids`` Ceres, Pallas, Vesta,
    Hygiea and Interamnia from their Type 21 segments.  Failed objects come back as NaN states; with
    ``suppress_errors=False`` the first failure raises ``ValueError``."""
    n_all = len(states) if not isinstance(states, State) else 1
    if non_gravs is not None and len(non_gravs) != n_all:
        raise ValueError("non_gravs must be the same length as states.")

    def run(batch, idx):
        if non_gravs is None:
            ng = None
        elif isinstance(non_gravs, np.ndarray):
            ng = np.ascontiguousarray(non_gravs[idx])
        else:
            ng = [non_gravs[i] for i in idx]
        r = get_context().n_body(batch, [jd], non_gravs=ng, include_asteroids=include_asteroids)
        st = r["status"]
        if not suppress_errors and np.any(st != 0):
            i = int(np.flatnonzero(st != 0)[0])
            raise ValueError(f"object {int(idx[i])}: {STATUS_NAMES.get(int(st[i]), st[i])}")
        if not suppress_impact_errors:
            from .naming import impact_message
            for i in np.flatnonzero(st == 6):
                d = batch.desigs[i] if batch.desigs else None
                print(impact_message(d, int(r["impact_id"][i]), float(r["impact_jd"][i])),
                      file=sys.stderr)
        return StateArray(np.full(len(batch), float(jd)), r["out"], batch.frame, batch.center_id,
                          batch.desigs)
    return _per_frame_group(states, run)
