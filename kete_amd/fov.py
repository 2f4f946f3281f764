"""Field-of-view checks: ``RectangleFOV`` and ``fov_state_check``.

Mirrors kete's ``RectangleFOV`` (``src/kete/rust/fovs/definitions.rs:
181-196,401-480``) and ``fov_state_check`` (``src/kete/rust/fovs/checks.rs:
24-110``) over the C ABI; the work runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .spk import FRAME_ECLIPTIC, FRAME_EQUATORIAL

_OBL = 0.40909280422232897


def _equ_to_ecl(v):
    c, s = math.cos(_OBL), math.sin(_OBL)
    v = np.asarray(v, dtype=np.float64)
    return np.array([v[0], c * v[1] + s * v[2], -s * v[1] + c * v[2]])


@dataclass
class RectangleFOV:
    """A rectangular field: pointing vector, rotation, widths (degrees, as
    kete's constructor takes them), and the observer's Sun-centred state.
    A field read from a kete file carries its four edge normals instead
    (``from_normals``)."""
    pointing: np.ndarray
    rotation: float
    observer_jd: float
    observer_pos: np.ndarray
    observer_vel: np.ndarray
    lon_width: float
    lat_width: float
    frame: int = FRAME_ECLIPTIC
    observer_desig: object = None
    observer_center: int = 10
    edge_normals: np.ndarray = None

    @classmethod
    def from_normals(cls, normals, rotation, observer_jd, observer_pos, observer_vel,
                     frame=FRAME_ECLIPTIC, observer_desig=None, observer_center=10):
        """A field given by its four great-circle edge normals (kete's
        SphericalPolygon<4>, ``fov/patches.rs:80-91``), rotation in degrees."""
        return cls(None, float(rotation), float(observer_jd),
                   np.asarray(observer_pos, dtype=np.float64),
                   np.asarray(observer_vel, dtype=np.float64), float("nan"), float("nan"),
                   frame, observer_desig, observer_center,
                   np.asarray(normals, dtype=np.float64).reshape(4, 3))

    def normals(self) -> np.ndarray:
        if self.edge_normals is not None:
            return self.edge_normals
        L = _lib.load()
        p = np.ascontiguousarray(self.pointing, dtype=np.float64).reshape(3)
        out = np.zeros(12)
        L.kh_fov_rect_normals(_lib.dptr(p), math.radians(self.rotation),
                              math.radians(self.lon_width), math.radians(self.lat_width),
                              _lib.dptr(out))
        return out.reshape(4, 3)

    def ecliptic(self):
        """Field and observer converted to the ecliptic frame
        (try_frame_change_mut, collection.rs:23-36)."""
        nrm = self.normals()
        if self.frame == FRAME_EQUATORIAL:
            nrm = np.array([_equ_to_ecl(n) for n in nrm])
            return nrm, _equ_to_ecl(self.observer_pos), _equ_to_ecl(self.observer_vel)
        return nrm, np.asarray(self.observer_pos, float), np.asarray(self.observer_vel, float)

    @property
    def jd(self):
        return self.observer_jd


@dataclass
class FovHits:
    """Visible objects: one row per (field, object), ordered as kete returns
    them (fields by epoch, objects by input index)."""
    field: np.ndarray
    obj: np.ndarray
    state: np.ndarray  # [7, n_hits]: jd, pos, vel (ecliptic, Sun-centred)

    def __len__(self):
        return len(self.field)


def fov_state_check(obj_state, fovs, dt_limit=3.0, include_asteroids=False, ctx=None,
                    cap=None, fields=None) -> FovHits:
    """kete ``fov_state_check``: which objects are visible in which fields.

    ``obj_state`` is a :class:`~kete_amd.StateArray` (one epoch, Sun-centred,
    ecliptic); ``fovs`` a list of :class:`RectangleFOV`, or ``fields`` a
    prepared tuple ``(jd[nf], obs[6, nf], normals[nf, 12])`` in the ecliptic
    frame.
    """
    from . import get_context
    ctx = ctx or get_context()
    L = ctx.L
    if obj_state.frame != FRAME_ECLIPTIC or obj_state.center_id != 10:
        raise ValueError("states must be Sun-centred (10) in the ecliptic frame")
    jd0 = float(obj_state.jd[0])
    if np.any(obj_state.jd != jd0):
        raise ValueError("all states must share one epoch (SimultaneousStates)")
    if fields is None:
        nf = len(fovs)
        fjd = np.array([f.jd for f in fovs], dtype=np.float64)
        obs = np.empty((6, nf))
        nrm = np.empty((nf, 12))
        for k, f in enumerate(fovs):
            n_, p_, v_ = f.ecliptic()
            nrm[k] = n_.reshape(12)
            obs[:3, k] = p_
            obs[3:, k] = v_
    else:
        fjd, obs, nrm = fields
        fjd = np.ascontiguousarray(fjd, dtype=np.float64)
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        nrm = np.ascontiguousarray(nrm, dtype=np.float64).reshape(-1, 12)
        nf = len(fjd)
    n = len(obj_state)
    # hit capacity: a call that overflows reports the size needed and is run
    # again, so start above the usual density (a 6.9 deg field covers 1.2e-3
    # of the sky)
    cap = cap or int(min(n * max(nf, 1), max(1 << 16, 2e-3 * n * nf + 4096)))
    while True:
        hf = np.empty(cap, dtype=np.int64)
        ho = np.empty(cap, dtype=np.int64)
        hs = np.empty((7, cap))
        nh = C.c_int64(0)
        rc = L.kh_fov_state_check(ctx.h, n, jd0, _lib.ptr_array(list(obj_state.xyz)),
                                  FRAME_ECLIPTIC, nf, _lib.dptr(fjd),
                                  _lib.ptr_array([obs[k] for k in range(6)]),
                                  _lib.dptr(nrm.reshape(-1)), float(dt_limit),
                                  int(bool(include_asteroids)), cap, C.byref(nh),
                                  _lib.iptr(hf, C.c_int64), _lib.iptr(ho, C.c_int64),
                                  _lib.ptr_array([hs[k] for k in range(7)]))
        if rc == -1 and nh.value > cap:
            cap = nh.value
            continue
        _lib.check(rc)
        k = nh.value
        # views of the first k entries (the untouched rest of the capacity
        # is never paged in)
        return FovHits(hf[:k], ho[:k], hs[:, :k])


def to_simultaneous_states(hits: FovHits, fovs, desigs=None):
    """kete's result form: one SimultaneousStates per field with hits, in
    field order, holding the visible objects' states (Sun-centred ecliptic,
    at the light-time-corrected epoch) and the field
    (``src/kete/rust/fovs/checks.rs:96-109``).  desigs[i] names object i."""
    from . import State
    from .io import SimultaneousStates
    out = []
    if len(hits) == 0:
        return out
    bounds = np.flatnonzero(np.diff(hits.field)) + 1
    for lo, hi in zip(np.r_[0, bounds], np.r_[bounds, len(hits)]):
        f = int(hits.field[lo])
        states = [State(desigs[int(hits.obj[k])] if desigs is not None else None,
                        float(hits.state[0, k]), hits.state[1:4, k].copy(),
                        hits.state[4:7, k].copy(), FRAME_ECLIPTIC, 10) for k in range(lo, hi)]
        fov = fovs[f] if fovs is not None else None
        out.append(SimultaneousStates(states, fov.jd if fov is not None else states[0].jd, 10,
                                      FRAME_ECLIPTIC, fov))
    return out


def fov_spk_check(obj_ids, fovs, ctx=None, fields=None, cap=None) -> FovHits:
    """kete ``fov_spk_check`` (``src/kete/rust/fovs/checks.rs:123-142``):
    which objects of the loaded ephemeris (NAIF ids, e.g. the Type 21 small
    bodies) are visible in which fields, with light time.  ``obj`` of the
    result indexes ``obj_ids``."""
    from . import get_context
    ctx = ctx or get_context()
    L = ctx.L
    ids = np.ascontiguousarray(obj_ids, dtype=np.int64)
    if fields is None:
        nf = len(fovs)
        fjd = np.array([f.jd for f in fovs], dtype=np.float64)
        obs = np.empty((6, nf))
        nrm = np.empty((nf, 12))
        for k, f in enumerate(fovs):
            n_, p_, v_ = f.ecliptic()
            nrm[k] = n_.reshape(12)
            obs[:3, k] = p_
            obs[3:, k] = v_
    else:
        fjd, obs, nrm = fields
        fjd = np.ascontiguousarray(fjd, dtype=np.float64)
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        nrm = np.ascontiguousarray(nrm, dtype=np.float64).reshape(-1, 12)
        nf = len(fjd)
    cap = cap or max(1024, min(len(ids) * max(nf, 1), 1 << 22))
    while True:
        hf = np.empty(cap, dtype=np.int64)
        ho = np.empty(cap, dtype=np.int64)
        hs = np.empty((7, cap))
        nh = C.c_int64(0)
        rc = L.kh_fov_spk_check(ctx.h, len(ids), _lib.iptr(ids, C.c_int64), FRAME_ECLIPTIC, nf,
                                _lib.dptr(fjd), _lib.ptr_array([obs[k] for k in range(6)]),
                                _lib.dptr(nrm.reshape(-1)), cap, C.byref(nh),
                                _lib.iptr(hf, C.c_int64), _lib.iptr(ho, C.c_int64),
                                _lib.ptr_array([hs[k] for k in range(7)]))
        if rc == -1 and nh.value > cap:
            cap = nh.value
            continue
        _lib.check(rc)
        k = nh.value
        return FovHits(hf[:k], ho[:k], hs[:, :k])


def ra_dec_vmag(ctx, state, obs, h_mag, g_param=0.15, frame=FRAME_ECLIPTIC):
    """RA, Dec, RA rate x cos(Dec), Dec rate (rad, rad/day) and V magnitude per
    detection -> array [5, n]."""
    state = np.ascontiguousarray(state, dtype=np.float64)
    obs = np.ascontiguousarray(obs, dtype=np.float64)
    h = np.ascontiguousarray(h_mag, dtype=np.float64)
    n = state.shape[1]
    out = np.empty((5, n))
    _lib.check(ctx.L.kh_radec_vmag(ctx.h, n, _lib.ptr_array(list(state)),
                                   _lib.ptr_array(list(obs)), frame, _lib.dptr(h),
                                   float(g_param), _lib.ptr_array([out[k] for k in range(5)])))
    return out
