"""kete's binary files: ``SimultaneousStates`` and ``FOVList``.

kete saves both through ``FileIO`` (``src/kete_core/src/io/mod.rs:15-50``):
serde + bincode 2 with ``config::legacy()`` -- little-endian fixed-width
integers, u64 lengths for strings and sequences, u32 enum variant indices,
a u8 tag for ``Option``, fixed-size arrays as bare tuples, struct fields in
declaration order.  The types written here, in their field order:

- ``SimultaneousStates`` (``simult_states.rs:14-25``): states: Vec<State>,
  jd: f64, center_id: i64, frame: Frame, fov: Option<FOV>.
- ``State`` (``state.rs:76-95``): desig: Desig, jd: f64, pos: [f64; 3],
  vel: [f64; 3], frame: Frame, center_id: i64.
- ``Desig`` (``state.rs:32-48``): Perm(u64) | Prov(String) | Name(String) |
  Naif(i64) | Empty.
- ``Frame`` (``frames/definitions.rs:20-45``): Ecliptic | Equatorial | FK4 |
  Galactic | Unknown(i32) | EclipticNonInertial(i32, [f64; 6]).
- ``FOV`` (``fov/mod.rs:25-49``), of which the rectangle this path checks is
  variant 4, ``GenericRectangle`` (``fov/generic.rs:14-22``): observer:
  State, patch: SphericalPolygon<4> {edge_normals: [[f64; 3]; 4], frame}
  (``fov/patches.rs:80-91``), rotation: f64 (radians).
- ``FOVList`` (``src/kete/rust/fovs/collection.rs:42-101``): Vec<FOV>.

Only the ecliptic and equatorial frames and the GenericRectangle field are
read and written (the states and fields this path handles); other variants
raise.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

from .spk import FRAME_ECLIPTIC, FRAME_EQUATORIAL

_DESIG = ("Perm", "Prov", "Name", "Naif", "Empty")
_FRAME_TAG = {FRAME_ECLIPTIC: 0, FRAME_EQUATORIAL: 1}
_TAG_FRAME = {0: FRAME_ECLIPTIC, 1: FRAME_EQUATORIAL}
_FOV_VARIANTS = ("Wise", "NeosCmos", "ZtfCcdQuad", "GenericCone", "GenericRectangle",
                 "ZtfField", "NeosVisit", "OmniDirectional")
_GENERIC_RECTANGLE = 4


@dataclass(frozen=True)
class Desig:
    """A kete designation with its variant ("Perm", "Prov", "Name", "Naif",
    "Empty"); plain strings are written as Name, ints as Naif, None as Empty."""
    kind: str
    value: object = None

    def __str__(self):
        return "" if self.value is None else str(self.value)


class _Writer:
    def __init__(self):
        self.parts = []

    def u8(self, v): self.parts.append(struct.pack("<B", v))
    def u32(self, v): self.parts.append(struct.pack("<I", v))
    def i32(self, v): self.parts.append(struct.pack("<i", v))
    def u64(self, v): self.parts.append(struct.pack("<Q", v))
    def i64(self, v): self.parts.append(struct.pack("<q", v))
    def f64(self, v): self.parts.append(struct.pack("<d", float(v)))

    def string(self, s):
        b = s.encode("utf-8")
        self.u64(len(b))
        self.parts.append(b)

    def bytes(self):
        return b"".join(self.parts)


class _Reader:
    def __init__(self, data: bytes):
        self.d = memoryview(data)
        self.o = 0

    def _take(self, fmt):
        v = struct.unpack_from(fmt, self.d, self.o)[0]
        self.o += struct.calcsize(fmt)
        return v

    def u8(self): return self._take("<B")
    def u32(self): return self._take("<I")
    def i32(self): return self._take("<i")
    def u64(self): return self._take("<Q")
    def i64(self): return self._take("<q")
    def f64(self): return self._take("<d")

    def string(self):
        n = self.u64()
        s = bytes(self.d[self.o:self.o + n]).decode("utf-8")
        self.o += n
        return s


def _write_desig(w, d):
    if d is None:
        d = Desig("Empty")
    elif isinstance(d, str):
        d = Desig("Name", d)
    elif isinstance(d, (int, np.integer)):
        d = Desig("Naif", int(d))
    w.u32(_DESIG.index(d.kind))
    if d.kind == "Perm":
        w.u64(d.value)
    elif d.kind in ("Prov", "Name"):
        w.string(d.value)
    elif d.kind == "Naif":
        w.i64(d.value)


def _read_desig(r):
    tag = r.u32()
    kind = _DESIG[tag]
    if kind == "Perm":
        return Desig(kind, r.u64())
    if kind in ("Prov", "Name"):
        return Desig(kind, r.string())
    if kind == "Naif":
        return Desig(kind, r.i64())
    return Desig(kind)


def _write_frame(w, frame):
    if frame not in _FRAME_TAG:
        raise ValueError(f"frame {frame} is not written (ecliptic or equatorial only)")
    w.u32(_FRAME_TAG[frame])


def _read_frame(r):
    tag = r.u32()
    if tag not in _TAG_FRAME:
        raise ValueError(f"frame variant {tag} is not read (ecliptic or equatorial only)")
    return _TAG_FRAME[tag]


def _write_state(w, desig, jd, pos, vel, frame, center_id):
    _write_desig(w, desig)
    w.f64(jd)
    for x in pos:
        w.f64(x)
    for x in vel:
        w.f64(x)
    _write_frame(w, frame)
    w.i64(center_id)


def _read_state(r):
    desig = _read_desig(r)
    jd = r.f64()
    pos = np.array([r.f64() for _ in range(3)])
    vel = np.array([r.f64() for _ in range(3)])
    frame = _read_frame(r)
    center = r.i64()
    return desig, jd, pos, vel, frame, center


def _write_fov(w, fov):
    from .fov import RectangleFOV
    if not isinstance(fov, RectangleFOV):
        raise ValueError("only RectangleFOV (kete's GenericRectangle) is written")
    w.u32(_GENERIC_RECTANGLE)
    _write_state(w, fov.observer_desig, fov.observer_jd, fov.observer_pos, fov.observer_vel,
                 fov.frame, fov.observer_center)
    for n in fov.normals():
        for x in n:
            w.f64(x)
    _write_frame(w, fov.frame)
    w.f64(np.radians(fov.rotation))


def _read_fov(r):
    from .fov import RectangleFOV
    tag = r.u32()
    if tag != _GENERIC_RECTANGLE:
        name = _FOV_VARIANTS[tag] if tag < len(_FOV_VARIANTS) else str(tag)
        raise ValueError(f"FOV variant {name} is not read (GenericRectangle only)")
    desig, jd, pos, vel, frame, center = _read_state(r)
    normals = np.array([[r.f64() for _ in range(3)] for _ in range(4)])
    patch_frame = _read_frame(r)
    rotation = r.f64()
    if patch_frame != frame:
        raise ValueError("field frame differs from its observer's")
    return RectangleFOV.from_normals(normals, np.degrees(rotation), jd, pos, vel, frame,
                                     observer_desig=desig, observer_center=center)


@dataclass
class SimultaneousStates:
    """kete's SimultaneousStates: states at one epoch (or seen by one field),
    one center and frame, the field they were seen in or None."""
    states: list = field(default_factory=list)   # kete_amd.State objects
    jd: float = 0.0
    center_id: int = 10
    frame: int = FRAME_ECLIPTIC
    fov: object = None

    def to_bytes(self) -> bytes:
        w = _Writer()
        w.u64(len(self.states))
        for s in self.states:
            _write_state(w, s.desig, s.jd, s.pos, s.vel, s.frame, s.center_id)
        w.f64(self.jd)
        w.i64(self.center_id)
        _write_frame(w, self.frame)
        if self.fov is None:
            w.u8(0)
        else:
            w.u8(1)
            _write_fov(w, self.fov)
        return w.bytes()

    @classmethod
    def from_bytes(cls, data: bytes):
        from . import State
        r = _Reader(data)
        states = []
        for _ in range(r.u64()):
            desig, jd, pos, vel, frame, center = _read_state(r)
            states.append(State(desig, jd, pos, vel, frame, center))
        jd = r.f64()
        center = r.i64()
        frame = _read_frame(r)
        fov = _read_fov(r) if r.u8() else None
        if r.o != len(data):
            raise ValueError("trailing bytes after SimultaneousStates")
        return cls(states, jd, center, frame, fov)

    def save(self, path):
        """FileIO::save (io/mod.rs:20-24)."""
        with open(path, "wb") as f:
            f.write(self.to_bytes())

    @classmethod
    def load(cls, path):
        """FileIO::load (io/mod.rs:26-30)."""
        with open(path, "rb") as f:
            return cls.from_bytes(f.read())


def fov_list_to_bytes(fovs) -> bytes:
    """FOVList::save (collection.rs:79-100): a Vec<FOV>."""
    w = _Writer()
    w.u64(len(fovs))
    for f in fovs:
        _write_fov(w, f)
    return w.bytes()


def fov_list_from_bytes(data: bytes):
    r = _Reader(data)
    out = [_read_fov(r) for _ in range(r.u64())]
    if r.o != len(data):
        raise ValueError("trailing bytes after FOVList")
    return out


def save_fov_list(fovs, path):
    with open(path, "wb") as f:
        f.write(fov_list_to_bytes(fovs))


def load_fov_list(path):
    with open(path, "rb") as f:
        return fov_list_from_bytes(f.read())
