"""DAF/SPK files with Type 2 and Type 21 segments: reading, writing, and the
device table layout.

Reading follows kete's DAF container reader (``spice/daf.rs:148-351``), its
Type 2 segment view (``spice/spk_segments.rs:316-397``) and its Type 21 view
(``spice/spk_segments.rs:958-1003``): a 1024-byte file record, summary
records chained by a forward pointer, 8-byte words addressed from 1, a
``[INIT, INTLEN, RSIZE, N]`` trailer on each Type 2 array, and on each Type 21
array ``N`` records of ``4*MAXDIM+11`` doubles, the ``N`` final epochs, a
directory of every 100th final epoch and a ``[MAXDIM, N]`` trailer.

The device table (:class:`Type2Table`) packs the records of every Type 2
segment one after another, ``[n_rec][rsize]`` per segment, so a kernel reads a
record as ``rsize`` consecutive doubles: ``[mid, radius, x[n], y[n], z[n]]``
in km and TDB seconds past J2000.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

FRAME_EQUATORIAL = 1
FRAME_ECLIPTIC = 17
J2000_JD = 2451545.0


def spice_jds_to_jd(sec: float) -> float:
    """Seconds past J2000 -> JD (TDB) (``spice/mod.rs:25-27``)."""
    return sec / 86400.0 + 2451545.0


def jd_to_spice_jd(jd: float) -> float:
    """JD (TDB) -> seconds past J2000 (``spice/mod.rs:30-34``)."""
    return (jd - 2451545.0) * 86400.0


@dataclass
class Segment:
    obj_id: int
    center_id: int
    frame: int
    seg_type: int
    start_s: float
    end_s: float
    data: np.ndarray = field(repr=False)

    @property
    def jd_start(self) -> float:
        return spice_jds_to_jd(self.start_s)

    @property
    def jd_end(self) -> float:
        return spice_jds_to_jd(self.end_s)

    # Type 2 trailer [INIT, INTLEN, RSIZE, N] (spk_segments.rs:379-397);
    # Type 21 trailer [MAXDIM, N] (spk_segments.rs:971-986)
    @property
    def rsize(self) -> int:
        if self.seg_type == 21:
            return 4 * self.n_coef + 11
        return int(self.data[-2])

    @property
    def intlen(self) -> float:
        return float(self.data[-3])

    @property
    def n_rec(self) -> int:
        return int(self.data[-1])

    @property
    def n_coef(self) -> int:
        """Chebyshev coefficients per axis (Type 2) or MAXDIM (Type 21)."""
        if self.seg_type == 21:
            return int(self.data[-2])
        return (self.rsize - 2) // 3

    def records(self) -> np.ndarray:
        return self.data[: self.n_rec * self.rsize].reshape(self.n_rec, self.rsize)

    def epochs(self) -> np.ndarray:
        """Type 21 final epochs, one per record (TDB seconds past J2000)."""
        if self.seg_type != 21:
            raise ValueError("final epochs exist for Type 21 segments only")
        o = self.n_rec * self.rsize
        return self.data[o: o + self.n_rec]


def type21_array(records: np.ndarray, epochs: np.ndarray) -> np.ndarray:
    """Type 21 segment array from ``[N][4*MAXDIM+11]`` records and their ``N``
    final epochs: records, epochs, the directory of every 100th final epoch,
    then ``[MAXDIM, N]`` (NAIF SPK Type 21 layout, as kete reads it in
    ``spk_segments.rs:971-1003``)."""
    records = np.ascontiguousarray(records, dtype=np.float64)
    n, rl = records.shape
    if (rl - 11) % 4 or len(epochs) != n:
        raise ValueError("records must be [N][4*MAXDIM+11] with N final epochs")
    epochs = np.asarray(epochs, dtype=np.float64)
    directory = epochs[99::100][: n // 100]
    return np.concatenate([records.reshape(-1), epochs, directory,
                           np.array([(rl - 11) // 4, n], dtype=np.float64)])


def read_spk(path: str) -> list[Segment]:
    """Parse every segment of a DAF/SPK file."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 1024 or not buf[:7] == b"DAF/SPK":
        raise ValueError(f"{path}: not a DAF/SPK file")
    fmt = buf[88:96].decode("ascii", "replace").lower()
    if fmt == "ltl-ieee":
        e = "<"
    elif fmt == "big-ieee":
        e = ">"
    else:
        raise ValueError("Expected little or big endian in DAF file, found neither")
    nd, ni = struct.unpack(e + "ii", buf[8:16])
    fward = struct.unpack(e + "i", buf[76:80])[0]
    if nd != 2 or ni != 6:
        raise ValueError("SPK File incorrectly Formatted.")
    ssize = nd + (ni + 1) // 2
    segs = []
    nxt = fward
    while nxt != 0:
        rec = buf[1024 * (nxt - 1): 1024 * nxt]
        nxt_f, _prev, nsum = struct.unpack(e + "ddd", rec[:24])
        nxt = int(nxt_f)
        for s in range(int(nsum)):
            o = 24 + s * ssize * 8
            f0, f1 = struct.unpack(e + "dd", rec[o: o + 16])
            ints = struct.unpack(e + "6i", rec[o + 16: o + 40])
            a0, a1 = ints[4], ints[5]
            data = np.frombuffer(buf, dtype=e + "f8", count=a1 - a0 + 1, offset=8 * (a0 - 1))
            segs.append(Segment(ints[0], ints[1], ints[2], ints[3], f0, f1,
                                np.array(data, dtype=np.float64)))
    return segs


def write_spk(path: str, segments: list[Segment], name: str = "SYNTHETIC") -> None:
    """Write segments (any type; ``data`` is the segment array) as a
    little-endian DAF/SPK file (NAIF DAF layout), in the order given."""
    nd, ni = 2, 6
    ssize = nd + (ni + 1) // 2
    if len(segments) > (1024 - 24) // (8 * ssize):
        raise ValueError("too many segments for one summary record")
    words = []
    addr = 3 * 128 + 1  # data starts at record 4
    summaries = []
    for sg in segments:
        d = np.ascontiguousarray(sg.data, dtype="<f8")
        a0, a1 = addr, addr + len(d) - 1
        summaries.append(struct.pack("<dd6i", sg.start_s, sg.end_s, sg.obj_id, sg.center_id,
                                     sg.frame, sg.seg_type, a0, a1))
        words.append(d.tobytes())
        addr = a1 + 1
    free = addr
    rec1 = bytearray(1024)
    rec1[0:8] = b"DAF/SPK "
    rec1[8:16] = struct.pack("<ii", nd, ni)
    rec1[16:76] = name.ljust(60)[:60].encode()
    rec1[76:88] = struct.pack("<iii", 2, 2, free)
    rec1[88:96] = b"LTL-IEEE"
    rec1[699:727] = b"FTPSTR:\r:\n:\r\n:\r\x00:\x81:\x10\xce:ENDFTP"[:28]
    rec2 = bytearray(1024)
    rec2[0:24] = struct.pack("<ddd", 0.0, 0.0, float(len(segments)))
    for k, s in enumerate(summaries):
        rec2[24 + 40 * k: 64 + 40 * k] = s
    rec3 = bytearray(1024)
    for k, sg in enumerate(segments):
        rec3[40 * k: 40 * k + 40] = f"{sg.obj_id} wrt {sg.center_id}".ljust(40).encode()
    body = b"".join(words)
    pad = (-len(body)) % 1024
    with open(path, "wb") as f:
        f.write(bytes(rec1) + bytes(rec2) + bytes(rec3) + body + b"\0" * pad)


write_spk_type2 = write_spk


class Type2Table:
    """Type 2 segments packed for the device.

    ``coef``: float64 ``[total_records * rsize]``.  ``meta``: per segment
    ``(obj_id, center_id, frame, rsize, n_coef, n_rec, rec_offset)`` as int64
    and ``(jd_start, jd_end, start_s, intlen_s)`` as float64.
    """

    def __init__(self, segments: list[Segment]):
        segs = [s for s in segments if s.seg_type == 2]
        if not segs:
            raise ValueError("no Type 2 segments")
        self.segments = segs
        offs, tot = [], 0
        for s in segs:
            offs.append(tot)
            tot += s.n_rec
        self.coef = np.concatenate([s.records().reshape(-1) for s in segs]).astype(np.float64)
        self.imeta = np.array([[s.obj_id, s.center_id, s.frame, s.rsize, s.n_coef, s.n_rec, o]
                               for s, o in zip(segs, offs)], dtype=np.int64)
        self.fmeta = np.array([[s.jd_start, s.jd_end, s.start_s, s.intlen] for s in segs],
                              dtype=np.float64)

    def find(self, obj_id: int) -> int:
        for k, s in enumerate(self.segments):
            if s.obj_id == obj_id:
                return k
        raise KeyError(obj_id)

    def state(self, obj_id: int, jd):
        """Position (AU) and velocity (AU/day) of ``obj_id`` w.r.t. its segment
        center, vectorised over ``jd`` (``spk_segments.rs:355-376``, evaluated
        by Clenshaw-free forward recurrences as ``interpolation.rs:25-77``)."""
        k = self.find(obj_id)
        s = self.segments[k]
        jd = np.atleast_1d(np.asarray(jd, dtype=np.float64))
        if np.any((jd < s.jd_start) | (jd > s.jd_end)):
            raise ValueError("JD is not present in this record.")
        js = (jd - 2451545.0) * 86400.0
        j0 = (s.jd_start - 2451545.0) * 86400.0
        ri = np.clip(np.floor((js - j0) / s.intlen).astype(np.int64), 0, s.n_rec - 1)
        rec = s.records()[ri]
        n = s.n_coef
        t = (js - rec[:, 0]) / rec[:, 1]
        scale = 86400.0 / rec[:, 1] / 149597870.7
        pos = np.empty((3, len(jd)))
        vel = np.empty((3, len(jd)))
        for ax in range(3):
            c = rec[:, 2 + ax * n: 2 + (ax + 1) * n]
            x2 = 2.0 * t
            val = c[:, 0] + c[:, 1] * t
            der = c[:, 1].copy()
            st, lt = np.ones_like(t), t.copy()
            su, lu = np.ones_like(t), x2.copy()
            for i in range(2, n):
                nt = x2 * lt - st
                val = val + c[:, i] * nt
                st, lt = lt, nt
                nu = x2 * lu - su
                der = der + c[:, i] * lu * float(i)
                su, lu = lu, nu
            pos[ax] = val / 149597870.7
            vel[ax] = der * scale
        return pos, vel
