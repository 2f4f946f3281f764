"""Text of kete's impact report (src/kete/rust/propagation.rs:129-138):
the body's NAIF name (spice::try_name_from_id over kete's naif_ids.csv,
src/kete_core/src/spice/naif_ids.rs:47-55) and the impact time in UTC as
an RFC 3339 string (Time::utc().to_iso(), src/kete_core/src/time/mod.rs:
100-160, scales.rs:60-84, leap_second.rs:62-68)."""
from __future__ import annotations

# the bodies the force models can report an impact with: kete's names
NAIF_NAMES = {
    1: "mercury barycenter", 2: "venus barycenter", 3: "earth barycenter",
    4: "mars barycenter", 5: "jupiter barycenter", 6: "saturn barycenter",
    7: "uranus barycenter", 8: "neptune barycenter", 9: "pluto barycenter", 10: "sun",
    199: "mercury", 299: "venus", 301: "moon", 399: "earth", 499: "mars", 599: "jupiter",
    699: "saturn", 799: "uranus", 899: "neptune",
    2000001: "ceres", 2000002: "pallas", 2000004: "vesta",
    20000001: "ceres", 20000002: "pallas", 20000004: "vesta", 20000010: "hygiea",
    20000704: "interamnia",
}

# TAI - UTC (s) from each MJD on (the IERS leap seconds since 1972); 0 before
LEAP_SECONDS = [
    (41317.0, 10), (41499.0, 11), (41683.0, 12), (42048.0, 13), (42413.0, 14),
    (42778.0, 15), (43144.0, 16), (43509.0, 17), (43874.0, 18), (44239.0, 19),
    (44786.0, 20), (45151.0, 21), (45516.0, 22), (46247.0, 23), (47161.0, 24),
    (47892.0, 25), (48257.0, 26), (48804.0, 27), (49169.0, 28), (49534.0, 29),
    (50083.0, 30), (50630.0, 31), (51179.0, 32), (53736.0, 33), (54832.0, 34),
    (56109.0, 35), (57204.0, 36), (57754.0, 37),
]
TT_TO_TAI = 32.184 / 86400.0
JD_TO_MJD = -2400000.5


def naif_name(naif_id: int) -> str:
    """try_name_from_id(id).unwrap_or(id.to_string())."""
    return NAIF_NAMES.get(int(naif_id), str(int(naif_id)))


def tai_to_utc_offset(mjd: float) -> float:
    """TAI - UTC in days at a TAI MJD: the entry at or before it (exact
    matches included), 0 before the table."""
    off = 0.0
    for m, s in LEAP_SECONDS:
        if m <= mjd:
            off = s
        else:
            break
    return off / 86400.0


def tdb_to_utc_jd(jd: float) -> float:
    """UTC::from_tdb: TDB is taken as TT; TT - 32.184 s is TAI; minus the
    leap seconds at that TAI."""
    jd = jd - TT_TO_TAI
    return jd - tai_to_utc_offset(jd + JD_TO_MJD)


def _year_month_day(jd: float):
    """Fliegel and Van Flandern's calendar algorithm (CACM 11, 657, 1968)
    -> (year, month, day, fraction of the day)."""
    offset = jd + 0.5
    frac = offset % 1.0
    l = int(offset // 1.0) + 68569
    n = (4 * l) // 146097
    l -= (146097 * n + 3) // 4
    i = (4000 * (l + 1)) // 1461001
    l -= (1461 * i) // 4 - 31
    k = (80 * l) // 2447
    day = l - (2447 * k) // 80
    l = k // 11
    month = k + 2 - 12 * l
    year = 100 * (n - 49) + i + l
    return year, month, day, frac


def utc_iso(jd_utc: float) -> str:
    """RFC 3339 text of a UTC Julian date at millisecond resolution (each
    field truncated), '+00:00' offset, fractional seconds only when the
    milliseconds are not zero -- chrono's to_rfc3339 of kete's datetime."""
    y, mo, d, f = _year_month_day(jd_utc)
    f *= 24.0
    h = int(f)
    f = (f - h) * 60.0
    mi = int(f)
    f = (f - mi) * 60.0
    s = int(f)
    ms = int((f - s) * 1000.0)
    frac = f".{ms:03d}" if ms else ""
    return f"{y:04d}-{mo:02d}-{d:02d}T{h:02d}:{mi:02d}:{s:02d}{frac}+00:00"


def desig_debug(d) -> str:
    """Rust's {:?} of kete's Desig for a designation given as a Python value."""
    if d is None:
        return "Empty"
    if isinstance(d, (int,)) and not isinstance(d, bool):
        return f"Perm({d})"
    return f'Name("{d}")'


def impact_message(desig, naif_id: int, jd_tdb: float) -> str:
    return (f"Impact detected between ({desig_debug(desig)}) <-> ({naif_name(naif_id)}) at time "
            f"{jd_tdb} ({utc_iso(tdb_to_utc_jd(jd_tdb))})")
