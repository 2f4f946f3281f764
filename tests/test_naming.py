"""kete's impact report text: NAIF names and UTC ISO times, with kete's own
known answers (time/mod.rs:219-300, time/leap_second.rs:70-100)."""
import numpy as np
import pytest

from kete_amd import naming


def test_year_month_day_known_answers():
    assert naming._year_month_day(2451545.0) == (2000, 1, 1, 0.5)
    assert naming._year_month_day(2000000.0) == (763, 9, 18, 0.5)
    y, m, d, _ = naming._year_month_day(-68774.4991992591)
    assert (y, m, d) == (-4901, 8, 8)


def test_leap_second_lookup_known_answers():
    assert naming.LEAP_SECONDS[0] == (41317.0, 10)
    assert naming.LEAP_SECONDS[27] == (57754.0, 37)
    for mjd, s in ((0.0, 0), (41317.0, 10), (41317.1, 10), (57753.9, 36), (57754.0, 37),
                   (57755.0, 37)):
        assert naming.tai_to_utc_offset(mjd) == s / 86400.0


def test_iso_known_answer():
    assert naming.utc_iso(2447154.5) == "1987-12-25T00:00:00+00:00"
    assert naming.utc_iso(2451544.75) == "2000-01-01T06:00:00+00:00"


def test_tdb_to_utc_offsets():
    # after 2017: UTC = TDB - 32.184 s - 37 s
    jd = 2460000.5
    assert abs((jd - naming.tdb_to_utc_jd(jd)) * 86400.0 - 69.184) < 1e-4
    # before 1972 kete's table gives no leap seconds: only the 32.184 s
    assert abs((2440000.5 - naming.tdb_to_utc_jd(2440000.5)) * 86400.0 - 32.184) < 1e-4


def test_impact_message_format():
    msg = naming.impact_message("2024 YR4", 399, 2460000.5)
    assert msg == ('Impact detected between (Name("2024 YR4")) <-> (earth) at time 2460000.5 '
                   '(2023-02-24T23:58:50.815+00:00)')
    assert naming.naif_name(20000704) == "interamnia"
    assert naming.naif_name(12345) == "12345"
    assert naming.desig_debug(None) == "Empty"
