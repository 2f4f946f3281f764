"""The C ABI library builds, loads, and exports every entry point the header
declares.  No compute calls: this runs without a GPU."""
import ctypes as C
import os
import re

import pytest

from kete_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kete_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kh_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_all_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        _lib.build()
    lib = C.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts():
    # kh_nongrav: int32 kind (padded to 8) + 10 doubles
    assert C.sizeof(_lib.KhNonGrav) == 8 + 8 * 10
    assert C.sizeof(_lib.KhCounters) == 16
    assert C.sizeof(_lib.KhSpk2Seg) == 8 + 8 + 4 + 4 + 8 + 8 + 8 + 8 + 8


def test_null_context_is_rejected():
    L = _lib.load()
    assert L.kh_last_kernel_ms(None) < 0
    assert L.kh_spk_upload_type2(None, None, 0) == -1
    assert b"null" in L.kh_last_error()


