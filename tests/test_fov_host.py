"""Host-side field construction: the library's rectangle normals equal the
oracle's (OnSkyRectangle::new, fov/patches.rs:122-155) bit for bit, and
satisfy kete's own rectangle known answers (patches.rs:366-396)."""
import math

import numpy as np

import oracle as o
from kete_amd import _lib


def lib_normals(p, rot, lon, lat):
    L = _lib.load()
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(12)
    L.kh_fov_rect_normals(_lib.dptr(p), rot, lon, lat, _lib.dptr(out))
    return out.reshape(4, 3)


def test_normals_match_oracle_bitwise():
    rng = np.random.default_rng(4)
    for _ in range(200):
        p = rng.normal(size=3)
        p /= np.linalg.norm(p)
        rot = rng.uniform(-3, 3) if rng.uniform() < 0.7 else 0.0
        w = math.radians(6.9)
        assert np.array_equal(lib_normals(p, rot, w, w), o.rect_normals(p, rot, w, w))


def test_library_normals_known_answers():
    nrm = lib_normals([1.0, 0, 0], 0.0, 0.1, 0.1)
    assert o.rect_contains(nrm, [1.0, 0.01, 0.01])[0]
    assert not o.rect_contains(nrm, [1.0, 0.1, 0.0])[0]
    rot = lib_normals([1.0, 0, 0], math.radians(45), 0.1, 0.1)
    assert not o.rect_contains(rot, [1.0, math.sin(0.05) * 0.99, math.sin(0.05) * 0.99])[0]
