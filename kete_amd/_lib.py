"""ctypes binding of libkete_hip.so (the C ABI in include/kete_hip.h).

The library is built in-tree by ``kete_amd/csrc/Makefile``.  There is no CPU
fallback: if the library is missing or no GPU is visible, every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkete_hip.so")
CSRC = os.path.join(HERE, "csrc")

KH_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP runtime failure", -3: "no ephemeris",
          -4: "no GPU visible"}

# Symbols declared in include/kete_hip.h
EXPORTS = [
    "kh_ctx_create", "kh_ctx_destroy", "kh_last_error", "kh_ctx_stream",
    "kh_spk_upload_type2", "kh_spk_upload_type21", "kh_spk_states", "kh_propagate_two_body",
    "kh_propagate_two_body_dev", "kh_propagate_n_body", "kh_propagate_n_body_dev",
    "kh_last_launch", "kh_last_kernel_ms", "kh_fov_rect_normals", "kh_fov_state_check",
    "kh_radec_vmag", "kh_radec_partials", "kh_radec_partials_dev", "kh_autotune_n_body",
    "kh_autotune_clear", "kh_fov_spk_check", "kh_propagate_n_body_long",
    "kh_propagate_n_body_long_dev", "kh_moid", "kh_moid_dev",
    "kh_propagate_n_body_dev_noalloc", "kh_fov_last_prefilter",
]

# Largest batch of kh_propagate_n_body_dev_noalloc (KH_NOALLOC_MAX_N)
NOALLOC_MAX_N = 2097152


class KhSpk2Seg(C.Structure):
    _fields_ = [("obj_id", C.c_int64), ("center_id", C.c_int64), ("frame", C.c_int32),
                ("rsize", C.c_int32), ("n_rec", C.c_int64), ("jd_start", C.c_double),
                ("jd_end", C.c_double), ("intlen_s", C.c_double),
                ("records", C.POINTER(C.c_double))]


class KhSpk21Seg(C.Structure):
    _fields_ = [("obj_id", C.c_int64), ("center_id", C.c_int64), ("frame", C.c_int32),
                ("maxdim", C.c_int32), ("n_rec", C.c_int64), ("jd_start", C.c_double),
                ("jd_end", C.c_double), ("records", C.POINTER(C.c_double)),
                ("epochs", C.POINTER(C.c_double))]


class KhNonGrav(C.Structure):
    _fields_ = [("kind", C.c_int32)] + [(k, C.c_double) for k in
                                        ("a1", "a2", "a3", "alpha", "r_0", "m", "n", "k", "dt",
                                         "beta")]


class KhCounters(C.Structure):
    _fields_ = [("n_force", C.c_int32), ("n_sweep", C.c_int32), ("n_acc", C.c_int32),
                ("n_fail", C.c_int32)]


class KhTuneResult(C.Structure):
    _fields_ = [("mode", C.c_int32), ("block", C.c_int32), ("window", C.c_int32),
                ("n_tried", C.c_int32), ("latency_k", C.c_int64), ("n_sample", C.c_int64),
                ("ms", C.c_double), ("default_ms", C.c_double)]


NONGRAV_DTYPE = np.dtype([("kind", np.int32), ("pad", np.int32)] +
                         [(k, np.float64) for k in ("a1", "a2", "a3", "alpha", "r_0", "m", "n",
                                                    "k", "dt", "beta")])
assert NONGRAV_DTYPE.itemsize == C.sizeof(KhNonGrav)

_lib = None


class KeteHipError(RuntimeError):
    pass


def build(force: bool = False) -> str:
    """Compile libkete_hip.so for gfx950 with hipcc (kete_amd/csrc/Makefile)."""
    cmd = ["make", "-s", "-C", CSRC, "-j8"]
    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True)
    subprocess.run(cmd, check=True)
    return LIB_PATH


def load():
    """Load the library (raising if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KeteHipError(f"{LIB_PATH} is not built; run kete_amd._lib.build() "
                           "(hipcc --offload-arch=gfx950)")
    # One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64
    # (soname libamdhip64.so.7, linked by the unversioned name): with torch
    # loaded first, this library binds to that copy; loaded the other way
    # round, torch would load a second runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P, PP, I32, I64, D, V = (C.POINTER(C.c_double), C.POINTER(C.POINTER(C.c_double)), C.c_int32,
                             C.c_int64, C.c_double, C.c_void_p)
    for name in EXPORTS:
        getattr(L, name)
    L.kh_ctx_create.argtypes = [I32, C.POINTER(V)]
    L.kh_ctx_create.restype = I32
    L.kh_ctx_destroy.argtypes = [V]
    L.kh_ctx_destroy.restype = I32
    L.kh_last_error.restype = C.c_char_p
    L.kh_ctx_stream.argtypes = [V]
    L.kh_ctx_stream.restype = V
    L.kh_spk_upload_type2.argtypes = [V, C.POINTER(KhSpk2Seg), I64]
    L.kh_spk_upload_type2.restype = I32
    L.kh_spk_upload_type21.argtypes = [V, C.POINTER(KhSpk21Seg), I64]
    L.kh_spk_upload_type21.restype = I32
    L.kh_spk_states.argtypes = [V, I64, I64, I32, I64, P, PP, C.POINTER(I32)]
    L.kh_spk_states.restype = I32
    tb = [V, I64, P, PP, I64, I32, I64, P, P, PP, C.POINTER(I32)]
    L.kh_propagate_two_body.argtypes = tb
    L.kh_propagate_two_body.restype = I32
    L.kh_propagate_two_body_dev.argtypes = tb + [C.POINTER(I32), I32]
    L.kh_propagate_two_body_dev.restype = I32
    nb = [V, I64, P, PP, I64, I32, P, I32, I32, C.POINTER(KhNonGrav), D, I32, PP,
          C.POINTER(I32), C.POINTER(I64), P, C.POINTER(KhCounters)]
    L.kh_propagate_n_body.argtypes = nb
    L.kh_propagate_n_body.restype = I32
    L.kh_propagate_n_body_dev.argtypes = nb + [I32]
    L.kh_propagate_n_body_dev.restype = I32
    L.kh_propagate_n_body_dev_noalloc.argtypes = nb + [I32]
    L.kh_propagate_n_body_dev_noalloc.restype = I32
    L.kh_last_launch.argtypes = [V, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]
    L.kh_last_launch.restype = I32
    L.kh_fov_rect_normals.argtypes = [P, D, D, D, P]
    L.kh_fov_rect_normals.restype = None
    L.kh_fov_state_check.argtypes = [V, I64, D, PP, I32, I64, P, PP, P, D, I32, I64,
                                     C.POINTER(I64), C.POINTER(I64), C.POINTER(I64), PP]
    L.kh_fov_state_check.restype = I32
    L.kh_fov_last_prefilter.argtypes = [V, C.POINTER(D), C.POINTER(I64), C.POINTER(I64)]
    L.kh_fov_last_prefilter.restype = I32
    L.kh_fov_spk_check.argtypes = [V, I64, C.POINTER(I64), I32, I64, P, PP, P, I64,
                                   C.POINTER(I64), C.POINTER(I64), C.POINTER(I64), PP]
    L.kh_fov_spk_check.restype = I32
    L.kh_propagate_n_body_long.argtypes = [V, I64, D, PP, I64, I32, D, P, C.POINTER(KhNonGrav),
                                           I32, PP, P, C.POINTER(I32)]
    L.kh_propagate_n_body_long.restype = I32
    L.kh_propagate_n_body_long_dev.argtypes = [V, I64, D, PP, I64, I32, D, P,
                                               C.POINTER(KhNonGrav), I32, PP, P,
                                               C.POINTER(I32), C.POINTER(KhCounters), I32]
    L.kh_propagate_n_body_long_dev.restype = I32
    L.kh_moid.argtypes = [V, I64, P, PP, I32, P, PP, P, C.POINTER(I32)]
    L.kh_moid.restype = I32
    L.kh_moid_dev.argtypes = [V, I64, P, PP, I32, P, PP, P, C.POINTER(I32), P, C.POINTER(I32),
                              C.POINTER(I32), I32]
    L.kh_moid_dev.restype = I32
    L.kh_radec_vmag.argtypes = [V, I64, PP, PP, I32, P, D, PP]
    L.kh_radec_vmag.restype = I32
    rp = [V, I64, P, PP, I32, C.POINTER(KhNonGrav), I32, P, PP, I32, I32, D, PP, PP, PP, PP,
          C.POINTER(I32), C.POINTER(KhCounters)]
    L.kh_radec_partials.argtypes = rp
    L.kh_radec_partials.restype = I32
    L.kh_radec_partials_dev.argtypes = rp + [I32]
    L.kh_radec_partials_dev.restype = I32
    L.kh_autotune_n_body.argtypes = [V, I64, P, PP, I64, I32, P, I32, I32, I64, I32,
                                     C.POINTER(KhTuneResult)]
    L.kh_autotune_n_body.restype = I32
    L.kh_autotune_clear.argtypes = [V]
    L.kh_autotune_clear.restype = I32
    L.kh_last_kernel_ms.argtypes = [V]
    L.kh_last_kernel_ms.restype = D
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != KH_OK:
        msg = load().kh_last_error().decode(errors="replace")
        raise KeteHipError(f"kete_hip error {rc} ({ERRORS.get(rc, '?')}): {msg}")


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray, ct):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def ptr_array(ptrs):
    """Array of double* from numpy arrays or raw device addresses (ints).
    The numpy arrays are kept alive with the pointer array (a pointer stored
    into a ctypes array does not hold its array, so a temporary passed here
    would otherwise be freed before the call reads it)."""
    arr = (C.POINTER(C.c_double) * len(ptrs))()
    for k, p in enumerate(ptrs):
        arr[k] = C.cast(C.c_void_p(p), C.POINTER(C.c_double)) if isinstance(p, int) else dptr(p)
    arr._keep = [p for p in ptrs if not isinstance(p, int)]
    return arr


def raw(p: int, ct=C.c_double):
    return C.cast(C.c_void_p(p), C.POINTER(ct)) if p else None
