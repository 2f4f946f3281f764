"""kh_propagate_n_body_dev_noalloc: the N-body entry point that allocates no
device memory (its order scratch is a device array built into the library)
returns exactly what kh_propagate_n_body_dev returns, on two calls in a row,
the second with a smaller batch (stale scratch from the first call must not
leak into the second), and rejects batches above its maximum before any
device work."""
import ctypes as C

import numpy as np
import pytest

import kete_amd as ka
from kete_amd import _lib
from kete_amd import synthetic as syn

JD0 = syn.EPOCH_JD


def test_rejects_batches_above_the_maximum():
    L = _lib.load()
    rc = L.kh_propagate_n_body_dev_noalloc(
        None, _lib.NOALLOC_MAX_N + 1, None, None, 10, 17, None, 0, 0, None, 1e-6, -1, None,
        None, None, None, None, 0)
    assert rc == -1
    assert str(_lib.NOALLOC_MAX_N).encode() in L.kh_last_error()


def _call(ctx, fn_name, n, xyz, jdf):
    import torch
    dev = torch.device("cuda:0")
    jd0 = torch.full((n,), JD0, dtype=torch.float64, device=dev)
    d_xyz = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev).contiguous()
    jf = torch.from_numpy(np.ascontiguousarray(jdf)).to(dev)
    out = torch.full((6, n), -7.0, dtype=torch.float64, device=dev)
    st = torch.full((n,), -7, dtype=torch.int32, device=dev)
    iid = torch.zeros(n, dtype=torch.int64, device=dev)
    ijd = torch.zeros(n, dtype=torch.float64, device=dev)
    cnt = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    xyz_p = _lib.ptr_array([d_xyz[k].data_ptr() for k in range(6)])
    out_p = _lib.ptr_array([out[k].data_ptr() for k in range(6)])
    rc = getattr(ctx.L, fn_name)(
        ctx.h, n, _lib.raw(jd0.data_ptr()), xyz_p, 10, ka.Frames.Ecliptic,
        _lib.raw(jf.data_ptr()), 1, 0, None, 1e-6, -1, out_p,
        _lib.raw(st.data_ptr(), C.c_int32), _lib.raw(iid.data_ptr(), C.c_int64),
        _lib.raw(ijd.data_ptr()), _lib.raw(cnt.data_ptr(), _lib.KhCounters), 1)
    _lib.check(rc)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st.cpu().numpy(), iid.cpu().numpy(), cnt.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_matches_the_allocating_entry_point(ctx):
    # uneven arcs of both signs (the cost order matters), then a smaller batch
    n1, n2 = 5000, 1300
    xyz = syn.n_body_workload(n=n1, neo_fraction=0.2, seed_index=3)
    span = syn.uneven_spans(n1) / 8.0
    jdf = JD0 + span
    print(f"noalloc: n={n1} then n={n2}, uneven spans", flush=True)
    a1 = _call(ctx, "kh_propagate_n_body_dev_noalloc", n1, xyz, jdf)
    a2 = _call(ctx, "kh_propagate_n_body_dev_noalloc", n2, xyz[:, -n2:], jdf[-n2:])
    b1 = _call(ctx, "kh_propagate_n_body_dev", n1, xyz, jdf)
    b2 = _call(ctx, "kh_propagate_n_body_dev", n2, xyz[:, -n2:], jdf[-n2:])
    for a, b in ((a1, b1), (a2, b2)):
        assert np.all(a[1] == 0)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
