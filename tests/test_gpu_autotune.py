"""The launch-shape autotuner (kh_autotune_n_body): it measures shapes on the
caller's batch, reports the fastest, and mode -1 calls then use it; the
results of a call do not depend on the shape chosen beyond rounding (the
three-lane kernel, and the sixteen-lane rows that finish the objects the
plain launch hands off, sum the force terms in another order; which
objects are handed off depends on the shape)."""
import numpy as np
import pytest

import kete_amd as ka
from kete_amd import synthetic as syn

pytestmark = pytest.mark.gpu

JD0 = syn.EPOCH_JD


@pytest.fixture(scope="module")
def batch():
    xyz = syn.n_body_workload(n=20000, neo_fraction=0.1, seed_index=12)
    return ka.StateArray(np.full(xyz.shape[1], JD0), xyz)


def test_autotune_reports_and_applies(ctx, batch):
    jf = JD0 + 120.0
    try:
        res = ctx.autotune(batch, jf, sample=0, apply=True)
        assert res["n_tried"] >= 5 and res["n_sample"] == len(batch)
        assert res["mode"] in (0, 1, 2, 3) and res["block"] in (64, 128, 256)
        assert 0 <= res["window"] <= 16
        assert 0 < res["ms"] <= res["default_ms"]
        tuned = ctx.n_body(batch, jf, mode=-1)
        _, _, used_mode = ctx.last_launch()
        assert used_mode == res["mode"]
        explicit = ctx.n_body(batch, jf, mode=res["mode"])
        assert np.array_equal(tuned["status"], explicit["status"])
        assert np.abs(tuned["out"][:3] - explicit["out"][:3]).max() < 1e-10
    finally:
        ctx.autotune_clear()
    ctx.n_body(batch, jf, mode=-1)
    assert ctx.last_launch()[2] == 0


def test_autotune_sampled_and_not_applied(ctx, batch):
    res = ctx.autotune(batch, JD0 - 60.0, sample=4096, apply=False)
    assert res["n_sample"] == 4096
    ctx.n_body(batch, JD0 - 60.0, mode=-1)
    assert ctx.last_launch()[2] == 0


def test_autotune_with_asteroids(ctx, batch):
    small = ka.StateArray(batch.jd[:3000], batch.xyz[:, :3000].copy())
    res = ctx.autotune(small, JD0 + 30.0, include_asteroids=True, apply=False)
    assert res["n_tried"] >= 5 and res["ms"] > 0
