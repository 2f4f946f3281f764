"""The rounding yardstick of the long-arc parity tests: the oracle built with
FMA contraction (-mfma -ffp-contract=fast) is the same algorithm with a
different, equally valid rounding.  It must agree with the default build to
the integrator's error level on one-year arcs (the 1e-9 AU bar) and differ
from it somewhere (else the yardstick would measure nothing), with
identical statuses."""
import numpy as np

import oracle
from kete_amd import synthetic as syn


def test_fma_build_is_a_second_rounding(spk_path):
    xyz = syn.n_body_workload(n=256, neo_fraction=0.25, seed_index=2)
    jf = syn.EPOCH_JD + 365.25
    assert oracle.spk_load(spk_path) == 0
    r = oracle.batch_n_body(syn.EPOCH_JD, xyz, jf)
    with oracle.use_library(oracle.FMA_LIB_PATH):
        assert oracle.spk_load(spk_path) == 0
        f = oracle.batch_n_body(syn.EPOCH_JD, xyz, jf)
        st, p, v = oracle.analytic_2_body(100.0, xyz[:3, 0], xyz[3:, 0])
    assert np.array_equal(r["status"], f["status"]) and np.all(r["status"] == 0)
    d = np.linalg.norm(f["out"][:3] - r["out"][:3], axis=0)
    assert d.max() < 1e-9
    assert np.count_nonzero(d) > 0.5 * len(d)
    # the two-body solver too agrees to its bar
    st2, p2, v2 = oracle.analytic_2_body(100.0, xyz[:3, 0], xyz[3:, 0])
    assert st == st2 == 0 and np.abs(p - p2).max() < 1e-10
