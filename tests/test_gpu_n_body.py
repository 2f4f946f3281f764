"""GPU parity of the Radau-15 N-body path (propagate_n_body) against the
oracle, on the synthetic ephemeris.

Tolerance: 1e-9 AU in position after a 1-year arc,
1e-11 AU/day in velocity.  The integrator's adaptive decisions depend on
last-ulp differences (FMA contraction, device libm), so step counts may
differ by a step on a few objects; the tolerance covers that.
"""
import numpy as np
import pytest

import kete_amd as ka
from kete_amd import synthetic as syn

pytestmark = pytest.mark.gpu

POS_TOL = 1e-9
VEL_TOL = 1e-11
JD0 = syn.EPOCH_JD


def oracle_spk_path():
    return syn.ephemeris_path()


@pytest.fixture(scope="module")
def mix():
    xyz = syn.n_body_workload(n=1200, neo_fraction=0.25, seed_index=2)
    return xyz


def run_both(ctx, oracle_spk, xyz, jd_final, mode=-1, frame=ka.Frames.Ecliptic, center=10,
             non_gravs=None, ng_oracle=None):
    n = xyz.shape[1]
    batch = ka.StateArray(np.full(n, JD0), xyz, frame, center)
    g = ctx.n_body(batch, jd_final, non_gravs=non_gravs, mode=mode, counters=True)
    r = oracle_spk.batch_n_body(JD0, xyz, jd_final, center=center, frame=frame,
                                nongravs=ng_oracle)
    return g, r


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_parity_one_year(ctx, oracle_spk, mix, mode):
    """mode 0: plain launch (the default), 1: persistent queue, both one lane
    per object; mode 2: persistent, three lanes per object.  All hand out the
    objects longest first."""
    g, r = run_both(ctx, oracle_spk, mix, JD0 + 365.25, mode=mode)
    assert np.array_equal(g["status"], r["status"])
    assert np.all(g["status"] == 0)
    dp = np.abs(g["out"][:3] - r["out"][:3]).max(axis=0)
    dv = np.abs(g["out"][3:] - r["out"][3:]).max(axis=0)
    assert dp.max() < POS_TOL, (dp.max(), int(dp.argmax()))
    assert dv.max() < VEL_TOL, dv.max()
    # The same algorithm: total work agrees to a few percent.  Per object the
    # sweep count is decided by |dB6|/|a| < 1e-14 (radau.rs:318), a threshold
    # ~50 ulp above roundoff, so FMA contraction and device libm flip it on
    # individual steps; accepted-step counts stay close.
    gc = g["counters"].astype(np.int64)
    rc = r["counters"]
    for k in range(3):
        assert abs(gc[:, k].sum() / rc[:, k].sum() - 1) < 0.05, k
    assert np.mean(np.abs(gc[:, 2] - rc[:, 2]) <= 1) > 0.95


@pytest.mark.parametrize("modes", [(0, 1), (3, 2)])
def test_plain_and_persistent_identical(ctx, mix, modes, monkeypatch):
    """The one-lane kernels (plain with the one-lane latency variant,
    persistent; three-lane plain and persistent) give bitwise-identical
    results (the plain launch's tail hand-off, which moves objects to
    sixteen lanes, is off here).  The sixteen-lane latency variant (the
    default for uneven arcs) sums the body terms in another order: the
    objects it takes (here all of them: a batch this small fits the latency
    kernel's cap) agree to rounding."""
    n = mix.shape[1]
    batch = ka.StateArray(np.full(n, JD0), mix)
    monkeypatch.setenv("KH_PARK", "0")
    monkeypatch.setenv("KH_LATENCY", "ilp")
    a = ctx.n_body(batch, JD0 - 200.0, mode=modes[0])
    b = ctx.n_body(batch, JD0 - 200.0, mode=modes[1])
    assert np.array_equal(a["status"], b["status"])
    assert np.array_equal(a["out"], b["out"])
    if modes[0] == 0:
        monkeypatch.setenv("KH_LATENCY", "wide")
        w = ctx.n_body(batch, JD0 - 200.0, mode=0)
        assert np.array_equal(w["status"], a["status"])
        d = np.abs(w["out"][:3] - a["out"][:3]).max(axis=0)
        assert d.max() < 1e-10


def test_tail_handoff(ctx, oracle_spk, monkeypatch):
    """The plain launch's tail hand-off: once a wavefront is down to four live
    lanes, objects with more than KH_PARK_LEFT force evaluations to go (by
    their rate so far; 200 here, so that the short arcs stay on their lanes
    -- the default, 0, parks every object still running) are parked at a
    step boundary and finished on sixteen lanes each.  A wavefront of 60
    main-belt objects on 40-day arcs and four NEOs on 3-year arcs: the NEOs
    are parked after the others end.  They must match the oracle as the
    one-lane path does, agree with the one-lane path (KH_PARK=0) to rounding
    -- not bitwise: the sixteen lanes sum the body terms in another order,
    which shows the hand-off ran -- keep their work counters running across
    it, and leave the short arcs bitwise alone."""
    mb = syn.n_body_workload(n=60, neo_fraction=0.0, seed_index=11)
    neo = syn.n_body_workload(n=4, neo_fraction=1.0, seed_index=12)
    xyz = np.concatenate([mb, neo], axis=1)
    n = xyz.shape[1]
    jf = np.concatenate([np.full(60, JD0 + 40.0), np.full(4, JD0 + 3 * 365.25)])
    batch = ka.StateArray(np.full(n, JD0), xyz)
    monkeypatch.setenv("KH_LATENCY_K", "0")
    monkeypatch.setenv("KH_PARK", "1")
    monkeypatch.setenv("KH_PARK_LEFT", "200")
    g = ctx.n_body(batch, jf, counters=True)
    monkeypatch.setenv("KH_PARK", "0")
    p = ctx.n_body(batch, jf, counters=True)
    r = oracle_spk.batch_n_body(JD0, xyz, jf)
    assert np.all(g["status"] == 0) and np.all(p["status"] == 0)
    assert np.array_equal(g["status"], r["status"])
    for res in (g, p):
        assert np.abs(res["out"][:3] - r["out"][:3]).max() < POS_TOL
        assert np.abs(res["out"][3:] - r["out"][3:]).max() < VEL_TOL
    # short arcs: untouched; long arcs: the sixteen-lane sums
    assert np.array_equal(g["out"][:, :60], p["out"][:, :60])
    d = np.abs(g["out"][:3, 60:] - p["out"][:3, 60:]).max(axis=0)
    # (the oracle's FMA build differs from its default build by 1-4e-11 AU
    # on these arcs)
    assert np.all(d > 0.0) and d.max() < 2e-10, d
    # counters continue across the hand-off: the same work to within the
    # rounding-decided sweeps and step failures (the oracle's two builds
    # differ by up to 9% in force evaluations on these arcs)
    gc = g["counters"][60:].astype(np.int64)
    pc = p["counters"][60:].astype(np.int64)
    assert np.all(np.abs(gc[:, 2] - pc[:, 2]) <= 2), (gc, pc)
    assert np.all(np.abs(gc[:, 0] / pc[:, 0] - 1) < 0.15), (gc, pc)


def test_three_lane_error_semantics(ctx, oracle_spk, mix):
    xyz = mix[:, :6].copy()
    xyz[0, 1] = np.nan
    xyz[:3, 2] = [1e-3, 0.0, 0.0]
    xyz[3:, 2] = [0.0, 0.0, 0.0]
    g, r = run_both(ctx, oracle_spk, xyz, JD0 + 30.0, mode=2)
    assert np.array_equal(g["status"], r["status"])
    assert g["status"][2] == 6 and g["impact_id"][2] == 10
    good = g["status"] == 0
    assert np.abs(g["out"][:3, good] - r["out"][:3, good]).max() < POS_TOL
    g, r = run_both(ctx, oracle_spk, mix[:, :4].copy(), JD0 + 5000.0, mode=2)
    assert np.all(g["status"] == 3)


@pytest.mark.parametrize("launch", ["default", "main_kernel"])
def test_uneven_spans_parity(ctx, oracle_spk, launch, monkeypatch):
    """Config 3 at its own spans: 4096 objects, |span| log-uniform in [1, 3650] d
    with random signs, in two launches: the default one (a batch this small
    goes to the sixteen-lane latency kernel almost whole) and the main
    kernel alone (KH_LATENCY_K=0), whose wavefronts move their LDS windows
    along with their lanes and hand their tails to sixteen-lane rows that
    read the window where the wavefront left it.  The GPU and the oracle run
    the same algorithm with different roundings (FMA contraction, operation
    order), and the adaptive Radau decisions (step sizes from a noisy b6,
    the corrector's sweep count) turn last-ulp differences into
    integration-error-sized ones, which planetary encounters on the 10-year
    NEO arcs amplify.  The yardstick is therefore the oracle against itself:
    the same restatement built with -ffp-contract=fast (a second, equally
    valid rounding).  Per decade of |span| the GPU's differences from the
    default oracle must be distributed like the FMA build's: quantiles
    50/90/99 % at most 2x, plus 1e-11 AU (the corrector's convergence level:
    one sweep more or fewer moves a short arc by up to 1.1e-11 AU).  The
    factor bounds the sampling spread of those statistics, measured on 20000
    objects in 4096-object chunks (quantile ratios 0.87 - 1.66).  Each
    object's difference must stay within 5x the decade's FMA-build maximum
    (the maximum's spread: 0.41 - 4.2) or within the one-year bar scaled by
    the arc, 1e-9 AU per 365.25 d: a step-failure flip (ten sweeps without
    the corrector meeting its ~50-ulp threshold, decided by the last bit)
    puts an arc on another step sequence whose truncation error differs by
    up to ~1e-9 AU per year -- object 521 of this sample (972 d) ends 2e-11
    AU from the converged answer in kete's rounding and 1.1e-9 AU with one
    such flip in another valid rounding, while objects 974 and 3509 go the
    other way (kete 5-6e-10 AU from it, the GPU 0.7-1e-10).  Statuses must
    be identical."""
    if launch == "main_kernel":
        monkeypatch.setenv("KH_LATENCY_K", "0")
    n = 4096
    xyz = syn.n_body_workload(n=n, neo_fraction=0.1, seed_index=3)
    span = syn.uneven_spans(n)
    jf = JD0 + span
    print(f"uneven spans: {n} objects, |span| {np.abs(span).min():.1f} - "
          f"{np.abs(span).max():.1f} d", flush=True)
    g, r = run_both(ctx, oracle_spk, xyz, jf)
    import oracle
    with oracle.use_library(oracle.FMA_LIB_PATH):
        assert oracle.spk_load(oracle_spk_path()) == 0
        f = oracle.batch_n_body(JD0, xyz, jf)
    assert np.array_equal(g["status"], r["status"])
    assert np.array_equal(f["status"], r["status"])
    ok = r["status"] == 0
    assert ok.mean() > 0.99
    dg = np.linalg.norm(g["out"][:3] - r["out"][:3], axis=0)
    df = np.linalg.norm(f["out"][:3] - r["out"][:3], axis=0)
    a = np.abs(span)
    for lo, hi in ((1, 10), (10, 100), (100, 1000), (1000, 3651)):
        s = ok & (a >= lo) & (a < hi)
        qg = np.quantile(dg[s], [0.5, 0.9, 0.99])
        qf = np.quantile(df[s], [0.5, 0.9, 0.99])
        print(f"  {lo}-{hi} d ({s.sum()}): gpu q50/90/99/max "
              + " ".join(f"{x:.1e}" for x in (*qg, dg[s].max()))
              + " | fma build " + " ".join(f"{x:.1e}" for x in (*qf, df[s].max())), flush=True)
        assert np.all(qg <= 2.0 * qf + 1e-11), (lo, hi)
        lim = np.maximum(5.0 * df[s].max() + 1e-11, 1e-9 * a[s] / 365.25)
        assert np.all(dg[s] <= lim), (lo, hi, np.nonzero(s)[0][dg[s] > lim])
    # the one-year bar holds on the arcs up to a year
    assert dg[ok & (a <= 366.0)].max() < POS_TOL


def test_frames_and_centers(ctx, oracle_spk, mix):
    xyz = mix[:, :200].copy()
    # equatorial input
    c, s = np.cos(0.40909280422232897), np.sin(0.40909280422232897)
    eq = xyz.copy()
    for k in (0, 3):
        eq[k + 1] = c * xyz[k + 1] - s * xyz[k + 2]
        eq[k + 2] = s * xyz[k + 1] + c * xyz[k + 2]
    g, r = run_both(ctx, oracle_spk, eq, JD0 + 100.0, frame=ka.Frames.Equatorial)
    assert np.all(g["status"] == 0) and np.array_equal(g["status"], r["status"])
    assert np.abs(g["out"][:3] - r["out"][:3]).max() < POS_TOL
    # SSB-centred input (center 0)
    sun, _ = ctx.spk_state(10, [JD0], 0, ka.Frames.Ecliptic)
    x0 = xyz + sun[:, :1]
    g, r = run_both(ctx, oracle_spk, x0, JD0 + 100.0, center=0)
    assert np.all(g["status"] == 0) and np.array_equal(g["status"], r["status"])
    assert np.abs(g["out"][:3] - r["out"][:3]).max() < POS_TOL
    # geocentric input (center 399, two-segment chain through the EMB)
    earth, _ = ctx.spk_state(399, [JD0], 10, ka.Frames.Ecliptic)
    xg = xyz - earth[:, :1]
    g, r = run_both(ctx, oracle_spk, xg, JD0 + 50.0, center=399)
    assert np.all(g["status"] == 0) and np.array_equal(g["status"], r["status"])
    assert np.abs(g["out"][:3] - r["out"][:3]).max() < POS_TOL


def test_error_semantics(ctx, oracle_spk, mix):
    xyz = mix[:, :6].copy()
    xyz[0, 1] = np.nan                    # non-finite input -> ValueError, NaN out
    xyz[:3, 2] = [1e-3, 0.0, 0.0]         # inside the Sun -> Impact(10) at the first eval
    xyz[3:, 2] = [0.0, 0.0, 0.0]
    g, r = run_both(ctx, oracle_spk, xyz, JD0 + 30.0)
    assert np.array_equal(g["status"], r["status"])
    assert g["status"][1] == 2 and np.all(np.isnan(g["out"][:, 1]))
    assert g["status"][2] == 6 and g["impact_id"][2] == 10
    assert r["impact_id"][2] == 10
    # beyond the ephemeris -> DAFLimits
    g, r = run_both(ctx, oracle_spk, mix[:, :4].copy(), JD0 + 5000.0)
    assert np.all(g["status"] == 3) and np.array_equal(g["status"], r["status"])
    assert np.all(np.isnan(g["out"]))
    # zero-length arc returns the input (radau.rs:183-189)
    g, r = run_both(ctx, oracle_spk, mix[:, :4].copy(), JD0)
    assert np.all(g["status"] == 0)
    assert np.abs(g["out"] - mix[:, :4]).max() < 1e-15


def test_impact_during_arc(ctx, oracle_spk, capsys):
    """An object aimed at the Earth impacts on an exact evaluation."""
    earth, _ = ctx.spk_state(399, [JD0 + 20.0], 10, ka.Frames.Ecliptic)
    e0, _ = ctx.spk_state(399, [JD0], 10, ka.Frames.Ecliptic)
    # start 0.02 AU away on a line towards where the Earth will be in 20 d
    d = earth[:3, 0] - e0[:3, 0]
    start = e0[:3, 0] - 0.01 * d / np.linalg.norm(d)
    vel = (earth[:3, 0] - start) / 20.0
    xyz = np.concatenate([start, vel]).reshape(6, 1)
    g, r = run_both(ctx, oracle_spk, xyz, JD0 + 40.0)
    assert r["status"][0] == g["status"][0]
    if r["status"][0] == 6:
        assert g["impact_id"][0] == r["impact_id"][0]
        assert abs(g["impact_jd"][0] - r["impact_jd"][0]) < 1e-6
        # kete's report on stderr: the body's name and the UTC time
        from kete_amd.naming import impact_message
        ka.propagate_n_body(ka.StateArray(np.full(1, JD0), xyz, desigs=["aimed"]), JD0 + 40.0)
        err = capsys.readouterr().err
        assert impact_message("aimed", int(g["impact_id"][0]), float(g["impact_jd"][0])) in err
        assert "<-> (earth)" in err or "<-> (moon)" in err


def test_nongrav_models(ctx, oracle_spk, mix):
    import oracle
    xyz = mix[:, -100:].copy()
    rng = np.random.default_rng(11)
    models, ng_o = [], []
    for i in range(100):
        if i % 2:
            a = rng.normal(0, 1e-8, 3)
            models.append(ka.NonGravModel.new_asteroid(*a))
            ng_o.append(oracle.NonGrav(1, a[0], a[1], a[2], 1.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0))
        else:
            models.append(ka.NonGravModel.new_dust(0.01))
            ng_o.append(oracle.NonGrav(2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0.01))
    g, r = run_both(ctx, oracle_spk, xyz, JD0 + 200.0, non_gravs=models, ng_oracle=ng_o)
    assert np.all(g["status"] == 0) and np.array_equal(g["status"], r["status"])
    assert np.abs(g["out"][:3] - r["out"][:3]).max() < POS_TOL


def test_python_surface(ctx):
    states = [ka.State(f"obj{i}", JD0, [2.0 + 0.1 * i, 0.1, 0.0], [0.0, 0.011, 0.0005])
              for i in range(5)]
    out = ka.propagate_n_body(states, JD0 + 10.0)
    assert len(out) == 5 and all(s.jd == JD0 + 10.0 for s in out)
    assert out[2].desig == "obj2" and out[2].is_finite
    out2 = ka.propagate_two_body(states, JD0 + 10.0)
    assert np.linalg.norm(out2[0].pos - out[0].pos) < 1e-4
    bad = [ka.State("x", JD0, [np.nan, 0, 0], [0, 0, 0])]
    with pytest.raises(ValueError):
        ka.propagate_n_body(bad, JD0 + 1.0, suppress_errors=False)
    assert not ka.propagate_n_body(bad, JD0 + 1.0)[0].is_finite


def test_mixed_frame_list(ctx, mix):
    """A state list mixing frames and centers (kete propagates each state on
    its own): each state's result equals that of a uniform batch."""
    c, s = np.cos(0.40909280422232897), np.sin(0.40909280422232897)
    xyz = mix[:, :6].copy()
    sts = []
    for k in range(6):
        p, v = xyz[:3, k].copy(), xyz[3:, k].copy()
        f = ka.Frames.Ecliptic
        if k % 2:
            p = np.array([p[0], c * p[1] - s * p[2], s * p[1] + c * p[2]])
            v = np.array([v[0], c * v[1] - s * v[2], s * v[1] + c * v[2]])
            f = ka.Frames.Equatorial
        sts.append(ka.State(f"m{k}", JD0, p, v, f, 10))
    out = ka.propagate_n_body(sts, JD0 + 50.0)
    ref = ka.propagate_n_body(ka.StateArray(np.full(6, JD0), xyz), JD0 + 50.0)
    for k in range(6):
        assert out[k].desig == f"m{k}" and out[k].frame == sts[k].frame
        p = out[k].pos
        if k % 2:
            p = np.array([p[0], c * p[1] + s * p[2], -s * p[1] + c * p[2]])
        # equatorial input rounds differently on the way in, which the
        # adaptive steps can carry to ~1e-11 AU over 50 d
        assert np.abs(p - ref.xyz[:3, k]).max() < 1e-10
