"""GPU parity of the observable/partials path (kh_radec_partials) and its
torch.autograd wrapper.

Checks, each against the oracle on the CPU:
  - central force: the state partials equal the two-body state transition
    matrix, taken by central differences of the analytic Kepler solution
    (kepler.rs:254-297), which has no step-size noise;
  - PLANETS force: values equal the oracle's observable chain (sequential
    propagate_n_body arcs, light-time two-body step, RA/Dec) and the partials
    equal central finite differences of that chain;
  - the autograd gradient of chi^2 equals J^T (d chi^2 / d radec).
Tolerances: values 1e-9 AU / 1e-9 rad; partials 1e-6 relative to the
column's largest entry (the bar for gradient accuracy), except where noted.
"""
import math

import numpy as np
import pytest

import kete_amd as ka
from kete_amd import synthetic as syn

pytestmark = pytest.mark.gpu

JD0 = syn.EPOCH_JD
C_INV = 0.005775518331436995   # C_AU_PER_DAY_INV (constants/universal.rs)
ROT_C, ROT_S = math.cos(0.40909280422232897), math.sin(0.40909280422232897)


def ecl_to_equ(v):
    v = np.asarray(v, dtype=np.float64)
    return np.array([v[0], ROT_C * v[1] - ROT_S * v[2], ROT_S * v[1] + ROT_C * v[2]])


def oracle_observables(oracle, jd0, xyz, jd_obs, obs, ng=None):
    """The observable chain on the oracle for objects xyz [6, n] (Sun-centred
    ecliptic at jd0): sequential propagate_n_body arcs to each epoch, then one
    two-body step back by the light time, then RA/Dec of obs -> object.
    -> radec [2, m, n], state [6, m, n] (Sun-centred ecliptic), status [n]."""
    n, m = xyz.shape[1], len(jd_obs)
    radec = np.full((2, m, n), np.nan)
    state = np.full((6, m, n), np.nan)
    status = np.zeros(n, dtype=np.int32)
    cur = xyz.copy()
    t = np.full(n, jd0)
    for j in range(m):
        r = oracle.batch_n_body(t, cur, jd_obs[j], center=10, frame=oracle.FRAME_ECLIPTIC,
                                nongravs=ng, n_threads=8)
        cur = r["out"]
        t = np.full(n, jd_obs[j])
        state[:, j] = cur
        ob = ecl_to_equ(obs[:, j])
        for i in range(n):
            if r["status"][i] != 0:
                status[i] = status[i] or r["status"][i]
                continue
            x, v = ecl_to_equ(cur[:3, i]), ecl_to_equ(cur[3:, i])
            tau = np.linalg.norm(x - ob) * C_INV
            st, p2, v2 = oracle.analytic_2_body(-tau, x, v)
            assert st == 0
            ra_dec = oracle.ra_dec_with_rates(p2, v2, ob, np.zeros(3))
            radec[:, j, i] = ra_dec[:2]
    return radec, state, status


def neo_sample(n, seed_index=7):
    return syn.n_body_workload(n=n, neo_fraction=1.0, seed_index=seed_index)


def test_central_partials_are_two_body_stm(ctx, oracle_spk):
    """force=1 (central_accel): d(state at t_j)/d(state at jd0) is the two-body
    state transition matrix."""
    oracle = oracle_spk
    xyz = neo_sample(6)
    jd_obs = JD0 + np.array([3.0, 17.0, 40.0, 95.0])
    obs = np.zeros((3, len(jd_obs)))
    obs[0] = 1.0                       # any observer; only the states are compared here
    g = ctx.radec_partials(JD0, xyz, jd_obs, obs, force=1, states=True)
    assert np.all(g["status"] == 0)
    for i in range(xyz.shape[1]):
        for j, tj in enumerate(jd_obs):
            st, p, v = oracle.analytic_2_body(tj - JD0, xyz[:3, i], xyz[3:, i])
            assert st == 0
            # values: Radau on the central force vs the Kepler solution
            assert np.abs(g["state"][:3, j, i] - p).max() < 1e-9
            phi = np.empty((6, 6))
            for k in range(6):
                h = 1e-6 if k < 3 else 1e-8
                xp, xm = xyz[:, i].copy(), xyz[:, i].copy()
                xp[k] += h
                xm[k] -= h
                _, pp, vp = oracle.analytic_2_body(tj - JD0, xp[:3], xp[3:])
                _, pm, vm = oracle.analytic_2_body(tj - JD0, xm[:3], xm[3:])
                phi[:, k] = np.concatenate([pp - pm, vp - vm]) / (2 * h)
            gpu = g["dstate"][:, :6, j, i]
            scale = np.abs(phi).max(axis=0)
            err = (np.abs(gpu - phi) / scale).max()
            assert err < 1e-6, (i, j, err)
            # no non-grav term in central_accel
            assert np.all(g["dstate"][:, 6:, j, i] == 0.0)


def test_central_partials_against_ketes_stm(ctx, oracle_spk):
    """The central-force columns against kete's own variational STM
    (compute_state_transition, state_transition.rs:52-112, restated in the
    oracle).  kete's arithmetic gives a different matrix -- its accel_grad
    is -1/2 of the true Jacobian, its trajectory runs with day-unit gravity
    in scaled time, and it returns the transpose -- so the comparison is
    made twice: with those defects switched off the columns agree to 1e-7
    of each column's largest entry; kete's own matrix differs by O(1), and
    the difference is recorded, not asserted away."""
    oracle = oracle_spk
    xyz = neo_sample(4, seed_index=11)
    jd_obs = JD0 + np.array([5.0, 30.0, 120.0])
    obs = np.zeros((3, len(jd_obs)))
    obs[0] = 1.0
    g = ctx.radec_partials(JD0, xyz, jd_obs, obs, force=1, states=True)
    assert np.all(g["status"] == 0)
    worst_fixed, worst_kete = 0.0, 0.0
    for i in range(xyz.shape[1]):
        for j, tj in enumerate(jd_obs):
            gpu = g["dstate"][:, :6, j, i]
            scale = np.abs(gpu).max(axis=0)
            # the equation is autonomous: times from 0 (at JD 2.46e6 the
            # scaled time's ulp, 7e-12, exceeds kete's 1e-12 end tolerance)
            st, _, _, m_fix, _ = oracle.compute_state_transition(
                xyz[:3, i], xyz[3:, i], 0.0, tj - JD0, 1.0,
                oracle.STM_FIX_GRAD | oracle.STM_FIX_ACCEL)
            assert st == 0
            st, _, _, m_kete, _ = oracle.compute_state_transition(xyz[:3, i], xyz[3:, i], 0.0,
                                                                  tj - JD0, 1.0, 0)
            assert st == 0
            worst_fixed = max(worst_fixed, (np.abs(gpu - m_fix.T) / scale).max())
            worst_kete = max(worst_kete, (np.abs(gpu - m_kete.T) / scale).max())
    print(f"  partials vs kete's STM: defects off {worst_fixed:.1e}, kete's arithmetic "
          f"{worst_kete:.2e} (relative to the column maximum)", flush=True)
    assert worst_fixed < 1e-7
    assert worst_kete > 1e-2


@pytest.mark.parametrize("restart", [True, False])
def test_planets_values_match_oracle_chain(ctx, oracle_spk, restart):
    """restart=True is the oracle's chain exactly (a new integrator per
    epoch); restart=False integrates through the epochs and agrees to the
    integrator's error level."""
    xyz = neo_sample(64)
    rng = np.random.default_rng(3)
    a = rng.normal(0.0, 1e-12, (3, 64))
    jd_obs = JD0 + np.sort(rng.uniform(0.0, 365.0, 8))
    earth, _ = ctx.spk_state(399, jd_obs, 10, ka.Frames.Ecliptic)
    models = [ka.NonGravModel.new_asteroid(*a[:, i]) for i in range(64)]
    import oracle
    ng_o = [oracle.NonGrav(1, *a[:, i], 1.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0) for i in range(64)]
    g = ctx.radec_partials(JD0, xyz, jd_obs, earth[:3], non_gravs=models, states=True,
                           counters=True, restart=restart)
    r_radec, r_state, r_status = oracle_observables(oracle_spk, JD0, xyz, jd_obs, earth[:3],
                                                    ng_o)
    assert np.array_equal(g["status"], r_status) and np.all(g["status"] == 0)
    assert np.abs(g["state"][:3] - r_state[:3]).max() < 1e-9
    # angular differences times the observer distance: 1e-9 AU on the sky
    rho = np.linalg.norm(r_state[:3] - earth[:3, :, None], axis=0)
    dra = np.abs((g["radec"][0] - r_radec[0] + np.pi) % (2 * np.pi) - np.pi)
    assert (dra * np.cos(r_radec[1]) * rho).max() < 2e-9
    assert (np.abs(g["radec"][1] - r_radec[1]) * rho).max() < 2e-9
    assert np.all(g["counters"][:, 0] > 0)


def test_planets_partials_match_central_differences(ctx, oracle_spk):
    """All 9 columns (state and A1..A3) of d(state)/dp and d(RA, Dec)/dp
    against central differences of the oracle's chain."""
    n = 3
    xyz = neo_sample(n, seed_index=9)
    rng = np.random.default_rng(5)
    a = rng.normal(0.0, 1e-12, (3, n))
    jd_obs = JD0 + np.array([20.0, 75.0, 160.0])
    earth, _ = ctx.spk_state(399, jd_obs, 10, ka.Frames.Ecliptic)
    import oracle

    def ng_list(aa):
        return [oracle.NonGrav(1, *aa[:, i], 1.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0)
                for i in range(aa.shape[1])]

    g = ctx.radec_partials(JD0, xyz, jd_obs, earth[:3],
                           non_gravs=[ka.NonGravModel.new_asteroid(*a[:, i]) for i in range(n)],
                           states=True)
    assert np.all(g["status"] == 0)
    # 5-point central differences, O(h^4).  The oracle's adaptive steps add
    # ~1e-12 relative noise to each evaluation, so the FD error grows as 1/h
    # below these steps (1e-6 AU gives 1e-6 relative); at them, 2-point and
    # 5-point stencils agree to 3e-8.
    steps = [1e-5] * 3 + [1e-6] * 3 + [1e-8] * 3
    for k in range(9):
        def chain(d):
            x, aa = xyz.copy(), a.copy()
            if k < 6:
                x[k] += d
            else:
                aa[k - 6] += d
            r, s_, _ = oracle_observables(oracle_spk, JD0, x, jd_obs, earth[:3], ng_list(aa))
            return r, s_
        h = steps[k]
        ev = {c: chain(c * h) for c in (-2, -1, 1, 2)}
        w = {-2: 1.0, -1: -8.0, 1: 8.0, 2: -1.0}
        fd_state = sum(w[c] * ev[c][1] for c in w) / (12 * h)
        base = ev[-2][0][0]
        ra = {c: base + (ev[c][0][0] - base + np.pi) % (2 * np.pi) - np.pi for c in w}
        fd_radec = np.stack([sum(w[c] * ra[c] for c in w),
                             sum(w[c] * ev[c][0][1] for c in w)]) / (12 * h)
        gs = g["dstate"][:, k]
        gr = g["jac"][:, k]
        for name, gpu, fd in (("state", gs, fd_state), ("radec", gr, fd_radec)):
            scale = np.abs(fd).max()
            err = np.abs(gpu - fd).max() / scale
            assert err < 1e-6, (name, k, err)


def test_autograd_chi2_gradient(ctx):
    torch = pytest.importorskip("torch")
    from kete_amd import astrometry as ast
    w = ast.synthetic_observations(n=16, n_obs=10, years=1.0, seed_index=5, ctx=ctx)
    p = w["params_true"].clone()
    p[:, 0] += 1e-7                                    # off the truth
    p.requires_grad_(True)
    f = ast.chi2(p, w["jd0"], w["jd_obs"], w["obs"], w["ra_obs"], w["dec_obs"], w["sigma"])
    f.backward()
    # the same gradient assembled by hand from the partials
    radec, jac, status = ast.radec_partials_dev(p.detach(), w["jd0"], w["jd_obs"], w["obs"])
    assert int((status != 0).sum()) == 0
    ra, dec = radec[0].t(), radec[1].t()               # [n, m]
    dra = torch.remainder(ra - w["ra_obs"] + math.pi, 2 * math.pi) - math.pi
    cd = torch.cos(w["dec_obs"])
    g_ra = 2 * dra * cd ** 2 / w["sigma"] ** 2          # d chi2 / d RA   [n, m]
    g_dec = 2 * (dec - w["dec_obs"]) / w["sigma"] ** 2  # d chi2 / d Dec  [n, m]
    manual = (jac[0] * g_ra.t()[None]).sum(dim=1) + (jac[1] * g_dec.t()[None]).sum(dim=1)
    manual = manual.t()                                  # [n, 9]
    assert torch.allclose(p.grad, manual, rtol=1e-9, atol=0.0)
    assert torch.isfinite(p.grad).all()


def test_partials_error_semantics(ctx):
    xyz = neo_sample(4)
    xyz[0, 1] = np.nan
    jd_obs = JD0 + np.array([5.0, 10.0])
    earth, _ = ctx.spk_state(399, jd_obs, 10, ka.Frames.Ecliptic)
    g = ctx.radec_partials(JD0, xyz, jd_obs, earth[:3])
    assert g["status"][1] == 2 and np.all(np.isnan(g["radec"][:, :, 1]))
    assert np.all(np.isnan(g["jac"][:, :, :, 1]))
    assert np.all(g["status"][[0, 2, 3]] == 0)
    # beyond the ephemeris: DAFLimits from the arc that leaves it on
    g = ctx.radec_partials(JD0, xyz[:, [0]], JD0 + np.array([10.0, 5000.0]), earth[:3])
    assert g["status"][0] == 3
    assert np.all(np.isfinite(g["radec"][:, 0, 0])) and np.all(np.isnan(g["radec"][:, 1, 0]))
    # dust models have no A1..A3: argument error
    with pytest.raises(ka.KeteHipError):
        ctx.radec_partials(JD0, xyz[:, [0]], jd_obs, earth[:3],
                           non_gravs=[ka.NonGravModel.new_dust(0.01)])


def planet_min_distance(oracle, jd0, xyz, nongravs, days):
    """Smallest distance of each object to Mercury, Venus, the Earth, the Moon
    and Mars, sampled every 0.1 day along the oracle's N-body integration."""
    jd = jd0 + np.arange(0.1, days, 0.1)
    r = oracle.chain_observe(jd0, xyz, jd, np.ones((3, len(jd))), nongravs=nongravs,
                             restart=False, n_threads=16)
    dmin = np.full(xyz.shape[1], np.inf)
    for pid in (1, 2, 399, 301, 4):
        st = [oracle.spk_state(pid, t, 10, oracle.FRAME_ECLIPTIC) for t in jd]
        assert all(x[0] == 0 for x in st)
        pp = np.array([x[1] for x in st]).T
        d = np.linalg.norm(r["state"][:3] - pp[:, :, None], axis=0)
        dmin = np.minimum(dmin, d.min(axis=0))
    return dmin


@pytest.mark.parametrize("restart", [True, False])
def test_config5_slice_against_frozen_step_differences(ctx, oracle_spk, restart):
    """256 orbits of the orbit-fit workload (NEO-like, A1..A3 ~ N(0, 1e-12)
    AU/d^2, 60 epochs over two years, the Earth as observer): all nine columns
    of d(RA, Dec)/dp against 5-point central differences of the oracle's
    chain with its step schedule frozen (recorded on the nominal orbit and
    replayed, so the map is smooth: differences at steps h and 2 h agree to
    5e-8 of each column's largest entry for the state columns and 4e-7 for
    A3, whose column is the smallest; steps: 2e-6 AU, 2e-8 AU/d, 1e-10
    AU/d^2, where the truncation error (O(h^4), growing through close
    approaches) and the rounding error (~1e-15 rad per evaluation over 12 h)
    cross).  The partials are the derivative of the computed map with the
    steps held fixed, so this is the reference they must match; plain
    differences of the adaptive map scatter by 1e-5 (its step decisions jump
    with the input) and, at the 1e-9 AU/d^2 steps a 2-year arc tolerates
    poorly, by up to 1e-3 near the celestial poles.  Bar: 1e-6 of each
    column's largest entry, on orbits that stay 0.05 AU from the inner
    planets and the Moon: a close approach amplifies the 1e-11 rad
    differences of the two integrators' step sequences (restart=False: one
    integration across all epochs) by up to 1e7, after which the two chains
    are different orbits.  Values: 1e-8 rad on those orbits."""
    oracle = oracle_spk
    n, m = 256, 60
    rng = np.random.default_rng(syn.SEED_BASE + 5)
    xyz = syn.n_body_workload(n=4096, neo_fraction=1.0, seed_index=5)[:, :n].copy()
    a = rng.normal(0.0, 1e-12, (4096, 3))[:n].T.copy()
    jd_obs = np.sort(JD0 + rng.uniform(0.0, 2.0 * 365.25, m))
    earth, _ = ctx.spk_state(399, jd_obs, 10, ka.Frames.Ecliptic)
    obs = earth[:3]
    print(f"config-5 slice: {n} orbits x {m} epochs, restart={restart}", flush=True)
    g = ctx.radec_partials(JD0, xyz, jd_obs, obs, restart=restart,
                           non_gravs=[ka.NonGravModel.new_asteroid(*a[:, i]) for i in range(n)])
    assert np.all(g["status"] == 0)

    def ng_list(aa):
        return [oracle.NonGrav(1, *aa[:, i], 1.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0)
                for i in range(aa.shape[1])]

    rec = oracle.chain_observe(JD0, xyz, jd_obs, obs, nongravs=ng_list(a), restart=restart,
                               record=True, n_threads=16)
    assert np.all(rec["status"] == 0)
    dmin = planet_min_distance(oracle, JD0, xyz, ng_list(a), 2 * 365.25 + 1)
    far = dmin > 0.05
    # values: the same chain to the integrators' agreement
    dra = (g["radec"][0] - rec["radec"][0] + np.pi) % (2 * np.pi) - np.pi
    dv = np.maximum(np.abs(dra * np.cos(rec["radec"][1])),
                    np.abs(g["radec"][1] - rec["radec"][1])).max(axis=0)
    print(f"  values: beyond 0.05 AU max {dv[far].max():.1e} rad; within: "
          + " ".join(f"{d:.3f} AU {v:.1e} rad" for d, v in zip(dmin[~far], dv[~far])), flush=True)
    assert dv[far].max() < 1e-8

    base = np.array([2e-6] * 3 + [2e-8] * 3 + [1e-10] * 3)
    w = {-2: 1.0, -1: -8.0, 1: 8.0, 2: -1.0}

    def frozen_fd(scale):
        J = np.zeros((2, 9, m, n))
        for k in range(9):
            h = base[k] * scale
            ev = {}
            for c in w:
                x, aa = xyz.copy(), a.copy()
                if k < 6:
                    x[k] += c * h
                else:
                    aa[k - 6] += c * h
                r = oracle.chain_observe(JD0, x, jd_obs, obs, nongravs=ng_list(aa),
                                         restart=restart, schedule=rec["schedule"],
                                         n_threads=16)
                ev[c] = r["radec"]
            ra = sum(w[c] * ((ev[c][0] - ev[-2][0] + np.pi) % (2 * np.pi) - np.pi) for c in w)
            J[0, k] = ra / (12 * h)
            J[1, k] = sum(w[c] * ev[c][1] for c in w) / (12 * h)
        return J

    j1, j2 = frozen_fd(1.0), frozen_fd(2.0)
    scale = np.abs(j1).max(axis=(0, 2), keepdims=True)       # per orbit and column

    def rel(x):
        return (np.abs(x - j1) / scale).max(axis=(0, 1, 2))

    err, fd_err = rel(g["jac"]), rel(j2)
    print(f"  orbits beyond 0.05 AU of the inner planets: {int(far.sum())} of {n}; "
          f"partials vs frozen FD: median {np.median(err[far]):.2e} max {err[far].max():.2e}; "
          f"FD h vs 2h max {fd_err[far].max():.2e}", flush=True)
    col = (np.abs(g["jac"] - j1) / scale)[..., far].max(axis=(0, 2, 3))
    print("  per column max: " + " ".join(f"{c:.1e}" for c in col), flush=True)
    if (~far).any():
        print(f"  within 0.05 AU: max {err[~far].max():.2e}", flush=True)
    assert far.sum() > 0.8 * n
    assert fd_err[far].max() < 5e-7
    assert err[far].max() < 1e-6


def test_device_models_checked_per_object(ctx, monkeypatch):
    """The device entry point checks each object's non-grav record on the
    device: a dust model ends that object with ValueError (2) and leaves the
    others untouched."""
    torch = pytest.importorskip("torch")
    from kete_amd import astrometry as ast
    w = ast.synthetic_observations(n=8, n_obs=6, years=0.5, seed_index=5, ctx=ctx)
    good = ast.radec_partials_dev(w["params_true"], w["jd0"], w["jd_obs"], w["obs"])
    make = ast._asteroid_models

    def with_dust(a123):
        rec = make(a123)
        rec.view(torch.int64)[3, 0] = 2            # kind 2: dust
        return rec

    monkeypatch.setattr(ast, "_asteroid_models", with_dust)
    radec, jac, status = ast.radec_partials_dev(w["params_true"], w["jd0"], w["jd_obs"],
                                                w["obs"])
    st = status.cpu().numpy()
    assert st[3] == 2 and np.all(np.delete(st, 3) == 0)
    assert torch.isnan(radec[:, :, 3]).all() and torch.isnan(jac[..., 3]).all()
    keep = [i for i in range(8) if i != 3]
    assert torch.equal(radec[:, :, keep], good[0][:, :, keep])
    assert torch.equal(jac[..., keep], good[1][..., keep])
