"""Benchmark: Radau-15 N-body propagation on one MI355X (object-years/s).

Default workload (the headline): 1e6 massless objects (9e5 main-belt +
1e5 NEO-like, seeded element draws), one-year arc from JD 2460000.5 TDB,
perturbed by the Sun, eight planet barycentres and the Moon with solar and
Jovian GR+J2 and Earth J2, read from a synthetic SPK Type 2 ephemeris
(degree-13 Chebyshev, 32-day records); Radau EPSILON 1e-6 (kete's constant);
FP64 throughout.

A step is one full propagation of the batch.  Inputs are resident in HBM
(torch tensors) before timing; the timed region is K steps between device
synchronisations.  The dominant kernel's duration is also measured with HIP
events on the library's stream.

Other workloads (--config): 1 two-body (1e5 states x 64 epochs, solves/s),
3 uneven spans (2e6 objects, |span| log-uniform in [1, 3650] d; also timed
through the persistent work queue), 4 field-of-view checks (1e6 states x 3000 fields, pairs/s),
5 orbit-fit gradients (4096 NEO-like orbits x 60 observations, gradients/s,
against finite differences of the plain propagators), 6 MOID of 1e6 objects
against the Earth (pairs/s), 7 propagate_n_body_long (1e5 objects in batches
of 10 with the nine planets, one year, object-years/s).

The default run (config 2) also carries the other workloads as sub-records
(--no-sub skips them; profiling runs of the headline kernel use it).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--config C] [--no-sub]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# flop model per unit of work, read off the reference's arithmetic:
F_FORCE = 1848.0   # force evaluation: 11 Chebyshev lookups + Newton + GR + J2
F_SWEEP = 1129.0   # predictor 7 substeps x 139 + B = G.C 147 + convergence 9
F_ACCEPT = 111.0   # state update 107 + step-size 4
F_FORCE_AST = 820.0  # include_asteroids: 5 x (Type 21 record 114 + Newton 29) + Sun 105
# per parameter column of the partials kernel (derivative of the same
# computation along one parameter): Jacobian-vector products of the force
# terms (Newton 10 x 19, GR 2 x 60, J2 3 x 50, non-grav 100), the sweep's
# linear predictor/corrector algebra, the accepted-step update
F_TAN_FORCE = 560.0
F_TAN_SWEEP = 1120.0
F_TAN_ACCEPT = 107.0
F_KEPLER0, F_KEPLER_IT = 97.0, 37.0   # universal Kepler: 97 + 37 per Newton iteration
# moid per pair: the 100 x 100 grid scan (3 sub, 3 mul, 2 add per entry) and
# the two element sets; the Kepler solves are counted by the kernels
F_MOID_SCAN, F_MOID_ELEM = 8.0e4, 120.0
# propagate_n_body_long per force evaluation of a batch: 24 flop per Newton
# pair term, 76 more for the GR and J2 terms of the Sun and Jupiter
F_LONG_PAIR, F_LONG_GRJ2 = 24.0, 76.0
# FOV prefilter per (state, field) pair: relative velocity and its norm, the
# light-time shift, the linear position, four edge dot products and the
# containment distance (52 add/mul/fma + 3 square roots)
F_PREFILTER = 55.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads of the CPU baseline: every CPU this process may run on."""
    return len(os.sched_getaffinity(0))


def host_cpu():
    """What the CPU baseline ran on: threads used (the affinity set),
    os.cpu_count(), the CPU model and the cgroup CPU quota (CPUs), if any."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            quota = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        pass
    return {"threads": cpu_threads(), "os_cpu_count": os.cpu_count(), "cpu_model": model,
            "cgroup_cpu_quota": quota}


def cpu_record(value, unit, sample, kind="port"):
    """cpu_baseline object: the rate, on all affinity threads of this host."""
    h = host_cpu()
    return {"value": value, "unit": unit, "cores": h["threads"], "kind": kind, "sample": sample,
            "os_cpu_count": h["os_cpu_count"], "cpu_model": h["cpu_model"],
            "cgroup_cpu_quota": h["cgroup_cpu_quota"]}


def beyond_bar(oracle, load, jd0, xyz, jdf, gpu_pos, ref_pos, include_asteroids=False,
               epsilon=1e-6, bar=1e-9, limit=16):
    """Objects whose GPU and oracle positions differ by more than `bar` AU,
    each set against two yardsticks: the oracle's FMA-contracted build (a
    second rounding of the same algorithm: how far rounding alone moves this
    object) and the oracle at epsilon 1e-10 (converged to ~1e-11 AU on
    one-year arcs), with the distance of both epsilon-1e-6 results from it.
    Where kete's own result is as far from the converged answer as the
    difference, the two took different step sequences through a close
    approach (or a convergence test decided by the last bit) and neither is
    the more accurate; `gpu_closer` says which one is nearer.  load() loads
    the ephemerides into the current oracle library."""
    d = np.abs(gpu_pos - ref_pos).max(axis=0)
    far = np.nonzero(d > bar)[0][:limit]
    out = []
    if len(far) == 0:
        return out
    jf = jdf if len(np.atleast_1d(jdf)) == 1 else np.atleast_1d(jdf)[far]
    old = oracle.radau_epsilon(1e-10)
    try:
        conv = oracle.batch_n_body(jd0, xyz[:, far], jf, include_asteroids=include_asteroids)
    finally:
        oracle.radau_epsilon(old)
    with oracle.use_library(oracle.FMA_LIB_PATH):
        load()
        old_f = oracle.radau_epsilon(epsilon)
        fma = oracle.batch_n_body(jd0, xyz[:, far], jf, include_asteroids=include_asteroids)
        oracle.radau_epsilon(old_f)
    load()
    for k, i in enumerate(far):
        eg = float(np.abs(gpu_pos[:, i] - conv["out"][:3, k]).max())
        er = float(np.abs(ref_pos[:, i] - conv["out"][:3, k]).max())
        ef = float(np.abs(ref_pos[:, i] - fma["out"][:3, k]).max())
        out.append({"sample_index": int(i), "diff_au": float(d[i]), "fma_build_vs_kete_au": ef,
                    "gpu_vs_converged_au": eg, "kete_vs_converged_au": er,
                    "gpu_closer": eg <= er})
    return out


def setup_nbody(n, xyz, jd0, jdf_host, stride):
    import torch
    dev = torch.device("cuda:0")
    t = dict(
        jd0=torch.full((n,), jd0, dtype=torch.float64, device=dev),
        xyz=torch.from_numpy(xyz).to(dev).contiguous(),
        jf=torch.from_numpy(np.ascontiguousarray(jdf_host, dtype=np.float64)).to(dev),
        out=torch.empty((6, n), dtype=torch.float64, device=dev),
        st=torch.empty(n, dtype=torch.int32, device=dev),
        iid=torch.empty(n, dtype=torch.int64, device=dev),
        ijd=torch.empty(n, dtype=torch.float64, device=dev),
        cnt=torch.zeros((n, 4), dtype=torch.int32, device=dev),
    )
    torch.cuda.synchronize()
    return t


def nbody_step_fn(ctx, t, n, stride, mode, include_asteroids=False, epsilon=1e-6):
    import kete_amd as ka
    from kete_amd import _lib
    xyz_p = _lib.ptr_array([t["xyz"][k].data_ptr() for k in range(6)])
    out_p = _lib.ptr_array([t["out"][k].data_ptr() for k in range(6)])

    def step(counters=False):
        rc = ctx.L.kh_propagate_n_body_dev(
            ctx.h, n, _lib.raw(t["jd0"].data_ptr()), xyz_p, 10, ka.Frames.Ecliptic,
            _lib.raw(t["jf"].data_ptr()), stride, int(include_asteroids), None, epsilon, mode, out_p,
            _lib.raw(t["st"].data_ptr(), C.c_int32), _lib.raw(t["iid"].data_ptr(), C.c_int64),
            _lib.raw(t["ijd"].data_ptr()),
            _lib.raw(t["cnt"].data_ptr(), _lib.KhCounters) if counters else None, 0)
        _lib.check(rc)
    return step


def time_steps(ctx, step, steps, warmup):
    import torch
    for _ in range(max(warmup, 0)):
        step()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        kernel_ms.append(ctx.last_kernel_ms())  # waits for this launch's end event
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1000.0 / steps, float(np.mean(kernel_ms))


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    FETCH_SIZE / WRITE_SIZE measurement (profiles/hbm_traffic.json, scaled by
    the copy-kernel calibration), or None when that workload was not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("workload") != workload:
        return None
    return float(rec["bytes_per_launch"])


def executed_roofline(kernel, workload, k_ms):
    """FP64 rate of the instructions the kernel executes (64 lanes x (2 FMA
    + MUL + ADD) per wave instruction, weighted by the active-lane fraction)
    from the committed counter pass on this workload
    (profiles/valu_counts.json) over this run's kernel time; None when that
    workload was not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "valu_counts.json")) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("workload") != workload:
        return None
    achieved = rec["executed_flop_per_launch"] / (k_ms / 1000.0) / 1e12
    return {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "flop_per_launch": rec["executed_flop_per_launch"], "kernel": kernel,
            "source": rec["source"]}


def radau_roofline(cnt, k_ms, workload, include_asteroids=False):
    n_force, n_sweep, n_acc = (int(cnt[:, k].sum()) for k in range(3))
    f_force = F_FORCE + (F_FORCE_AST if include_asteroids else 0.0)
    flop = f_force * n_force + F_SWEEP * n_sweep + F_ACCEPT * n_acc
    achieved = flop / (k_ms / 1000.0) / 1e12
    return {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic("radau_nbody_kernel", workload),
            "kernel": "radau_nbody_kernel", "kernel_ms": k_ms, "flop_per_launch": flop,
            "work": {"force_evals": n_force, "sweeps": n_sweep, "accepted_steps": n_acc}}


def run_nbody(args, uneven=False, epsilon=1e-6, extras=True):
    """Config 2 (uneven=False) or config 3; epsilon: the Radau tolerance
    (kete's 1e-6; 1e-9 for the labelled run); extras: the host-array and
    include_asteroids legs of config 2."""
    import torch

    import kete_amd as ka
    from kete_amd import synthetic as syn
    n = args.n or (2_000_000 if uneven else 1_000_000)
    seed_index = 3 if uneven else 2
    log(f"[bench] generating workload n={n} uneven={uneven}")
    xyz = syn.n_body_workload(n=n, neo_fraction=0.1, seed_index=seed_index)
    jd0 = syn.EPOCH_JD
    if uneven:
        span = syn.uneven_spans(n)
        jdf = jd0 + span
        stride = 1
    else:
        jdf = np.array([jd0 + 365.25])
        stride = 0
    ctx = ka.get_context()
    t = setup_nbody(n, xyz, jd0, jdf, stride)
    mode = args.mode
    step = nbody_step_fn(ctx, t, n, stride, mode, epsilon=epsilon)
    log("[bench] warmup (with work counters)")
    step(counters=True)
    torch.cuda.synchronize()
    cnt = t["cnt"].cpu().numpy().astype(np.int64)
    status = t["st"].cpu().numpy()
    log(f"[bench] timing {args.steps} steps")
    ms_per_step, k_ms = time_steps(ctx, step, args.steps, max(args.warmup - 1, 0))
    block, grid, used_mode = ctx.last_launch()
    spans = np.abs(np.broadcast_to(jdf, (n,)) - jd0)
    obj_years = float(spans.sum() / 365.25)
    value = obj_years / (ms_per_step / 1000.0)
    workload = "radau15_nbody_uneven_spans_2e6" if uneven else "radau15_nbody_1e6_objects_1yr"
    if args.n:
        workload += f"_n{n}"
    if epsilon != 1e-6:
        workload += f"_eps{epsilon:g}"
    roof = radau_roofline(cnt, k_ms, workload)
    roof_x = executed_roofline("radau_nbody_kernel", workload, k_ms)
    log(f"[bench] {ms_per_step:.1f} ms/step, kernel {k_ms:.1f} ms, {value:.3e} obj-yr/s, "
        f"{roof['achieved']:.2f} TFLOP/s FP64 ({100 * roof['frac']:.1f}% of peak)")
    res_gpu = t["out"].cpu().numpy()
    extra = {}
    if uneven:
        # the same batch through the persistent work queue, for comparison
        pers = nbody_step_fn(ctx, t, n, stride, 1, epsilon=epsilon)
        ms_p, k_p = time_steps(ctx, pers, 1, 0)
        extra["persistent_queue_launch"] = {"ms_per_step": ms_p, "kernel_ms": k_p,
                                            "value": obj_years / (ms_p / 1000.0)}
        log(f"[bench] persistent queue launch: {ms_p:.1f} ms/step")

    if not uneven and extras:
        # the host-array entry point (inputs and outputs cross PCIe), once
        batch = ka.StateArray(np.full(n, jd0), xyz)
        t0 = time.perf_counter()
        ctx.n_body(batch, jdf[0], mode=mode)
        ms_host = (time.perf_counter() - t0) * 1000.0
        extra["host_arrays"] = {"ms_per_step": ms_host,
                                "value": obj_years / (ms_host / 1000.0),
                                "note": "kh_propagate_n_body with host arrays: staging "
                                        "copies over PCIe included"}
        log(f"[bench] host-array entry point: {ms_host:.1f} ms/step")
        # the same batch with include_asteroids (kete's MASSES: five Type 21
        # perturbers after the planets)
        ast_step = nbody_step_fn(ctx, t, n, stride, mode, include_asteroids=True)
        ast_step(counters=True)
        torch.cuda.synchronize()
        cnt_a = t["cnt"].cpu().numpy().astype(np.int64)
        st_a = t["st"].cpu().numpy()
        out_a = t["out"].cpu().numpy()
        ms_a, k_a = time_steps(ctx, ast_step, max(1, args.steps // 2), 0)
        roof_a = radau_roofline(cnt_a, k_a, workload + "_masses", include_asteroids=True)
        roof_a.pop("traffic")
        extra["include_asteroids"] = {
            "ms_per_step": ms_a, "kernel_ms": k_a, "value": obj_years / (ms_a / 1000.0),
            "cost_vs_planets": ms_a / ms_per_step, "roofline": roof_a,
            "status_counts": {str(k): int(v) for k, v in zip(*np.unique(st_a, return_counts=True))},
            "note": "same workload, force list MASSES (PLANETS + Ceres, Pallas, Vesta, "
                    "Hygiea, Interamnia from synthetic SPK Type 21 segments)"}
        log(f"[bench] include_asteroids: {ms_a:.1f} ms/step ({ms_a / ms_per_step:.2f}x), "
            f"{100 * roof_a['frac']:.1f}% of FP64 peak")

    cpu = parity = None
    if not args.no_cpu:
        import oracle
        oracle.spk_load(syn.ephemeris_path())
        threads = cpu_threads()
        eps_old = oracle.radau_epsilon(epsilon)
        try:
            # grow the evenly spaced sample until one run takes about
            # cpu_seconds of wall time on all threads
            m = 256
            while True:
                idx = np.linspace(0, n - 1, m).astype(np.int64)
                tc = time.perf_counter()
                jf_s = jdf[idx] if stride else jdf
                r = oracle.batch_n_body(jd0, xyz[:, idx], jf_s, n_threads=threads)
                dt = time.perf_counter() - tc
                if dt > 0.6 * args.cpu_seconds or m >= n:
                    break
                m = min(n, int(m * min(8.0, args.cpu_seconds / max(dt, 1e-3))))
        finally:
            oracle.radau_epsilon(eps_old)
        cpu_val = float(spans[idx].sum() / 365.25) / dt
        ok = (r["status"] == 0) & (status[idx] == 0)
        dpos = np.abs(res_gpu[:3, idx] - r["out"][:3]).max(axis=0)
        dp = float(dpos[ok].max()) if ok.any() else float("nan")
        parity = {"sample": int(len(idx)), "max_pos_err_au": dp,
                  "status_match": bool(np.array_equal(r["status"], status[idx]))}
        if not uneven:
            parity["beyond_1e-9"] = beyond_bar(
                oracle, lambda: oracle.spk_load(syn.ephemeris_path()), jd0, xyz[:, idx][:, ok],
                jdf, res_gpu[:3, idx][:, ok], r["out"][:3, ok], epsilon=epsilon)
        else:
            # arcs up to ten years: the objects past 1e-8 AU, with the same
            # yardsticks (the uneven-span test bounds the distribution)
            parity["beyond_1e-8"] = beyond_bar(
                oracle, lambda: oracle.spk_load(syn.ephemeris_path()), jd0, xyz[:, idx][:, ok],
                jdf[idx][ok], res_gpu[:3, idx][:, ok], r["out"][:3, ok], epsilon=epsilon,
                bar=1e-8)
        if uneven:
            # per decade of |span|: the error grows with the arc (close
            # approaches over 10-year arcs); the bound is derived in the
            # uneven-span parity test
            sp = spans[idx]
            parity["max_pos_err_au_by_span"] = {
                f"{lo:g}-{hi:g}d": float(dpos[ok & (sp >= lo) & (sp < hi)].max())
                for lo, hi in ((1, 10), (10, 100), (100, 1000), (1000, 3651))
                if np.any(ok & (sp >= lo) & (sp < hi))}
        cpu = cpu_record(cpu_val, "object-years/s",
                         f"{len(idx)} evenly spaced objects of the same workload and arcs, "
                         f"{dt:.1f} s wall on {threads} threads (C restatement, -O3)")
        log(f"[bench] cpu {cpu_val:.1f} obj-yr/s on {threads} threads; parity {parity}")
        if "include_asteroids" in extra:
            # include_asteroids on a quarter of the sample: CPU rate and parity
            ia = idx[::4]
            oracle.spk_load(syn.ephemeris_path(), syn.asteroid_ephemeris_path())
            tc = time.perf_counter()
            ra = oracle.batch_n_body(jd0, xyz[:, ia], jdf, n_threads=threads,
                                     include_asteroids=True)
            dta = time.perf_counter() - tc
            oka = (ra["status"] == 0) & (st_a[ia] == 0)
            ea = extra["include_asteroids"]
            ga, rfa = out_a[:3, ia][:, oka], ra["out"][:3, oka]
            ea["parity"] = {"sample": int(len(ia)),
                            "max_pos_err_au": float(np.abs(ga - rfa).max()),
                            "status_match": bool(np.array_equal(ra["status"], st_a[ia])),
                            "beyond_1e-9": beyond_bar(
                                oracle, lambda: oracle.spk_load(syn.ephemeris_path(),
                                                                syn.asteroid_ephemeris_path()),
                                jd0, xyz[:, ia][:, oka], jdf, ga, rfa, include_asteroids=True)}
            ea["cpu_baseline"] = cpu_record(float(spans[ia].sum() / 365.25) / dta,
                                            "object-years/s", f"{len(ia)} objects, {dta:.1f} s wall")
            log(f"[bench] include_asteroids cpu {ea['cpu_baseline']['value']:.1f} obj-yr/s; "
                f"parity {ea['parity']}")
            oracle.spk_load(syn.ephemeris_path())

    line = {
        "metric": "object-years/s, FP64 Radau (<=1e-9 AU vs ref)",
        "value": value,
        "unit": "object-years/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "objects": n, "main_belt": n - int(round(n * 0.1)),
                   "neo": int(round(n * 0.1)),
                   "arc_days": "log-uniform |span| in [1, 3650], random sign" if uneven
                   else 365.25,
                   "object_years": obj_years, "epsilon": epsilon,
                   "ephemeris": "synthetic SPK type 2, 11 segments, degree 13, 32 d",
                   "force": "Sun+8 planets+Moon, GR+J2 (Sun, Jupiter), J2 (Earth)",
                   "seed": syn.SEED_BASE + seed_index,
                   "launch": {"block": block, "grid": grid, "mode": used_mode}},
        "roofline": roof,
        "roofline_executed": roof_x,
        "cpu_baseline": cpu,
        "parity": parity,
        "status_counts": {str(k): int(v) for k, v in zip(*np.unique(status, return_counts=True))},
    }
    line.update(extra)
    return line


def run_adjoint(args):
    """Gradients of an orbit-fit chi^2 (4096 NEO-like orbits, 60 observations
    each over 2 years) with respect to the 6 state and 3 non-grav parameters,
    through torch.autograd and the partials kernel; compared with central
    finite differences of the plain propagators (13 evaluations: state only;
    19: all 9 parameters), batched into one call and as separate calls."""
    import torch

    import kete_amd as ka
    from kete_amd import astrometry as ast
    from kete_amd import synthetic as syn
    n = args.n or 4096
    m = 60
    ctx = ka.get_context()
    log(f"[bench] adjoint workload n={n} m={m}")
    w = ast.synthetic_observations(n=n, n_obs=m, years=2.0, noise_arcsec=0.1, seed_index=5,
                                   ctx=ctx)
    p0 = w["params_true"]
    args_chi = (w["jd0"], w["jd_obs"], w["obs"], w["ra_obs"], w["dec_obs"], w["sigma"])

    def grad_step():
        p = p0.clone().requires_grad_(True)
        f = ast.chi2(p, *args_chi)
        f.backward()
        return p.grad

    for _ in range(max(args.warmup, 1)):
        g_tan = grad_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k_ms = []
    for _ in range(args.steps):
        g_tan = grad_step()
        k_ms.append(ctx.last_kernel_ms())
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    k_ms = float(np.mean(k_ms))
    # the same gradient with a new integrator at every epoch (successive
    # propagate_n_body calls), for comparison
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p = p0.clone().requires_grad_(True)
    ast.chi2(p, *args_chi, restart=True).backward()
    torch.cuda.synchronize()
    ms_restart = (time.perf_counter() - t0) * 1000.0
    cnt = torch.zeros((n, 4), dtype=torch.int32, device="cuda:0")
    _, _, status = ast.radec_partials_dev(p0, w["jd0"], w["jd_obs"], w["obs"], counters=cnt)
    cnt = cnt.cpu().numpy().astype(np.int64)
    n_bad = int((status != 0).sum())
    log(f"[bench] autograd gradient: {ms:.1f} ms/step (kernel {k_ms:.1f} ms), "
        f"{n / (ms / 1000.0):.0f} gradients/s, {n_bad} failed objects")

    # finite differences of the plain propagators
    jd_obs = w["jd_obs"][:, 0].cpu().numpy()
    obs = w["obs"][:, :, 0].cpu().numpy()
    ra_obs, dec_obs, sigma = w["ra_obs"], w["dec_obs"], w["sigma"]
    steps_fd = torch.tensor([1e-6] * 3 + [1e-7] * 3 + [1e-9] * 3, dtype=torch.float64,
                            device="cuda:0")

    def wrap(x):
        return torch.remainder(x + math.pi, 2 * math.pi) - math.pi

    def fd_gradient(n_par, batched):
        """Central-difference Jacobian of the predicted RA/Dec (2 n_par + 1
        evaluations of the plain chain), then the chi^2 gradient
        2 sum r J / sigma^2 from it, as a finite-difference orbit fit does."""
        copies = [p0]
        for k in range(n_par):
            for sgn in (1.0, -1.0):
                q = p0.clone()
                q[:, k] += sgn * steps_fd[k]
                copies.append(q)
        if batched:
            P = torch.cat(copies, dim=0)
            radec, _ = ast.radec_plain(P, w["jd0"].repeat(len(copies)), jd_obs, obs, ctx=ctx)
            rads = [radec[i * n:(i + 1) * n] for i in range(len(copies))]
        else:
            rads = [ast.radec_plain(q, w["jd0"], jd_obs, obs, ctx=ctx)[0] for q in copies]
        J = []
        for i in range(n_par):
            d = rads[1 + 2 * i] - rads[2 + 2 * i]
            d = torch.stack([wrap(d[..., 0]), d[..., 1]], dim=-1)
            J.append(d / (2 * steps_fd[i]))
        J = torch.stack(J)                                  # [n_par, n, m, 2]
        dra = wrap(rads[0][..., 0] - ra_obs)
        ddec = rads[0][..., 1] - dec_obs
        g = 2.0 * ((dra * torch.cos(dec_obs) ** 2)[None] * J[..., 0] +
                   ddec[None] * J[..., 1]).sum(dim=2) / sigma ** 2   # [n_par, n]
        return g.t(), J

    radec_k, jac_k, _ = ast.radec_partials_dev(p0, w["jd0"], w["jd_obs"], w["obs"])
    ok = (status == 0)
    fd = {}
    for n_par in (6, 9):
        fd_gradient(n_par, True)                # warm-up (allocations, code objects)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g_fd, J_fd = fd_gradient(n_par, True)
        torch.cuda.synchronize()
        ms_b = (time.perf_counter() - t0) * 1000.0
        t0 = time.perf_counter()
        fd_gradient(n_par, False)
        torch.cuda.synchronize()
        ms_s = (time.perf_counter() - t0) * 1000.0
        scale = g_tan[:, :n_par].abs().amax(dim=1).clamp_min(1e-300)
        rel = ((g_fd - g_tan[:, :n_par]).abs().amax(dim=1) / scale)[ok]
        # Jacobian agreement, per object relative to its largest entry per column
        jk = jac_k[:, :n_par].permute(1, 3, 2, 0)           # [n_par, n, m, 2]
        jscale = jk.abs().amax(dim=(2, 3), keepdim=True).clamp_min(1e-300)
        jrel = ((J_fd - jk).abs() / jscale).amax(dim=(0, 2, 3))[ok]
        fd[f"fd{2 * n_par + 1}"] = {
            "evaluations": 2 * n_par + 1, "params": n_par,
            "steps": [float(x) for x in steps_fd[:n_par]],
            "batched_ms": ms_b, "sequential_ms": ms_s,
            "batched_gradients_per_s": n / (ms_b / 1000.0),
            "sequential_gradients_per_s": n / (ms_s / 1000.0),
            "speedup_vs_batched": ms_b / ms, "speedup_vs_sequential": ms_s / ms,
            "grad_rel_diff": {"median": float(rel.median()),
                              "p90": float(rel.quantile(0.9)), "max": float(rel.max())},
            "jacobian_rel_diff": {"median": float(jrel.median()),
                                  "p90": float(jrel.quantile(0.9)),
                                  "max": float(jrel.max())}}
        log(f"[bench] FD-{2 * n_par + 1}: batched {ms_b:.1f} ms, sequential {ms_s:.1f} ms; "
            f"jacobian rel. diff median {float(jrel.median()):.2e} p90 "
            f"{float(jrel.quantile(0.9)):.2e}; gradient rel. diff median "
            f"{float(rel.median()):.2e} p90 {float(rel.quantile(0.9)):.2e}")

    frozen = cpu = None
    if not args.no_cpu:
        cpu = cpu_fd_gradients(w, args.cpu_seconds)
        log(f"[bench] cpu FD-13 gradients: {cpu['value']:.0f} gradients/s ({cpu['sample']})")
        frozen = frozen_fd_check(w, radec_k, jac_k, g_tan, status, n_sample=64)
        log(f"[bench] partials vs frozen-step FD of the oracle chain ({frozen['sample']} "
            f"orbits): jacobian median {frozen['jacobian_rel_diff']['median']:.2e} max "
            f"{frozen['jacobian_rel_diff']['max']:.2e}; gradient max "
            f"{frozen['grad_rel_diff']['max']:.2e}")

    n_force, n_sweep, n_acc = (int(cnt[:, k].sum()) for k in range(3))
    flop = (F_FORCE * n_force + F_SWEEP * n_sweep + F_ACCEPT * n_acc +
            9 * (F_TAN_FORCE * n_force + F_TAN_SWEEP * n_sweep + F_TAN_ACCEPT * n_acc))
    achieved = flop / (k_ms / 1000.0) / 1e12
    return {"metric": "adjoint gradients/s (chi^2 over 6 state + 3 non-grav params)",
            "value": n / (ms / 1000.0), "unit": "gradients/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "orbit_fit_gradients_4096_neo_x_60_obs_2yr", "objects": n,
                       "observations": m, "years": 2.0, "noise_arcsec": 0.1,
                       "nongrav": "A1..A3 ~ N(0, 1e-12) AU/d^2, new_asteroid defaults",
                       "seed": syn.SEED_BASE + 5,
                       "method": "forward-mode partials in the propagation kernel, "
                                 "torch.autograd.Function",
                       "chain": "one integration through the 60 epochs"},
            "restart_per_epoch_ms": ms_restart,
            "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                         "traffic": None,
                         "kernel": "radau_partials_wide_kernel+partials_observe_kernel",
                         "kernel_ms": k_ms, "flop_per_launch": flop,
                         "work": {"force_evals": n_force, "sweeps": n_sweep,
                                  "accepted_steps": n_acc}},
            "roofline_executed": executed_roofline(
                "radau_partials_wide_kernel+partials_observe_kernel",
                "orbit_fit_gradients_4096_neo_x_60_obs_2yr", k_ms),
            "cpu_baseline": cpu,
            "finite_differences": fd, "failed_objects": n_bad,
            "accuracy_vs_frozen_step_fd": frozen}


def min_planet_distance(oracle, jd0, xyz, nongravs, days):
    """Smallest distance of each object to Mercury, Venus, the Earth, the
    Moon and Mars, every 0.1 day along the oracle's integration."""
    jd = jd0 + np.arange(0.1, days, 0.1)
    r = oracle.chain_observe(jd0, xyz, jd, np.ones((3, len(jd))), nongravs=nongravs,
                             restart=False, n_threads=cpu_threads())
    dmin = np.full(xyz.shape[1], np.inf)
    for pid in (1, 2, 399, 301, 4):
        pp = np.array([oracle.spk_state(pid, t, 10, oracle.FRAME_ECLIPTIC)[1] for t in jd]).T
        dmin = np.minimum(dmin, np.linalg.norm(r["state"][:3] - pp[:, :, None],
                                               axis=0).min(axis=0))
    return dmin


def cpu_fd_gradients(w, cpu_seconds):
    """CPU baseline of config 5: chi^2 gradients by 13-evaluation central
    differences (state columns) of the oracle's observable chain -- what an
    orbit fit on the CPU path does -- on evenly spaced orbits, sized to
    about cpu_seconds of wall time on all threads."""
    import oracle
    from kete_amd import synthetic as syn
    oracle.spk_load(syn.ephemeris_path())
    n = w["params_true"].shape[0]
    P = w["params_true"].cpu().numpy()
    jd0 = float(w["jd0"][0])
    jd_obs = w["jd_obs"][:, 0].cpu().numpy()
    obs = w["obs"][:, :, 0].cpu().numpy()
    steps = [1e-6] * 3 + [1e-7] * 3
    th = cpu_threads()

    def run(k):
        idx = np.linspace(0, n - 1, k).astype(np.int64)
        xyz = np.ascontiguousarray(P[idx, :6].T)
        ng = [oracle.NonGrav(1, *P[i, 6:], 1.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0) for i in idx]
        tc = time.perf_counter()
        oracle.chain_observe(jd0, xyz, jd_obs, obs, nongravs=ng, restart=False, n_threads=th)
        for c in range(6):
            for sgn in (1.0, -1.0):
                x = xyz.copy()
                x[c] += sgn * steps[c]
                oracle.chain_observe(jd0, x, jd_obs, obs, nongravs=ng, restart=False,
                                     n_threads=th)
        return time.perf_counter() - tc

    k = 64
    dt = run(k)
    k = int(min(n, max(k, k * 0.8 * cpu_seconds / max(dt, 1e-3))))
    dt = run(k)
    return cpu_record(k / dt, "gradients/s",
                      f"{k} evenly spaced orbits, 13 chain evaluations each (one integration "
                      f"through the 60 epochs), {dt:.1f} s wall")


def frozen_fd_check(w, radec_k, jac_k, g_tan, status, n_sample=64):
    """Jacobian and chi^2 gradient of the partials kernel against 5-point
    central differences of the oracle's observable chain with its step
    schedule frozen (recorded on the nominal orbit, replayed on each
    perturbed one), on evenly spaced orbits of the workload that stay 0.05
    AU from the inner planets.  Relative to each orbit's largest entry per
    column (Jacobian) or per gradient.  The gradient check applies the
    frozen-step Jacobian to the kernel's own residuals, so that it measures
    the Jacobian and not the 1e-10 rad difference of the two chains' values
    (which, against 5e-7 rad residuals, moves a gradient by ~1e-3)."""
    import oracle
    from kete_amd import synthetic as syn
    oracle.spk_load(syn.ephemeris_path())
    n = w["params_true"].shape[0]
    idx = np.linspace(0, n - 1, min(n, n_sample)).astype(np.int64)
    p = w["params_true"].cpu().numpy()[idx]                  # [k, 9]
    xyz = np.ascontiguousarray(p[:, :6].T)
    a = np.ascontiguousarray(p[:, 6:].T)
    jd0 = float(w["jd0"][0])
    jd_obs = w["jd_obs"][:, 0].cpu().numpy()
    obs = w["obs"][:, :, 0].cpu().numpy()
    m, k = len(jd_obs), len(idx)

    def ngl(aa):
        return [oracle.NonGrav(1, *aa[:, i], 1.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0) for i in range(k)]

    th = cpu_threads()
    tc = time.perf_counter()
    rec = oracle.chain_observe(jd0, xyz, jd_obs, obs, nongravs=ngl(a), restart=False,
                               record=True, n_threads=th)
    steps = np.array([2e-6] * 3 + [2e-8] * 3 + [1e-10] * 3)
    wts = {-2: 1.0, -1: -8.0, 1: 8.0, 2: -1.0}
    J = np.zeros((2, 9, m, k))
    for c in range(9):
        ev = {}
        for q in wts:
            x, aa = xyz.copy(), a.copy()
            if c < 6:
                x[c] += q * steps[c]
            else:
                aa[c - 6] += q * steps[c]
            ev[q] = oracle.chain_observe(jd0, x, jd_obs, obs, nongravs=ngl(aa), restart=False,
                                         schedule=rec["schedule"], n_threads=th)["radec"]
        ra = sum(wts[q] * ((ev[q][0] - ev[-2][0] + np.pi) % (2 * np.pi) - np.pi) for q in wts)
        J[0, c] = ra / (12 * steps[c])
        J[1, c] = sum(wts[q] * ev[q][1] for q in wts) / (12 * steps[c])
    dmin = min_planet_distance(oracle, jd0, xyz, ngl(a), float(jd_obs.max() - jd0) + 1.0)
    dt = time.perf_counter() - tc
    jk = jac_k.cpu().numpy()[..., idx]                         # [2, 9, m, k]
    scale = np.abs(J).max(axis=(0, 2), keepdims=True)
    jrel = (np.abs(jk - J) / scale).max(axis=(0, 1, 2))
    # chi^2 gradient from the frozen-step Jacobian and the oracle's residuals
    ra_obs = w["ra_obs"].cpu().numpy()[idx].T                  # [m, k]
    dec_obs = w["dec_obs"].cpu().numpy()[idx].T
    sig = w["sigma"]
    sig = float(sig) if not hasattr(sig, "shape") or sig.dim() == 0 else \
        sig.cpu().numpy()[idx].T
    rk = radec_k.cpu().numpy()[..., idx]                       # [2, m, k]
    dra = (rk[0] - ra_obs + np.pi) % (2 * np.pi) - np.pi
    ddec = rk[1] - dec_obs
    g_fd = 2.0 * ((dra * np.cos(dec_obs) ** 2 / sig ** 2)[None] * J[0] +
                  (ddec / sig ** 2)[None] * J[1]).sum(axis=1)   # [9, k]
    gk = g_tan.cpu().numpy()[idx].T
    grel = np.abs(gk - g_fd).max(axis=0) / np.abs(g_fd).max(axis=0)
    ok = (dmin > 0.05) & (rec["status"] == 0) & (status.cpu().numpy()[idx] == 0)
    return {"sample": int(k), "beyond_0.05au": int(ok.sum()),
            "steps": [float(x) for x in steps],
            "jacobian_rel_diff": {"median": float(np.median(jrel[ok])),
                                  "max": float(jrel[ok].max())},
            "grad_rel_diff": {"median": float(np.median(grel[ok])), "max": float(grel[ok].max())},
            "within_0.05au_jacobian_max": float(jrel[~ok].max()) if (~ok).any() else None,
            "cpu_seconds": dt,
            "note": "reference: 5-point differences of the oracle chain (one integration "
                    "through the epochs) with the recorded step schedule replayed"}


def run_two_body(args):
    import torch

    import kete_amd as ka
    from kete_amd import _lib
    from kete_amd import synthetic as syn
    n = args.n or 100_000
    m = 64
    xyz, targets = syn.two_body_workload(n=n, m=m)
    ctx = ka.get_context()
    dev = torch.device("cuda:0")
    d_jd0 = torch.full((n,), syn.EPOCH_JD, dtype=torch.float64, device=dev)
    d_xyz = torch.from_numpy(xyz).to(dev).contiguous()
    d_t = torch.from_numpy(targets).to(dev)
    d_out = torch.empty((6, m, n), dtype=torch.float64, device=dev)
    d_st = torch.empty((m, n), dtype=torch.int32, device=dev)
    d_it = torch.zeros((m, n), dtype=torch.int32, device=dev)
    xyz_p = _lib.ptr_array([d_xyz[k].data_ptr() for k in range(6)])
    out_p = _lib.ptr_array([d_out[k].data_ptr() for k in range(6)])

    def step(iters=False):
        _lib.check(ctx.L.kh_propagate_two_body_dev(
            ctx.h, n, _lib.raw(d_jd0.data_ptr()), xyz_p, 10, ka.Frames.Ecliptic, m,
            _lib.raw(d_t.data_ptr()), None, out_p, _lib.raw(d_st.data_ptr(), C.c_int32),
            _lib.raw(d_it.data_ptr(), C.c_int32) if iters else None, 0))

    step(iters=True)                     # Newton iteration counts, untimed
    torch.cuda.synchronize()
    n_newton = int(d_it.sum().item())
    steps = max(args.steps, 20)          # a 2 ms launch: time enough of them
    ms, k_ms = time_steps(ctx, step, steps, args.warmup)
    solves = n * m
    out = d_out.cpu().numpy()
    bytes_out = solves * 48
    flop = F_KEPLER0 * solves + F_KEPLER_IT * n_newton
    achieved = flop / (k_ms / 1000.0) / 1e12
    line = {"metric": "two-body solves/s", "value": solves / (ms / 1000.0),
            "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "universal_kepler_1e5_states_x_64_epochs", "states": n,
                       "epochs": m},
            "roofline": {"bound": "fp64-valu", "achieved": achieved,
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                         "kernel": "two_body_kernel", "kernel_ms": k_ms,
                         "flop_per_launch": flop,
                         "transcendentals_per_launch": 7 * solves + 6 * n_newton,
                         "newton_iterations_per_solve": n_newton / solves,
                         "hbm_write_gbs": bytes_out / (k_ms / 1000.0) / 1e9},
            "roofline_executed": executed_roofline(
                "two_body_kernel", "universal_kepler_1e5_states_x_64_epochs", k_ms)}
    log(f"[bench] {ms:.3f} ms/step, kernel {k_ms:.3f} ms, {solves / (ms / 1000.0):.3e} "
        f"solves/s, {n_newton / solves:.2f} Newton iterations per solve, "
        f"{achieved:.2f} TFLOP/s")
    if not args.no_cpu:
        import oracle
        # the whole workload fits the CPU budget: repeated until it has run
        # for a third of cpu_seconds
        idx = np.arange(n)
        reps, dt = 0, 0.0
        while dt < args.cpu_seconds / 3 or reps == 0:
            tc = time.perf_counter()
            ref, _ = oracle.batch_two_body(syn.EPOCH_JD, xyz, targets, n_threads=cpu_threads())
            dt += time.perf_counter() - tc
            reps += 1
        line["cpu_baseline"] = cpu_record(reps * n * m / dt, "solves/s",
                                          f"the whole workload ({n} states x {m} epochs) "
                                          f"{reps} times, {dt:.1f} s wall")
        line["parity"] = {"max_pos_err_au": float(np.abs(out[:3][:, :, idx] - ref[:3]).max())}
    return line


def fov_edge_distance(nrm, obs, state):
    """Angular distance (rad) of the object's direction (state [jd, x, y, z,
    ...]) from the nearest edge plane of a rectangular field."""
    rel = np.asarray(state[1:4]) - obs[:3]
    u = rel / np.linalg.norm(rel)
    n = nrm.reshape(4, 3)
    return float(np.min(np.abs(n @ u) / np.linalg.norm(n, axis=1)))


def run_fov(args):
    """Config 4: fov_state_check of 1e6 states against 3000 rectangular
    fields (pairs/s); visible-set identity with the oracle on a sample of
    the states (all fields), and the prefilter's roofline from its HIP
    event timings."""
    import kete_amd as ka
    from kete_amd import synthetic as syn
    from kete_amd.fov import fov_state_check
    n = args.n or 1_000_000
    nf = args.fields
    xyz, fjd, obs, pts, h = syn.fov_workload(n=n, n_fields=nf)
    L = ka._lib.load()
    w = math.radians(6.9)
    nrm = np.empty((nf, 12))
    for k in range(nf):
        out = np.zeros(12)
        L.kh_fov_rect_normals(ka._lib.dptr(np.ascontiguousarray(pts[k])), 0.0, w, w,
                              ka._lib.dptr(out))
        nrm[k] = out
    batch = ka.StateArray(np.full(n, syn.EPOCH_JD), xyz)
    ctx = ka.get_context()
    fov_state_check(batch, None, fields=(fjd[:8], obs[:, :8], nrm[:8]))  # warm-up
    t0 = time.perf_counter()
    hits = fov_state_check(batch, None, fields=(fjd, obs, nrm))
    dt = time.perf_counter() - t0
    pf_ms, pf_launches, pf_pairs = ctx.last_fov_prefilter()
    # prefilter: per launch it reads each state once (epoch + 6 doubles) and
    # scans the tile's fields from LDS; F_PREFILTER flop per (state, field)
    pf_bytes = 56.0 * n * pf_launches
    pf_gbs = pf_bytes / (pf_ms / 1000.0) / 1e9
    pf_tflops = F_PREFILTER * pf_pairs / (pf_ms / 1000.0) / 1e12
    roof = {"bound": "fp64-valu", "achieved": pf_tflops, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": pf_tflops / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic("fov_prefilter_kernel",
                                        "fov_rect_6.9deg_1e6_states_x_3000_fields"),
            "kernel": "fov_prefilter_kernel", "kernel_ms": pf_ms / max(pf_launches, 1),
            "launches": pf_launches, "pairs": pf_pairs,
            "hbm": {"achieved": pf_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": pf_gbs / HBM_PEAK_GBS, "bytes_per_launch": 56.0 * n}}
    log(f"[bench] fov {dt * 1000:.1f} ms, {n * nf / dt:.3e} pairs/s, {len(hits)} hits; "
        f"prefilter {pf_launches} launches, {pf_ms:.2f} ms total, {pf_gbs:.0f} GB/s, "
        f"{pf_tflops:.2f} TFLOP/s")
    line = {"metric": "fov (state x field) pairs/s", "value": n * nf / dt, "unit": "pairs/s",
            "n_gpus": 1, "steps": 1, "ms_per_step": dt * 1000, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "fov_rect_6.9deg_1e6_states_x_3000_fields", "states": n,
                       "fields": nf, "dt_limit": 3.0,
                       "schedule": os.environ.get("KH_FOV_SCHEDULE", "continuous")},
            "roofline": roof,
            "roofline_executed": executed_roofline(
                "fov_prefilter_kernel", "fov_rect_6.9deg_1e6_states_x_3000_fields",
                pf_ms / max(pf_launches, 1)),
            "hits": len(hits)}
    if not args.no_cpu:
        # kete's schedule on the CPU restatement, a bounded sample of the
        # states against all fields; the visible sets must be identical
        # except for objects within 1e-9 rad of a field edge
        import oracle
        oracle.spk_load(syn.ephemeris_path())
        m = min(n, max(500, int(args.cpu_seconds * 250)))
        idx = np.linspace(0, n - 1, m).astype(np.int64)
        tc = time.perf_counter()
        of, oo, os_ = oracle.fov_state_check(syn.EPOCH_JD, np.ascontiguousarray(xyz[:, idx]),
                                             fjd, obs.T.copy(), nrm, dt_limit=3.0,
                                             n_threads=cpu_threads())
        dtc = time.perf_counter() - tc
        sel = np.isin(hits.obj, idx)
        g = {(int(f), int(i)): k for k, (f, i) in enumerate(zip(hits.field[sel], hits.obj[sel]))}
        gk = np.flatnonzero(sel)
        r = {(int(f), int(idx[i])): k for k, (f, i) in enumerate(zip(of, oo))}
        only = set(g) ^ set(r)
        edge = [fov_edge_distance(nrm[f], obs[:, f],
                                  os_[r[(f, i)]] if (f, i) in r else hits.state[:, gk[g[(f, i)]]])
                for f, i in only]
        common = sorted(set(g) & set(r))
        dpos = max((float(np.abs(hits.state[1:4, gk[g[k]]] - os_[r[k]][1:4]).max())
                    for k in common), default=0.0)
        line["cpu_baseline"] = cpu_record(m * nf / dtc, "pairs/s",
                                          f"{m} evenly spaced states x {nf} fields, kete's "
                                          f"schedule, {dtc:.1f} s wall")
        line["parity"] = {"sample_states": m, "hits_gpu": len(g), "hits_cpu": len(r),
                          "common": len(common), "only_one_side": len(only),
                          "only_one_side_max_edge_rad": max(edge, default=0.0),
                          "sets_identical_up_to_edge_1e-9": all(e < 1e-9 for e in edge),
                          "max_state_pos_diff_au": dpos}
        log(f"[bench] fov cpu {m * nf / dtc:.3e} pairs/s; parity {line['parity']}")
    return line


def run_moid(args):
    """kete moid of each object against the Earth (moid_py's default) for the
    config-2 population: pairs/s."""
    import torch

    import kete_amd as ka
    from kete_amd import _lib
    from kete_amd import synthetic as syn
    n = args.n or 1_000_000
    xyz = syn.n_body_workload(n=n, neo_fraction=0.1, seed_index=2)
    ctx = ka.get_context()
    dev = torch.device("cuda:0")
    d_jd = torch.full((n,), syn.EPOCH_JD, dtype=torch.float64, device=dev)
    d_xyz = torch.from_numpy(xyz).to(dev).contiguous()
    d_moid = torch.empty(n, dtype=torch.float64, device=dev)
    d_st = torch.empty(n, dtype=torch.int32, device=dev)
    d_it = torch.empty(n, dtype=torch.int32, device=dev)
    d_work = torch.zeros((2, n), dtype=torch.int32, device=dev)
    xyz_p = _lib.ptr_array([d_xyz[k].data_ptr() for k in range(6)])

    def step(work=False):
        _lib.check(ctx.L.kh_moid_dev(
            ctx.h, n, _lib.raw(d_jd.data_ptr()), xyz_p, ka.Frames.Ecliptic, None, None,
            _lib.raw(d_moid.data_ptr()), _lib.raw(d_st.data_ptr(), C.c_int32), None,
            _lib.raw(d_it.data_ptr(), C.c_int32) if work else None,
            _lib.raw(d_work.data_ptr(), C.c_int32) if work else None, 0))

    step(work=True)                      # work counters, untimed
    torch.cuda.synchronize()
    solves, newton = (int(d_work[k].sum().item()) for k in range(2))
    steps = max(args.steps, 5)
    ms, k_ms = time_steps(ctx, step, steps, args.warmup)
    moid = d_moid.cpu().numpy()
    st = d_st.cpu().numpy()
    flop = F_KEPLER0 * solves + F_KEPLER_IT * newton + (F_MOID_SCAN + F_MOID_ELEM) * n
    achieved = flop / (k_ms / 1000.0) / 1e12
    line = {"metric": "moid pairs/s", "value": n / (ms / 1000.0), "unit": "pairs/s",
            "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "moid_1e6_objects_vs_earth", "objects": n,
                       "second_orbit": "earth_at_epoch"},
            "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                         "traffic": None, "kernel": "moid_grid_kernel+moid_nm_kernel",
                         "kernel_ms": k_ms, "flop_per_launch": flop,
                         "work": {"kepler_solves": solves, "newton_iterations": newton,
                                  "nm_iterations_median": float(np.median(
                                      d_it.cpu().numpy()))}},
            "roofline_executed": executed_roofline(
                "moid_grid_kernel+moid_nm_kernel", "moid_1e6_objects_vs_earth", k_ms),
            "failed": int((st != 0).sum())}
    log(f"[bench] moid {ms:.2f} ms/step, kernels {k_ms:.2f} ms, {n / (ms / 1000.0):.3e} pairs/s, "
        f"{solves / n:.1f} solves and {newton / max(solves, 1):.2f} Newton iterations per solve, "
        f"{achieved:.2f} TFLOP/s")
    if not args.no_cpu:
        import oracle
        oracle.spk_load(syn.ephemeris_path())
        m = min(n, int(args.cpu_seconds * 2000 * cpu_threads()))
        idx = np.linspace(0, n - 1, m).astype(np.int64)
        tc = time.perf_counter()
        ref, rst, _, _ = oracle.moid_batch(syn.EPOCH_JD, np.ascontiguousarray(xyz[:, idx]),
                                           n_threads=cpu_threads())
        dtc = time.perf_counter() - tc
        ok = (rst == 0) & (st[idx] == 0)
        d = np.abs(moid[idx] - ref)[ok]
        line["cpu_baseline"] = cpu_record(m / dtc, "pairs/s",
                                          f"{m} evenly spaced objects, {dtc:.1f} s wall")
        line["parity"] = {"sample": m, "status_match": bool(np.array_equal(rst, st[idx])),
                          "max_abs_diff_au": float(d.max()),
                          "q99_abs_diff_au": float(np.quantile(d, 0.99))}
        log(f"[bench] moid cpu {m / dtc:.3e} pairs/s; parity {line['parity']}")
    return line


def run_long(args):
    """propagate_n_body_long: particles integrated in batches of kete's default
    10 with the nine SIMPLE_PLANETS, one year: object-years/s."""
    import torch

    import kete_amd as ka
    from kete_amd import _lib
    from kete_amd import synthetic as syn
    n = args.n or 100_000
    batch = 10
    span = 365.25
    xyz = syn.n_body_workload(n=n, neo_fraction=0.1, seed_index=2)
    ctx = ka.get_context()
    dev = torch.device("cuda:0")
    d_xyz = torch.from_numpy(xyz).to(dev).contiguous()
    d_out = torch.empty((6, n), dtype=torch.float64, device=dev)
    nb = (n + batch - 1) // batch
    d_st = torch.empty(nb, dtype=torch.int32, device=dev)
    d_cnt = torch.zeros((nb, 4), dtype=torch.int32, device=dev)
    xyz_p = _lib.ptr_array([d_xyz[k].data_ptr() for k in range(6)])
    out_p = _lib.ptr_array([d_out[k].data_ptr() for k in range(6)])
    jf = syn.EPOCH_JD + span

    def step(counters=False):
        _lib.check(ctx.L.kh_propagate_n_body_long_dev(
            ctx.h, n, syn.EPOCH_JD, xyz_p, 10, ka.Frames.Ecliptic, jf, None, None, batch,
            out_p, None, _lib.raw(d_st.data_ptr(), C.c_int32),
            _lib.raw(d_cnt.data_ptr(), _lib.KhCounters) if counters else None, 0))

    step(counters=True)
    torch.cuda.synchronize()
    cnt = d_cnt.cpu().numpy().astype(np.int64)
    ms, k_ms = time_steps(ctx, step, max(args.steps, 3), args.warmup)
    m_b = np.full(nb, batch, dtype=np.int64)
    m_b[-1] = n - batch * (nb - 1)
    bodies = 9 + m_b
    pairs = 72 + 9 * m_b
    gr_pairs = 2 * (8 + m_b)
    flop = float(np.sum(cnt[:, 0] * (F_LONG_PAIR * pairs + F_LONG_GRJ2 * gr_pairs)
                        + cnt[:, 1] * F_SWEEP * bodies + cnt[:, 2] * F_ACCEPT * bodies))
    achieved = flop / (k_ms / 1000.0) / 1e12
    st = d_st.cpu().numpy()
    out = d_out.cpu().numpy()
    value = n * span / 365.25 / (ms / 1000.0)
    line = {"metric": "n-body long object-years/s", "value": value, "unit": "object-years/s",
            "n_gpus": 1, "steps": max(args.steps, 3), "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "n_body_long_1e5_objects_1yr_batch10", "objects": n,
                       "span_days": span, "batch_size": batch},
            "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                         "traffic": None, "kernel": "radau_long_kernel", "kernel_ms": k_ms,
                         "flop_per_launch": flop,
                         "work": {"force_evals": int(cnt[:, 0].sum()),
                                  "sweeps": int(cnt[:, 1].sum()),
                                  "accepted_steps": int(cnt[:, 2].sum())}},
            "roofline_executed": executed_roofline(
                "radau_long_kernel", "n_body_long_1e5_objects_1yr_batch10", k_ms),
            "failed_batches": int((st != 0).sum())}
    log(f"[bench] long {ms:.2f} ms/step, kernel {k_ms:.2f} ms, {value:.3e} obj-yr/s, "
        f"{achieved:.2f} TFLOP/s")
    if not args.no_cpu:
        import oracle
        oracle.spk_load(syn.ephemeris_path())
        m = min(n, max(batch, int(args.cpu_seconds * 300 * cpu_threads()) // batch * batch))
        tc = time.perf_counter()
        rst, ref, _ = oracle.propagate_n_body_long(syn.EPOCH_JD, np.ascontiguousarray(xyz[:, :m]),
                                                   jf, batch_size=batch,
                                                   n_threads=cpu_threads())
        dtc = time.perf_counter() - tc
        line["cpu_baseline"] = cpu_record(m * span / 365.25 / dtc, "object-years/s",
                                          f"the first {m} objects, {dtc:.1f} s wall")
        line["parity"] = {"sample": m, "status": int(rst),
                          "max_pos_err_au": float(np.abs(out[:3, :m] - ref[:3]).max())}
        log(f"[bench] long cpu {m * span / 365.25 / dtc:.3e} obj-yr/s; parity {line['parity']}")
    return line


def run_copy_calibration(args):
    """A plain device copy of known size (2^27 doubles read, 2^27 written) for
    scaling the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = 1 << 27
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"metric": "copy calibration", "bytes_read": n * 8, "bytes_written": n * 8,
            "ms": dt * 1000.0, "gbs": 2 * n * 8 / dt / 1e9}


def run_sub_records(args):
    """The other workloads of the metric, after the headline, each with its
    own roofline, CPU baseline and parity: two-body (config 1), uneven spans
    (config 3), field of view (config 4), orbit-fit gradients (config 5),
    the headline at epsilon 1e-9, MOID and propagate_n_body_long.  Shorter
    timing (at most 3 steps) and CPU samples (about 6 s each)."""
    import copy
    import traceback
    sub = copy.copy(args)
    sub.steps, sub.warmup, sub.n = min(args.steps, 3), 1, 0
    sub.cpu_seconds = min(args.cpu_seconds, 6.0)
    legs = [("two_body", lambda: run_two_body(sub)),
            ("uneven_spans", lambda: run_nbody(sub, uneven=True)),
            ("fov", lambda: run_fov(sub)),
            ("orbit_fit_gradients", lambda: run_adjoint(sub)),
            ("epsilon_1e-9", lambda: run_nbody(sub, epsilon=1e-9, extras=False)),
            ("moid", lambda: run_moid(sub)),
            ("n_body_long", lambda: run_long(sub))]
    out = {}
    for name, fn in legs:
        t0 = time.perf_counter()
        log(f"[bench] sub-record {name}")
        try:
            out[name] = fn()
        except Exception as e:          # keep the headline line whatever a leg does
            traceback.print_exc()
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        out[name]["wall_s"] = time.perf_counter() - t0
        log(f"[bench] sub-record {name} done in {out[name]['wall_s']:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--fields", type=int, default=3000)
    ap.add_argument("--config", type=int, default=2, choices=[0, 1, 2, 3, 4, 5, 6, 7])
    ap.add_argument("--mode", type=int, default=-1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU time budget of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--epsilon", type=float, default=1e-6,
                    help="Radau tolerance of configs 2 and 3 (kete's 1e-6 by default)")
    ap.add_argument("--no-sub", action="store_true",
                    help="config 2 only: skip the other workloads' sub-records")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark measures one GPU; --gpus must be 1 "
                                   "(the propagation shards by object across GPUs)"}))
        sys.exit(2)
    if args.config == 2:
        line = run_nbody(args, epsilon=args.epsilon)
        if not args.no_sub:
            line["sub_records"] = run_sub_records(args)
    elif args.config == 3:
        line = run_nbody(args, uneven=True, epsilon=args.epsilon)
    elif args.config == 1:
        line = run_two_body(args)
    elif args.config == 0:
        line = run_copy_calibration(args)
    elif args.config == 5:
        line = run_adjoint(args)
    elif args.config == 6:
        line = run_moid(args)
    elif args.config == 7:
        line = run_long(args)
    else:
        line = run_fov(args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
