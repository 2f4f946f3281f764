"""Benchmark: Radau-15 N-body propagation on one MI355X (object-years/s).

Workload: 1e6 massless objects (9e5 main-belt + 1e5 NEO-like, seeded element
draws), one-year arc from JD 2460000.5 TDB, perturbed by the Sun, eight planet
barycentres and the Moon with solar and Jovian GR+J2 and Earth J2, read from a
synthetic SPK Type 2 ephemeris (degree-13 Chebyshev, 32-day records); Radau
EPSILON 1e-6 (kete's constant); FP64 throughout.

A step is one full propagation of the batch.  Inputs are resident in HBM
(torch tensors) before timing; the timed region is K steps between device
synchronisations.  The dominant kernel's duration is also measured with HIP
events on the library's stream.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# flop model per unit of work, from the reference's arithmetic:
# force evaluation (PLANETS, 11 Chebyshev lookups + Newton + GR + J2)
F_FORCE = 1848.0
F_SWEEP = 1129.0   # predictor 7 substeps x 139 + B = G.C 147 + convergence 9
F_ACCEPT = 111.0   # state update 107 + step-size 4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--mode", type=int, default=-1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU time budget of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md)"}))
        sys.exit(2)

    import torch

    import kete_amd as ka
    from kete_amd import _lib
    from kete_amd import synthetic as syn

    n = args.n
    log(f"[bench] generating workload n={n}")
    xyz = syn.n_body_workload(n=n, neo_fraction=0.1, seed_index=2)
    jd0 = syn.EPOCH_JD
    jdf = jd0 + 365.25
    ctx = ka.get_context()
    L = ctx.L
    dev = torch.device("cuda:0")
    d_jd0 = torch.full((n,), jd0, dtype=torch.float64, device=dev)
    d_xyz = torch.from_numpy(xyz).to(dev).contiguous()
    d_jf = torch.tensor([jdf], dtype=torch.float64, device=dev)
    d_out = torch.empty((6, n), dtype=torch.float64, device=dev)
    d_st = torch.empty(n, dtype=torch.int32, device=dev)
    d_iid = torch.empty(n, dtype=torch.int64, device=dev)
    d_ijd = torch.empty(n, dtype=torch.float64, device=dev)
    d_cnt = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    xyz_p = _lib.ptr_array([d_xyz[k].data_ptr() for k in range(6)])
    out_p = _lib.ptr_array([d_out[k].data_ptr() for k in range(6)])

    def step(counters=False):
        rc = L.kh_propagate_n_body_dev(
            ctx.h, n, _lib.raw(d_jd0.data_ptr()), xyz_p, 10, ka.Frames.Ecliptic,
            _lib.raw(d_jf.data_ptr()), 0, None, 1e-6, args.mode, out_p,
            _lib.raw(d_st.data_ptr(), C.c_int32), _lib.raw(d_iid.data_ptr(), C.c_int64),
            _lib.raw(d_ijd.data_ptr()),
            _lib.raw(d_cnt.data_ptr(), _lib.KhCounters) if counters else None, 0)
        _lib.check(rc)

    log("[bench] warmup")
    step(counters=True)
    torch.cuda.synchronize()
    cnt = d_cnt.cpu().numpy().astype(np.int64)
    status = d_st.cpu().numpy()
    for _ in range(max(args.warmup - 1, 0)):
        step()
    torch.cuda.synchronize()

    log(f"[bench] timing {args.steps} steps")
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kernel_ms.append(ctx.last_kernel_ms())  # waits for this launch's end event
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    k_ms = float(np.mean(kernel_ms))
    block, grid, mode = ctx.last_launch()
    obj_years = n * (jdf - jd0) / 365.25
    value = obj_years / (ms_per_step / 1000.0)

    n_force = cnt[:, 0].sum()
    n_sweep = cnt[:, 1].sum()
    n_acc = cnt[:, 2].sum()
    flop = F_FORCE * n_force + F_SWEEP * n_sweep + F_ACCEPT * n_acc
    achieved = flop / (k_ms / 1000.0) / 1e12
    log(f"[bench] {ms_per_step:.1f} ms/step, kernel {k_ms:.1f} ms, {value:.3e} obj-yr/s, "
        f"{achieved:.2f} TFLOP/s FP64")

    res_gpu = d_out.cpu().numpy()
    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        oracle.spk_load(syn.ephemeris_path())
        threads = min(16, len(os.sched_getaffinity(0)))
        # a sample that keeps the main-belt / NEO mix: every k-th object
        m = 256
        while True:
            idx = np.linspace(0, n - 1, m).astype(np.int64)
            tc = time.perf_counter()
            r = oracle.batch_n_body(jd0, xyz[:, idx], jdf, n_threads=threads)
            dt = time.perf_counter() - tc
            if dt > args.cpu_seconds / 4 or m >= n:
                break
            m = min(n, int(m * max(2.0, min(8.0, args.cpu_seconds / max(dt, 1e-3)))))
        cpu_val = len(idx) * (jdf - jd0) / 365.25 / dt
        ok = (r["status"] == 0) & (status[idx] == 0)
        dp = np.abs(res_gpu[:3, idx][:, ok] - r["out"][:3, ok]).max() if ok.any() else float("nan")
        parity = {"sample": int(len(idx)), "max_pos_err_au": float(dp),
                  "status_match": bool(np.array_equal(r["status"], status[idx]))}
        cpu = {"value": cpu_val, "unit": "object-years/s", "cores": threads, "kind": "port",
               "sample": f"{len(idx)} evenly spaced objects of the same workload, same arc, "
                         f"{dt:.1f} s wall on {threads} threads (C restatement, -O3)"}
        log(f"[bench] cpu {cpu_val:.1f} obj-yr/s on {threads} threads; parity {parity}")

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
        "config": {"workload": "radau15_nbody_1e6_objects_1yr",
                   "objects": n, "main_belt": n - int(round(n * 0.1)),
                   "neo": int(round(n * 0.1)), "arc_days": 365.25, "epsilon": 1e-6,
                   "ephemeris": "synthetic SPK type 2, 11 segments, degree 13, 32 d",
                   "force": "Sun+8 planets+Moon, GR+J2 (Sun, Jupiter), J2 (Earth)",
                   "seed": syn.SEED_BASE + 2, "launch": {"block": block, "grid": grid,
                                                         "mode": mode}},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "radau_nbody_kernel", "kernel_ms": k_ms,
                     "flop_per_launch": flop,
                     "work": {"force_evals": int(n_force), "sweeps": int(n_sweep),
                              "accepted_steps": int(n_acc)}},
        "cpu_baseline": cpu,
        "parity": parity,
        "status_counts": {str(k): int(v) for k, v in zip(*np.unique(status, return_counts=True))},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
