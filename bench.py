"""Benchmark of the multigrid V-cycle on one MI355X.

Workload (one "step" = one V-cycle): 256^3 cells, periodic in x and y,
Neumann walls in z, double-sided tanh stretching beta = 1.5 in z (the grid of
the 256^3 Rayleigh-Benard configuration), RB-SOR omega = 1, V(2,2), coarsest
2^3 level solved to 1e-2 ("Solve Coarsest" defaults).  The headline V-cycle
is the FP32 inner cycle of the mixed-precision solver; the FP64 cycle of the
plain path is reported beside it.

value: algorithmic HBM bytes per V-cycle / device time per V-cycle, with the
per-level byte model  B_l = N_l s (3 (nu1 + nu2) + 2 + 1/8 + 2 + 1/8)
(3 words per smoothing sweep, residual+restriction reads x, b and writes
b_{l+1}, prolongation reads e_{l+1} and reads/writes x), s = bytes per word.

solve: time to a 1e-10 relative L-inf residual on 256^3 and 512^3 periodic
uniform grids, RHS = 64 Fourier modes with amplitudes ~ |k|^-2 and zero mean,
plain FP64 vs mixed precision, inputs already resident in HBM; the mixed
solution's relative L2 / L-inf difference from the FP64 solution (which
reproduces the FP64 CPU oracle bit for bit), modulo the mean.

roofline: the dominant kernel (one colour pass of the finest level on
colour-split rows): algorithmic 1.5 words per cell per launch (half of a
sweep's 3) / average launch time (HIP events around the finest-level sweeps).
cpu_baseline: the FP64 CPU oracle's V-cycle on the same workload.

serving: the batched request loop on the Poisson request mix (75 % 3-D
n^3, n log-uniform in [32, 256]; 25 % 2-D n^2, n in [64, 2048]; multiples of
16; Dirichlet / Neumann / periodic; 8-mode Fourier RHS; tolerance 1e-8):
solves/s with every request queued at once (saturation; free slots take the
waiting request of largest footprint first), p50/p99 latency of
a seeded Poisson arrival stream at 70 % of that rate, and the same requests
solved one at a time with plain FP64 solves (one slot, no graphs), with a
solver built per request (the plain way) and with resident hierarchies; the
server's hierarchies are resident after one warm-up pass; RHS arrays are
generated beforehand.

rbc_256_step: the 256^3 Rayleigh-Benard time step (Ra 1e7, Pr 1, z-tanh 1.5,
dScheme 2, Morinishi, RK3 with Jacobi CN, V(3,6) pressure solves of 20
cycles): device time per full step, FP64 and with mixed-precision pressure
solves.  "Solve Coarsest" is off: with it on, the reference algorithm's
pressure solve diverges on this stretched Neumann grid in the first stage.
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
DOMINANT = "k_rb_split<float,3,true,false>"
# HBM bytes per launch of the dominant kernel (one colour pass of the finest
# level) at 256^3 from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per
# dispatch), calibrated against the copy kernel (FETCH x 2.0, WRITE x 1.0):
# (2 * 50246.5 + 32768.0) KiB.  Source: profiles/sweep_hbm_traffic_256cubed.md
DOMINANT_TRAFFIC_BYTES = (2 * 50246.5 + 32768.0) * 1024


def bytes_per_vcycle(shapes, s, nu1, nu2):
    tot = 0.0
    for shp in shapes[:-1]:
        tot += float(np.prod(shp)) * s * (3 * (nu1 + nu2) + 2 + 1 / 8 + 2 + 1 / 8)
    return tot + float(np.prod(shapes[-1])) * s * 3 * (nu1 + nu2)


def rbc_grid():
    return S.grid_desc((N, N, N), periodic=(1, 1, 0), gtype=(0, 0, S.GRID_TANH_DOUBLE),
                       beta=(1.0, 1.0, 1.5), pad=1)


def rbc_rhs(seed=2):
    rng = np.random.default_rng(seed)
    b = np.zeros((N + 2,) * 3)
    c = rng.standard_normal((N, N, N))
    b[1:-1, 1:-1, 1:-1] = c - c.mean()
    return b


def fourier_rhs(n, modes=64, kmax=16, seed=3):
    """sum_m A_m cos(2 pi k_m . x + phi_m), A_m = U(0.5, 1) |k_m|^-2, zero mean.
    Each mode is Re(e_x(i) * [e_y(j) e_z(k) e^{i phi}]): the sum over modes is one
    (n x 2M) @ (2M x n^2) product of real and imaginary parts."""
    rng = np.random.default_rng(seed)
    xc = (np.arange(n) + 0.5) / n
    A = np.zeros((n, 2 * modes))
    B = np.zeros((2 * modes, n * n))
    for m in range(modes):
        while True:
            k = rng.integers(-kmax, kmax + 1, 3)
            if np.abs(k).max() >= 1:
                break
        amp = rng.uniform(0.5, 1.0) / float(k @ k)
        phi = rng.uniform(0, 2 * np.pi)
        ex = np.exp(2j * np.pi * k[0] * xc)
        ey = np.exp(2j * np.pi * k[1] * xc)
        ez = np.exp(2j * np.pi * k[2] * xc) * np.exp(1j * phi)
        yz = np.outer(ey, ez).ravel()
        A[:, 2 * m], A[:, 2 * m + 1] = amp * ex.real, -amp * ex.imag
        B[2 * m], B[2 * m + 1] = yz.real, yz.imag
    acc = (A @ B).reshape(n, n, n)
    acc -= acc.mean()
    b = np.zeros((n + 2,) * 3)
    b[1:-1, 1:-1, 1:-1] = acc
    return b


def cycle_rates(args, nu1, nu2):
    g = rbc_grid()
    b = rbc_rhs()
    out = {}
    for name, prec in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
        p = S.mg_params(vc_count=1, vc_tol=0.0, mg_tol=1e-2, pre=nu1, post=nu2, bc=S.BC_NEUMANN,
                        precision=prec, inner_cycles=1, max_refine=1)
        mg = S.Multigrid(g, p)
        mg.solve(np.zeros_like(b), b)            # puts b and a first iterate on the device
        shapes = [mg.level_shape(lv) for lv in range(mg.levels)]
        s = 8 if prec == S.PREC_FP64 else 4
        mg.bench_cycles(args.warmup)
        ms = mg.bench_cycles(args.steps) / args.steps
        reps = 20
        mg.bench_sweeps(2)
        # a sweep of the finest level is two launches of the colour-pass
        # kernel (red, black) of equal work: per launch, half the sweep's
        # 3 words per cell and half its time
        launch_ms = mg.bench_sweeps(reps) / reps / 2
        n0 = float(np.prod(shapes[0]))
        out[name] = dict(ms=ms, bv=bytes_per_vcycle(shapes, s, nu1, nu2), launch_ms=launch_ms,
                         launch_bytes=1.5 * n0 * s, shapes=shapes, s=s)
        mg.close()
    return out


def profile_sweeps(args, nu1, nu2):
    """Finest-level sweeps only: one launch per sweep, `steps` launches per
    precision, plus 10 launches of the plain copy kernel over 2^28 doubles."""
    g = rbc_grid()
    b = rbc_rhs()
    for prec in (S.PREC_MIXED, S.PREC_FP64):
        p = S.mg_params(vc_count=1, vc_tol=0.0, mg_tol=1e-2, pre=nu1, post=nu2, bc=S.BC_NEUMANN,
                        precision=prec, inner_cycles=1, max_refine=1)
        mg = S.Multigrid(g, p)
        mg.solve(np.zeros_like(b), b)
        ms = mg.bench_sweeps(args.steps) / args.steps
        print(json.dumps({"precision": "mixed" if prec else "fp64", "sweep_ms": ms}), flush=True)
        mg.close()
    ms = S.bench_stream_copy(1 << 28, 10) / 10
    print(json.dumps({"copy_ms": ms, "copy_bytes": 16 * (1 << 28)}), flush=True)


def solve_compare(nu1, nu2, tol=1e-10, n=N, reps=3, inner=4):
    """Plain FP64 and mixed-precision solves of the periodic Fourier problem to
    the same relative L-inf residual, inputs resident in HBM; the mixed
    solution's difference from the FP64 one (the FP64 path reproduces the
    FP64 CPU oracle bit for bit), modulo the mean."""
    g = S.grid_desc((n, n, n), periodic=(1, 1, 1))
    b = fourier_rhs(n)
    bd = S.DeviceArray.from_host(b)
    res, sols = {}, {}
    for name, prec in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
        p = S.mg_params(vc_count=100, vc_tol=tol, mg_tol=1e-2, pre=nu1, post=nu2, bc=S.BC_NEUMANN,
                        precision=prec, inner_cycles=inner, max_refine=40)
        mg = S.Multigrid(g, p)
        best = None
        for r in range(reps):
            xd = S.DeviceArray.from_host(np.zeros_like(b))
            st, status = mg.solve_device(xd, bd)
            if best is None or st["device_ms"] < best["device_ms"]:
                best = st
            if r == reps - 1:
                sols[name] = xd.to_host(b.shape)[1:-1, 1:-1, 1:-1]
            xd.free()
        res[name] = dict(ms=best["device_ms"], cycles=best["cycles"], refinements=best["refinements"],
                         residual=best["residual"], converged=best["converged"])
        mg.close()
    bd.free()
    d = (sols["mixed"] - sols["mixed"].mean()) - (sols["fp64"] - sols["fp64"].mean())
    ref = sols["fp64"] - sols["fp64"].mean()
    res["diff_l2_rel"] = float(np.linalg.norm(d) / np.linalg.norm(ref))
    res["diff_linf_rel"] = float(np.abs(d).max() / np.abs(ref).max())
    return res


def rbc_fields(n, pad, seed=2):
    """T = 1 - z + U(-1e-3, 1e-3) (seeded), u = 0, P = 1, Pp = 0 on the z-tanh grid."""
    sh = (n + 2 * pad,) * 3
    g = S.grid_desc(n, periodic=(1, 1, 0), gtype=(0, 0, S.GRID_TANH_DOUBLE), beta=(1, 1, 1.5), pad=pad)
    z = S.level_metrics(g, 0, 2)["x"][1:-1]          # cell centres, grid.cc:411-508
    rng = np.random.default_rng(seed)
    T = np.zeros(sh)
    c = (slice(pad, -pad),) * 3
    T[c] = 1.0 - z[None, None, :] + rng.uniform(-1e-3, 1e-3, (n, n, n))
    zero = np.zeros(sh)
    return g, dict(vx=zero, vy=zero, vz=zero, p=np.ones(sh), t=T, pp=zero)


def rbc_step(steps, warmup, n=256):
    pad = 2
    g, f = rbc_fields(n, pad)
    nu = (1.0 / 1e7) ** 0.5
    prm = S.nse_params(dscheme=2, nlscheme=3, has_T=1, nu=nu, kappa=nu, dt=1e-3, cn_tol=1e-6,
                       buoy=(0.0, 0.0, -1.0))
    out = {}
    for name, prec in (("fp64", S.PREC_FP64), ("mixed", S.PREC_MIXED)):
        # every pressure solve of this run takes all 20 cycles (its residual
        # plateaus near 1e-4), so the mixed solver may run up to 5 FP32 cycles
        # per refinement (3 + 5 + 5 + 5 + 2)
        mg = S.mg_params(vc_count=20, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2, pre=3, post=6,
                         bc=S.BC_NEUMANN, precision=prec, inner_cycles=5, max_refine=10)
        s = S.Nse(g, mg, prm)
        s.set(**f)
        for _ in range(warmup):
            r, st = s.step()
            S.check(r)
        r, ms = s.bench_steps(steps)
        S.check(r)
        r, st = s.step()
        S.check(r)
        d = s.diag()
        out[name] = dict(ms_per_step=round(ms / steps, 3), steps=steps, jacobi_iters=st["jacobi_iters"],
                         pressure_cycles=st["mg"]["cycles"], pressure_residual=st["mg"]["residual"],
                         ke=d["ke"], divmax=d["divmax"], nu_vol=d["nuvol"], nu_wall=d["nuwall"])
        s.close()
    return out


def les_step(steps, warmup, n=384, noise=0.01, seed=5):
    """Decaying Taylor-Green vortex in a periodic (2 pi)^3 box, Re = 1600
    (nu = 1/1600, unit velocity and length), seeded uniform noise of 1 % of
    the peak velocity, spiral sub-grid model, RK3/CN, FP64 stencils and the
    mixed-precision pressure solve.  The model enters once solTime > 5 dt, so
    at least six warm-up steps run before the timed ones."""
    pad = 2
    L = 2 * np.pi
    g = S.grid_desc(n, periodic=(1, 1, 1), length=(L, L, L), pad=pad)
    x = (np.arange(-pad, n + pad) + 0.5) * (L / n)
    rng = np.random.default_rng(seed)
    sh = (n + 2 * pad,) * 3
    sx, cx = np.sin(x), np.cos(x)
    u = sx[:, None, None] * (cx[:, None] * cx[None, :])[None, :, :]
    u += noise * rng.uniform(-1, 1, sh)
    v = -(cx[:, None, None] * (sx[:, None] * cx[None, :])[None, :, :])
    v += noise * rng.uniform(-1, 1, sh)
    w = noise * rng.uniform(-1, 1, sh)
    zero = np.zeros(sh)
    prm = S.nse_params(dscheme=2, nlscheme=3, has_T=0, nu=1.0 / 1600, kappa=1.0 / 1600, dt=1e-3, cn_tol=1e-6,
                       les=1, vbc=[([S.FBC_DIRICHLET] * 6, [0.0] * 6)] * 3)
    # these solves stop before the 20-cycle cap, so the residual is checked
    # (one refinement) every 2 FP32 cycles
    mg = S.mg_params(vc_count=20, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2, pre=3, post=6, bc=S.BC_NEUMANN,
                     precision=S.PREC_MIXED, inner_cycles=2, max_refine=10)
    s = S.Nse(g, mg, prm)
    s.set(vx=u, vy=v, vz=w, p=zero, t=zero, pp=zero)
    del u, v, w
    for _ in range(max(warmup, 6)):
        r, st = s.step()
        S.check(r)
    r, ms = s.bench_steps(steps)
    S.check(r)
    d = s.diag()
    sg, t = s.sgs()
    vol = L ** 3
    s.close()
    return dict(ms_per_step=round(ms / steps, 3), steps=steps, cells=[n, n, n], jacobi_iters=st["jacobi_iters"],
                pressure_cycles=st["mg"]["cycles"], pressure_residual=st["mg"]["residual"], time=t, ke=d["ke"],
                divmax=d["divmax"],
                sgs_ke=sg["sgke"] / vol, sgs_dissipation=sg["disp"] / vol, nu_sgs=sg["nusg"] / vol)


def serving(count=240, seed=4, slots=8, budget=96 << 30):
    from saras_amd import serving as V

    def note(msg):
        print(f"[serving] {msg}", file=sys.stderr, flush=True)

    reqs = V.draw_requests(count, seed=seed)
    rhs = V.RhsCache()
    for r in reqs:
        rhs.get(r)
    note(f"{count} requests, right-hand sides ready")
    srv = S.Server(budget, slots=slots, use_graphs=1, largest_first=1)
    w = V.summarize(V.run_stream(srv, reqs, rhs, S.PREC_MIXED)[0])     # warm-up: hierarchies + graphs
    note(f"warm-up pass {w['seconds']:.2f} s")
    sat = V.summarize(V.run_stream(srv, reqs, rhs, S.PREC_MIXED)[0])
    note(f"saturation {sat['solves_per_s']:.1f} solves/s")
    rate = 0.7 * sat["solves_per_s"]
    opn = V.summarize(V.run_stream(srv, reqs, rhs, S.PREC_MIXED, rate=rate)[0])
    note(f"open loop at {rate:.1f}/s: p50 {opn['p50_ms']:.2f} ms, p99 {opn['p99_ms']:.2f} ms")
    srv.close()
    # the plain way: one request at a time, FP64, a solver built for each
    # request and released after it (poisson's constructor + mgSolve)
    cold = S.Server(budget, slots=1, use_graphs=0, keep_resident=0)
    one_cold = V.summarize(V.run_stream(cold, reqs, rhs, S.PREC_FP64)[0])
    cold.close()
    note(f"one at a time, FP64, solver per request: {one_cold['solves_per_s']:.1f} solves/s")
    # ... and with every hierarchy kept resident
    base = S.Server(budget, slots=1, use_graphs=0)
    V.run_stream(base, reqs, rhs, S.PREC_FP64)                          # warm-up
    one = V.summarize(V.run_stream(base, reqs, rhs, S.PREC_FP64)[0])
    note(f"one at a time, FP64, resident hierarchies: {one['solves_per_s']:.1f} solves/s")
    base.close()
    n3 = sum(1 for r in reqs if r["dim"] == 3)
    return {"requests": count, "requests_3d": n3, "slots": slots, "mem_budget_gb": budget / 2**30,
            "saturation": sat, "open_loop": dict(rate_per_s=round(rate, 1), **opn),
            "one_at_a_time_fp64": one_cold, "one_at_a_time_fp64_resident": one,
            "speedup_vs_one_at_a_time": round(sat["solves_per_s"] / one_cold["solves_per_s"], 3),
            "speedup_vs_one_at_a_time_resident": round(sat["solves_per_s"] / one["solves_per_s"], 3)}


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


def cpu_serving_baseline(count=240, seed=4, budget_s=15.0):
    """The serving mix solved one request at a time by the FP64 CPU oracle
    (same grids, parameters and right-hand sides as the GPU legs), over the
    requests that fit a bounded time budget, in stream order."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from saras_amd import serving as V

    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    reqs = V.draw_requests(count, seed=seed)
    rhs = V.RhsCache()
    t0 = time.perf_counter()
    done = 0
    for r in reqs:
        g = V.request_grid(r)
        p = V.request_params(r, S.PREC_FP64)
        go = O.grid_desc(tuple(g.n), dim=g.dim, periodic=tuple(g.periodic), pad=1)
        po = O.mg_params(vc_count=p.vc_count, vc_tol=p.vc_tol, solve_flag=p.solve_flag, mg_tol=p.mg_tol,
                         pre=p.pre_smooth, post=p.post_smooth, res_type=p.res_type, bc=p.bc)
        b = rhs.get(r)
        mg = O.Multigrid(go, po)
        mg.solve(np.zeros_like(b), b, 1)
        del mg
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return dict(value=round(done / el, 3), unit="solves/s", cores=threads, kind="port",
                sample=f"the first {done} requests of the serving stream, one at a time, FP64 oracle "
                       f"({threads} OpenMP threads), {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--profile-sweep", action="store_true",
                    help="only finest-level sweeps (FP32 and FP64) and the copy kernel, for counter runs")
    ap.add_argument("--rbc-only", action="store_true",
                    help="only the 256^3 Rayleigh-Benard time steps (warmup + steps, FP64 and mixed)")
    ap.add_argument("--no-rbc", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-les", action="store_true")
    ap.add_argument("--les-only", action="store_true", help="only the 384^3 Taylor-Green LES step")
    ap.add_argument("--serve-only", action="store_true", help="only the batched serving measurement")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--requests", type=int, default=240)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU path "
              "(multi-GPU sharding is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    nu1 = nu2 = 2
    if args.profile_sweep:
        profile_sweeps(args, nu1, nu2)
        return
    if args.rbc_only:
        print(json.dumps({"rbc_256_step": rbc_step(args.steps, args.warmup)}))
        return
    if args.les_only:
        print(json.dumps({"les_384_step": les_step(args.steps, args.warmup)}))
        return
    if args.serve_only:
        print(json.dumps({"serving": serving(count=args.requests, slots=args.slots)}))
        return
    rates = cycle_rates(args, nu1, nu2)
    head, f64 = rates["mixed"], rates["fp64"]
    value = head["bv"] / (head["ms"] * 1e-3) / 1e12
    achieved = head["launch_bytes"] / (head["launch_ms"] * 1e-3) / 1e9
    copy_n = 1 << 28
    S.bench_stream_copy(copy_n, 2)
    copy_ms = S.bench_stream_copy(copy_n, 10) / 10
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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "poisson_vcycle_256^3_PPN_z-tanh1.5_neumann_V(2,2)", "cells": [N, N, N],
                   "levels": len(head["shapes"]), "cycle_precision": "fp32 inner cycle of the mixed solver"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": round(DOMINANT_TRAFFIC_BYTES),
                     "kernel": DOMINANT, "launch_ms": round(head["launch_ms"], 5),
                     "bytes_per_launch": head["launch_bytes"]},
        "fp64_vcycle": {"ms": round(f64["ms"], 4),
                        "tbps": round(f64["bv"] / (f64["ms"] * 1e-3) / 1e12, 4),
                        "pass_launch_ms": round(f64["launch_ms"], 5),
                        "pass_gbs": round(f64["launch_bytes"] / (f64["launch_ms"] * 1e-3) / 1e9, 1)},
        "stream_copy_gbs": round(copy_gbs, 1),
    }
    if not args.no_solve:
        for n, reps in ((256, 3), (512, 2)):
            sc = solve_compare(nu1, nu2, n=n, reps=reps)
            out[f"solve_{n}^3_periodic_to_1e-10"] = {
                "fp64_ms": round(sc["fp64"]["ms"], 3), "fp64_cycles": sc["fp64"]["cycles"],
                "fp64_residual": sc["fp64"]["residual"], "fp64_converged": sc["fp64"]["converged"],
                "mixed_ms": round(sc["mixed"]["ms"], 3), "mixed_fp32_cycles": sc["mixed"]["cycles"],
                "mixed_refinements": sc["mixed"]["refinements"], "mixed_residual": sc["mixed"]["residual"],
                "mixed_converged": sc["mixed"]["converged"],
                "speedup": round(sc["fp64"]["ms"] / sc["mixed"]["ms"], 3),
                "mixed_vs_fp64_l2_rel": sc["diff_l2_rel"], "mixed_vs_fp64_linf_rel": sc["diff_linf_rel"]}
            print(f"solve {n}^3 done", file=sys.stderr, flush=True)
    if not args.no_serve:
        out["serving"] = serving(count=args.requests, slots=args.slots)
    if not args.no_rbc:
        out["rbc_256_step"] = rbc_step(3, 1)
    if not args.no_les:
        out["les_384_step"] = les_step(3, 6)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nu1, nu2, f64["shapes"])
        if not args.no_serve:
            out["cpu_baseline"]["serving"] = cpu_serving_baseline(count=args.requests)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
