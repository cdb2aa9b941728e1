"""Benchmark of the multigrid Poisson solve on one MI355X.

Workload (one "step" = one solve): the 256^3 Poisson problem on the periodic
unit cube, RHS = 64 Fourier modes with random phases and amplitudes
~ U(0.5, 1) |k|^-2 (seeded, zero mean), RB-SOR omega = 1, V(2,2) cycles,
coarsest 2^3 level solved to 1e-2, solved from a zero initial guess to a
relative L-inf residual of 1e-10 (the FP64 accuracy the plain path reaches)
by the mixed-precision solver: FP32 V-cycles inside an FP64 residual and
iterative-refinement loop, inputs and outputs resident in HBM.  The plain
FP64 solve of the same problem is timed beside it.

value: algorithmic HBM bytes of the mixed solve / its device time, in TB/s:
FP32 cycles x B_V + refinement passes x 32 N + the initial residual 20 N,
with the per-level V-cycle model  B_l = N_l s (3 (nu1 + nu2) + 2 + 1/8 + 2 + 1/8)
(3 words per smoothing sweep, residual+restriction reads x, b and writes
b_{l+1}, prolongation reads e_{l+1} and reads/writes x), s = bytes per word;
a refinement pass reads x64, b64, e32 and writes x64, r32; the initial
residual reads x64, b64 and writes r32.
fp64_equivalent_tbps: the bytes the plain FP64 solve moves for the same
result (its cycles x (B_V + 16 N for computeError)) / the mixed solve's time.

fp32_vcycle / fp64_vcycle: one V-cycle of each precision on the grid of the
256^3 Rayleigh-Benard configuration (periodic x, y; Neumann z walls; z-tanh
beta = 1.5), B_V / time per cycle.
solve_512^3: the same comparison at 512^3 (config 3's grid and RHS).

roofline: the dominant kernel of the headline solve (one colour pass of the
finest level on colour-split rows, FP32): algorithmic 1.5 words per cell per
launch (half of a sweep's 3) / average launch time (HIP events around
finest-level sweeps on the solve's own handle and stream).
cpu_baseline: the FP64 CPU oracle solving the same 256^3 problem to the same
tolerance (the reference algorithm, OpenMP threads = the host's CPU share).

serving: the batched request loop on the Poisson request mix (75 % 3-D
n^3, n log-uniform in [32, 256]; 25 % 2-D n^2, n in [64, 2048]; rounded to
multigrid-friendly multiples of 16; Dirichlet / Neumann / periodic; 8-mode
Fourier RHS; tolerance 1e-8; "Solve Coarsest" on): converged solves/s with
every request queued at once (saturation; free slots take the
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

les_384_step: the 384^3 decaying Taylor-Green LES (Re 1600, spiral sub-grid
model, RK3/CN) per step in the deck's precision recipe (FP32 V-cycles in
the FP64 pressure refinement, the sub-grid structure-function integral in
FP32, every other stencil FP64) and in FP64 throughout, with the fields'
difference between the two.
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
DOMINANT = "k_rb_split_zp<float,3,true,false,true,false>"
# HBM bytes per launch of the dominant kernel (one FP32 colour pass of the
# finest level of the z-periodic 256^3 grid, SOR parameter 1) from rocprofv3
# --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch), calibrated against the
# copy kernel (FETCH x 2.0, WRITE x 1.0): (2 * 33873.2 + 32768.0) KiB.
# Source: profiles/headline_sweep_hbm_traffic.md
DOMINANT_TRAFFIC_BYTES = (2 * 33873.2 + 32768.0) * 1024


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


def solve_bytes(shapes, nu1, nu2, prec, cycles, refinements):
    """Algorithmic HBM bytes of a whole solve (see the module docstring)."""
    n0 = float(np.prod(shapes[0]))
    if prec == S.PREC_FP64:
        return cycles * (bytes_per_vcycle(shapes, 8, nu1, nu2) + 16 * n0)
    return cycles * bytes_per_vcycle(shapes, 4, nu1, nu2) + refinements * 32 * n0 + 20 * n0


def headline_solve(steps, warmup, nu1, nu2, n=N, tol=1e-10, inner=4):
    """The headline: `steps` timed mixed-precision solves of the 256^3 periodic
    Fourier problem (each from a zero initial guess, device arrays), after
    `warmup` untimed ones; then the plain FP64 solve of the same problem, the
    dominant kernel's launch time on the mixed handle, and the difference of
    the two solutions."""
    g = S.grid_desc((n, n, n), periodic=(1, 1, 1))
    b = fourier_rhs(n)
    bd = S.DeviceArray.from_host(b)
    xd = S.DeviceArray(b.nbytes)
    out, sols = {}, {}
    for name, prec in (("mixed", S.PREC_MIXED), ("fp64", S.PREC_FP64)):
        p = S.mg_params(vc_count=100, vc_tol=tol, mg_tol=1e-2, pre=nu1, post=nu2, bc=S.BC_NEUMANN,
                        precision=prec, inner_cycles=inner, max_refine=40)
        mg = S.Multigrid(g, p)
        shapes = [mg.level_shape(lv) for lv in range(mg.levels)]
        k = steps if prec == S.PREC_MIXED else max(3, steps // 4)
        for _ in range(warmup):
            xd.zero()
            mg.solve_device(xd, bd)
        S.device_sync()
        t0 = time.perf_counter()
        dev_ms, stats = [], []
        for _ in range(k):
            xd.zero()
            st, status = mg.solve_device(xd, bd)
            S.check(status)
            dev_ms.append(st["device_ms"])
            stats.append(st)
        S.device_sync()
        wall = (time.perf_counter() - t0) * 1e3 / k
        sols[name] = xd.to_host(b.shape)[1:-1, 1:-1, 1:-1]
        st = stats[-1]
        assert all(s["cycles"] == st["cycles"] and s["refinements"] == st["refinements"] for s in stats)
        ms = float(np.mean(dev_ms))
        bts = solve_bytes(shapes, nu1, nu2, prec, st["cycles"], st["refinements"])
        r = dict(ms=ms, ms_min=float(np.min(dev_ms)), wall_ms=wall, solves=k, cycles=st["cycles"],
                 refinements=st["refinements"], residual=st["residual"], converged=st["converged"],
                 bytes=bts, tbps=bts / (ms * 1e-3) / 1e12, shapes=shapes)
        if prec == S.PREC_MIXED:
            # the dominant kernel: finest-level FP32 colour passes on this
            # handle's stream (a sweep = two launches of equal work)
            mg.bench_sweeps(2)
            reps = 20
            r["launch_ms"] = mg.bench_sweeps(reps) / reps / 2
            r["launch_bytes"] = 1.5 * float(np.prod(shapes[0])) * 4
        out[name] = r
        mg.close()
    bd.free()
    xd.free()
    d = (sols["mixed"] - sols["mixed"].mean()) - (sols["fp64"] - sols["fp64"].mean())
    ref = sols["fp64"] - sols["fp64"].mean()
    out["diff_l2_rel"] = float(np.linalg.norm(d) / np.linalg.norm(ref))
    out["diff_linf_rel"] = float(np.abs(d).max() / np.abs(ref).max())
    return out


def field_diffs(a, b):
    """Per field: L2 and L-inf of a - b over the core, each relative to b's
    (absolute where b is zero).  The pressure of these runs (periodic or
    Neumann walls) is defined up to a constant: it is compared modulo its mean."""
    out = {}
    for k in b:
        x, y = a[k], b[k]
        if k == "p":
            x, y = x - x.mean(), y - y.mean()
        d = x - y
        l2, li = float(np.sqrt(np.mean(d * d))), float(np.abs(d).max())
        n2, ni = float(np.sqrt(np.mean(y * y))), float(np.abs(y).max())
        out[k] = dict(l2_rel=l2 / n2 if n2 > 0 else l2, linf_rel=li / ni if ni > 0 else li)
    return out


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
    ph = S.nse_physics(5, ra=1e7, pr=1.0, rbc_type=2, g=(0.0, 0.0, -1.0))   # Ra 1e7, Pr 1, rbcType 2
    prm = S.nse_params(dscheme=2, nlscheme=3, has_T=1, nu=ph["nu"], kappa=ph["kappa"], dt=1e-3, cn_tol=1e-6,
                       buoy=ph["buoy"])
    out, fields = {}, {}
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
        fields[name] = {k: v[(slice(pad, -pad),) * 3].copy() for k, v in s.get().items() if k != "pp"}
        s.close()
    out["mixed_vs_fp64"] = field_diffs(fields["mixed"], fields["fp64"])
    out["steps_compared"] = warmup + steps + 1
    return out


def tgv_les_case(n, noise=0.01, seed=5, sf32=0):
    """Decaying Taylor-Green vortex in a periodic (2 pi)^3 box, Re = 1600
    (nu = 1/1600, unit velocity and length), seeded uniform noise of 1 % of
    the peak velocity, spiral sub-grid model, RK3/CN with FP64 stencils
    (sf32: the sub-grid model's structure-function integral in FP32).
    Returns the grid, the parameters and the initial fields."""
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
    nu = S.nse_physics(2, re=1600.0)["nu"]
    prm = S.nse_params(dscheme=2, nlscheme=3, has_T=0, nu=nu, kappa=nu, dt=1e-3, cn_tol=1e-6,
                       les=1, vbc=[([S.FBC_DIRICHLET] * 6, [0.0] * 6)] * 3, les_sf32=sf32)
    return g, prm, dict(vx=u, vy=v, vz=w, p=zero, t=zero, pp=zero)


def tgv_les_mg(prec):
    # these solves stop before the 20-cycle cap, so the residual is checked
    # (one refinement) every 2 FP32 cycles
    return S.mg_params(vc_count=20, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2, pre=3, post=6, bc=S.BC_NEUMANN,
                       precision=prec, inner_cycles=2, max_refine=10)


def les_step(steps, warmup, n=384):
    """The Taylor-Green LES in the deck's precision recipe, timed: FP32
    V-cycles inside the FP64 residual and refinement of the pressure solve,
    the sub-grid structure-function integral in FP32; the same run in FP64
    throughout beside it.  The model enters once solTime > 5 dt, so at least
    six warm-up steps run before the timed ones."""
    pad = 2
    L = 2 * np.pi
    res, fields = {}, {}
    for name, prec, sf32 in (("mixed", S.PREC_MIXED, 1), ("fp64", S.PREC_FP64, 0)):
        g, prm, f0 = tgv_les_case(n, sf32=sf32)
        s = S.Nse(g, tgv_les_mg(prec), prm)
        s.set(**f0)
        for _ in range(max(warmup, 6)):
            r, st = s.step()
            S.check(r)
        r, ms = s.bench_steps(steps)
        S.check(r)
        d = s.diag()
        sg, t = s.sgs()
        res[name] = dict(ms=ms, st=st, d=d, sg=sg, t=t, sgs_ms=s.bench_sgs(3))
        fields[name] = {k: v[(slice(pad, -pad),) * 3].copy() for k, v in s.get().items() if k in ("vx", "vy", "vz", "p")}
        s.close()
    vol = L ** 3
    m = res["mixed"]
    return dict(ms_per_step=round(m["ms"] / steps, 3), fp64_ms_per_step=round(res["fp64"]["ms"] / steps, 3),
                steps=steps, cells=[n, n, n], jacobi_iters=m["st"]["jacobi_iters"],
                pressure_cycles=m["st"]["mg"]["cycles"], pressure_residual=m["st"]["mg"]["residual"], time=m["t"],
                ke=m["d"]["ke"], ke_fp64=res["fp64"]["d"]["ke"], divmax=m["d"]["divmax"],
                sgs_ke=m["sg"]["sgke"] / vol, sgs_dissipation=m["sg"]["disp"] / vol, nu_sgs=m["sg"]["nusg"] / vol,
                sgs_ke_fp64=res["fp64"]["sg"]["sgke"] / vol, mixed_vs_fp64=field_diffs(fields["mixed"], fields["fp64"]),
                steps_compared=max(warmup, 6) + steps, sgs_stress=sgs_rate(res["fp64"]["sgs_ms"], n),
                sgs_stress_sf32_ms=round(m["sgs_ms"], 4))


def les_history(steps, n=384, sf32=1):
    """The Taylor-Green LES over `steps` steps, once in the deck's precision
    recipe (mixed-precision pressure solves; with sf32, the sub-grid
    structure-function integral in FP32) and once in FP64 throughout (the
    FP64 path follows the oracle's history to 1e-8 at 64^3 over 500 steps):
    per step the run time, KE, max|div| and the sub-grid diagnostics, and the
    dissipation eps = -dKE/dt by centred differences.  Reports the largest
    differences of the two histories, a few samples of them, and the final
    fields' differences."""
    pad = 2
    L = 2 * np.pi
    vol = L ** 3
    hist, fields, wall = {}, {}, {}
    for name, prec, sf in (("mixed", S.PREC_MIXED, sf32), ("fp64", S.PREC_FP64, 0)):
        g, prm, f0 = tgv_les_case(n, sf32=sf)
        s = S.Nse(g, tgv_les_mg(prec), prm)
        s.set(**f0)
        h = []
        t0 = time.time()
        for step in range(steps):
            r, st = s.step()
            S.check(r)
            d = s.diag()
            sg, t = s.sgs()
            h.append(dict(time=t, ke=d["ke"], divmax=d["divmax"], sgs_ke=sg["sgke"] / vol,
                          sgs_dissipation=sg["disp"] / vol, nu_sgs=sg["nusg"] / vol))
            if (step + 1) % 50 == 0:
                print(f"les history {name} step {step + 1}: ke {d['ke']:.12e}", file=sys.stderr, flush=True)
        wall[name] = time.time() - t0
        hist[name] = h
        fields[name] = {k: v[(slice(pad, -pad),) * 3].copy() for k, v in s.get().items() if k in ("vx", "vy", "vz", "p")}
        s.close()
    a, b = hist["mixed"], hist["fp64"]
    worst = {}
    for k in ("ke", "sgs_ke", "sgs_dissipation", "nu_sgs"):
        worst[k] = max(abs(x[k] - y[k]) / max(abs(y[k]), 1e-300) for x, y in zip(a, b))
    worst["divmax"] = max(abs(x["divmax"] - y["divmax"]) for x, y in zip(a, b)) / max(y["divmax"] for y in b)
    dt = 1e-3
    ke_a = np.array([x["ke"] for x in a])
    ke_b = np.array([y["ke"] for y in b])
    eps_a = -(ke_a[2:] - ke_a[:-2]) / (2 * dt)
    eps_b = -(ke_b[2:] - ke_b[:-2]) / (2 * dt)
    worst["eps"] = float(np.abs(eps_a - eps_b).max() / np.abs(eps_b).max())
    marks = sorted(set([0, steps // 4, steps // 2, 3 * steps // 4, steps - 1]))
    samples = [dict(step=i + 1, time=b[i]["time"], ke_fp64=b[i]["ke"], ke_mixed=a[i]["ke"],
                    eps_fp64=float(eps_b[i - 1]) if 1 <= i <= steps - 2 else None,
                    eps_mixed=float(eps_a[i - 1]) if 1 <= i <= steps - 2 else None,
                    sgs_dissipation_fp64=b[i]["sgs_dissipation"], sgs_dissipation_mixed=a[i]["sgs_dissipation"])
               for i in marks]
    return dict(cells=[n, n, n], steps=steps, sgs_sf32=bool(sf32), wall_s={k: round(v, 1) for k, v in wall.items()},
                mixed_vs_fp64_history_max_rel=worst, samples=samples,
                mixed_vs_fp64_final_fields=field_diffs(fields["mixed"], fields["fp64"]))


# FP64 operations per cell of the sub-grid stress kernel at 384^3, from
# rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 (wave instructions x
# 64 lanes, an FMA counted twice): 199.2 GFLOP per launch.
# Source: profiles/les_step_384cubed.md
SGS_FLOP_PER_CELL = 3518.1
FP64_VECTOR_PEAK_TFLOPS = 78.6


def sgs_rate(ms, n):
    """The compute-bound sub-grid stress kernel, outside the TB/s metric:
    its launch time on the handle's stream and the FP64 rate it reaches."""
    tf = SGS_FLOP_PER_CELL * float(n) ** 3 / (ms * 1e-3) / 1e12
    return dict(ms_per_launch=round(ms, 4), flop_per_cell=SGS_FLOP_PER_CELL, tflops_fp64=round(tf, 2),
                peak_fp64_vector_tflops=FP64_VECTOR_PEAK_TFLOPS, frac=round(tf / FP64_VECTOR_PEAK_TFLOPS, 4))


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
    # the same server with plain FP64 solves: the batching gain alone
    s64 = S.Server(budget, slots=slots, use_graphs=1, largest_first=1)
    V.run_stream(s64, reqs, rhs, S.PREC_FP64)                           # warm-up
    sat64 = V.summarize(V.run_stream(s64, reqs, rhs, S.PREC_FP64)[0])
    s64.close()
    note(f"saturation, FP64 solves {sat64['solves_per_s']:.1f} solves/s")
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
            "saturation": sat, "open_loop": dict(rate_per_s=round(rate, 1), **opn), "saturation_fp64": sat64,
            "one_at_a_time_fp64": one_cold, "one_at_a_time_fp64_resident": one,
            "speedup_vs_one_at_a_time": round(sat["solves_per_s"] / one_cold["solves_per_s"], 3),
            "speedup_vs_one_at_a_time_resident": round(sat["solves_per_s"] / one["solves_per_s"], 3),
            "speedup_fp64_server_vs_one_at_a_time": round(sat64["solves_per_s"] / one_cold["solves_per_s"], 3),
            "note": "solves_per_s counts converged solves; the mixed server against one-at-a-time FP64 combines the "
                    "batching and precision gains, the FP64 server against it shows batching alone"}


def host_threads():
    """Threads for the CPU legs: one per physical core of this process's CPU
    set, at most OMP_NUM_THREADS (the host's CPU share for one GPU where the
    environment sets it).  Returns (threads, description)."""
    cpus = sorted(os.sched_getaffinity(0))
    phys = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            phys.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            phys.add(("?", str(c)))
    model = "?"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    n = len(phys) if share <= 0 else min(len(phys), share)
    desc = (f"{model}; {len(cpus)} logical CPUs = {len(phys)} physical cores in this process's CPU set, "
            f"OMP_NUM_THREADS={share or 'unset'}: {n} threads, one per physical core")
    return max(1, n), desc


def cpu_baseline(nu1, nu2, n=N, tol=1e-10, budget_s=20.0):
    """The FP64 CPU oracle (the reference algorithm) solving the headline
    problem to the same tolerance, as many whole solves as fit the budget
    (at least one)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads, desc = host_threads()
    O.set_threads(threads)
    g = O.grid_desc((n, n, n), periodic=(1, 1, 1))
    mg = O.Multigrid(g, O.mg_params(vc_count=100, vc_tol=tol, mg_tol=1e-2, pre=nu1, post=nu2, bc=O.BC_NEUMANN))
    b = fourier_rhs(n)
    shapes = [mg.level_shape(lv) for lv in range(mg.levels)]
    t0 = time.perf_counter()
    k, cycles = 0, 0
    while True:
        _, st = mg.solve(np.zeros_like(b), b, 1)
        k += 1
        cycles = st["cycles"]
        el = time.perf_counter() - t0
        if el > budget_s * 0.5 or k >= 5:
            break
    per = el / k
    bts = solve_bytes(shapes, nu1, nu2, S.PREC_FP64, cycles, 0)
    return dict(value=round(bts / per / 1e12, 5), unit="TB/s", cores=threads, kind="port",
                ms_per_solve=round(per * 1e3, 1),
                sample=f"{k} whole FP64 solves of the 256^3 headline problem to 1e-10 ({cycles} V(2,2) cycles each) "
                       f"on the FP64 oracle, {per * 1e3:.0f} ms per solve; {desc}")


def cpu_serving_baseline(count=240, seed=4, budget_s=20.0):
    """The serving mix solved one request at a time by the FP64 CPU oracle
    (same grids, parameters and right-hand sides as the GPU legs) on a seeded
    random sample of the stream that fits the time budget, so the sample spans
    the size range.  value: the stream's solves/s estimated from the sample's
    cell throughput (a solve costs about cells x cycles, and the cycle count
    is nearly the same across the mix): count / (stream cells / sample cells
    per second)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from saras_amd import serving as V

    threads, desc = host_threads()
    O.set_threads(threads)
    reqs = V.draw_requests(count, seed=seed)
    cells = lambda r: float(r["n"]) ** r["dim"]  # noqa: E731
    order = np.random.default_rng(seed + 1).permutation(len(reqs))
    rhs = V.RhsCache()
    t_solve, done, c_done, dims = 0.0, 0, 0.0, []
    t_all = time.perf_counter()
    for q in order:
        r = reqs[int(q)]
        g = V.request_grid(r)
        p = V.request_params(r, S.PREC_FP64)
        go = O.grid_desc(tuple(g.n), dim=g.dim, periodic=tuple(g.periodic), pad=1)
        po = O.mg_params(vc_count=p.vc_count, vc_tol=p.vc_tol, solve_flag=p.solve_flag, mg_tol=p.mg_tol,
                         pre=p.pre_smooth, post=p.post_smooth, res_type=p.res_type, bc=p.bc)
        b = rhs.get(r)
        t0 = time.perf_counter()
        mg = O.Multigrid(go, po)
        mg.solve(np.zeros_like(b), b, 1)
        del mg
        t_solve += time.perf_counter() - t0
        done += 1
        c_done += cells(r)
        dims.append((r["dim"], r["n"]))
        if time.perf_counter() - t_all > budget_s:
            break
    c_all = sum(cells(r) for r in reqs)
    est = count / (c_all / (c_done / t_solve))
    n3 = [n for d, n in dims if d == 3]
    n2 = [n for d, n in dims if d == 2]
    return dict(value=round(est, 3), unit="solves/s", cores=threads, kind="port",
                sample=f"{done} requests drawn at random from the stream ({len(n3)} 3-D, n {min(n3, default=0)}-"
                       f"{max(n3, default=0)}; {len(n2)} 2-D, n {min(n2, default=0)}-{max(n2, default=0)}), one at a "
                       f"time on the FP64 oracle, {t_solve:.1f} s of solves = {done / t_solve:.3f} solves/s on the "
                       f"sample; scaled to the stream by cells; {desc}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the 512^3 solve comparison")
    ap.add_argument("--profile-sweep", action="store_true",
                    help="only finest-level sweeps (FP32 and FP64) and the copy kernel, for counter runs")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline solves (mixed and FP64), for profiling")
    ap.add_argument("--n", type=int, default=N, help="grid extent of the --headline-only solves")
    ap.add_argument("--rbc-only", action="store_true",
                    help="only the 256^3 Rayleigh-Benard time steps (warmup + steps, FP64 and mixed)")
    ap.add_argument("--no-rbc", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-les", action="store_true")
    ap.add_argument("--les-only", action="store_true", help="only the 384^3 Taylor-Green LES step")
    ap.add_argument("--les-history", type=int, default=0, metavar="STEPS",
                    help="only the 384^3 Taylor-Green LES histories (mixed against FP64 pressure) over STEPS steps")
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
    if args.les_history:
        print(json.dumps({"les_384_history": les_history(args.les_history)}))
        return
    if args.les_only:
        print(json.dumps({"les_384_step": les_step(args.steps, args.warmup)}))
        return
    if args.serve_only:
        print(json.dumps({"serving": serving(count=args.requests, slots=args.slots)}))
        return
    hs = headline_solve(args.steps, args.warmup, nu1, nu2, n=args.n if args.headline_only else N)
    mx, f64 = hs["mixed"], hs["fp64"]
    if args.headline_only:
        print(json.dumps({"mixed_ms": mx["ms"], "fp64_ms": f64["ms"]}))
        return
    print("headline solves done", file=sys.stderr, flush=True)
    value = mx["tbps"]
    achieved = mx["launch_bytes"] / (mx["launch_ms"] * 1e-3) / 1e9
    rates = cycle_rates(args, nu1, nu2)
    copy_n = 1 << 28
    S.bench_stream_copy(copy_n, 2)
    copy_ms = S.bench_stream_copy(copy_n, 10) / 10
    copy_gbs = 16.0 * copy_n / (copy_ms * 1e-3) / 1e9
    out = {
        "metric": "V-cycle HBM TB/s at 256^3; batched solves/s + p99 latency vs plain FP64",
        "value": round(value, 4),
        "unit": "TB/s",
        "n_gpus": 1,
        "steps": mx["solves"],
        "warmup": args.warmup,
        "ms_per_step": round(mx["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycles + f64 residual/refinement (f64-accurate result)",
        "data": "synthetic",
        "config": {"workload": "poisson_solve_256^3_periodic_64-mode-fourier_V(2,2)_to_1e-10_mixed",
                   "cells": [N, N, N], "levels": len(mx["shapes"]), "tolerance": 1e-10,
                   "step": "one whole solve from a zero initial guess, device-resident arrays"},
        "solve": {"mixed_ms": round(mx["ms"], 4), "mixed_ms_min": round(mx["ms_min"], 4),
                  "mixed_wall_ms": round(mx["wall_ms"], 4), "mixed_fp32_cycles": mx["cycles"],
                  "mixed_refinements": mx["refinements"], "mixed_residual": mx["residual"],
                  "mixed_converged": mx["converged"], "mixed_bytes": mx["bytes"],
                  "fp64_ms": round(f64["ms"], 4), "fp64_cycles": f64["cycles"], "fp64_residual": f64["residual"],
                  "fp64_converged": f64["converged"], "fp64_bytes": f64["bytes"],
                  "fp64_tbps": round(f64["tbps"], 4),
                  "fp64_equivalent_tbps": round(f64["bytes"] / (mx["ms"] * 1e-3) / 1e12, 4),
                  "speedup_mixed_vs_fp64": round(f64["ms"] / mx["ms"], 3),
                  "mixed_vs_fp64_l2_rel": hs["diff_l2_rel"], "mixed_vs_fp64_linf_rel": hs["diff_linf_rel"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": round(DOMINANT_TRAFFIC_BYTES),
                     "kernel": DOMINANT, "launch_ms": round(mx["launch_ms"], 5),
                     "bytes_per_launch": mx["launch_bytes"]},
        "fp32_vcycle": {"grid": "256^3 PPN z-tanh 1.5 (Rayleigh-Benard)", "ms": round(rates["mixed"]["ms"], 4),
                        "tbps": round(rates["mixed"]["bv"] / (rates["mixed"]["ms"] * 1e-3) / 1e12, 4)},
        "fp64_vcycle": {"grid": "256^3 PPN z-tanh 1.5 (Rayleigh-Benard)", "ms": round(rates["fp64"]["ms"], 4),
                        "tbps": round(rates["fp64"]["bv"] / (rates["fp64"]["ms"] * 1e-3) / 1e12, 4),
                        "pass_launch_ms": round(rates["fp64"]["launch_ms"], 5),
                        "pass_gbs": round(rates["fp64"]["launch_bytes"] / (rates["fp64"]["launch_ms"] * 1e-3) / 1e9,
                                          1)},
        "stream_copy_gbs": round(copy_gbs, 1),
    }
    if not args.no_solve:
        sc = solve_compare(nu1, nu2, n=512, reps=2)
        out["solve_512^3_periodic_to_1e-10"] = {
            "fp64_ms": round(sc["fp64"]["ms"], 3), "fp64_cycles": sc["fp64"]["cycles"],
            "fp64_residual": sc["fp64"]["residual"], "fp64_converged": sc["fp64"]["converged"],
            "mixed_ms": round(sc["mixed"]["ms"], 3), "mixed_fp32_cycles": sc["mixed"]["cycles"],
            "mixed_refinements": sc["mixed"]["refinements"], "mixed_residual": sc["mixed"]["residual"],
            "mixed_converged": sc["mixed"]["converged"],
            "speedup": round(sc["fp64"]["ms"] / sc["mixed"]["ms"], 3),
            "mixed_vs_fp64_l2_rel": sc["diff_l2_rel"], "mixed_vs_fp64_linf_rel": sc["diff_linf_rel"]}
        print("solve 512^3 done", file=sys.stderr, flush=True)
    if not args.no_serve:
        out["serving"] = serving(count=args.requests, slots=args.slots)
    if not args.no_rbc:
        out["rbc_256_step"] = rbc_step(3, 1)
        print("rbc step done", file=sys.stderr, flush=True)
    if not args.no_les:
        out["les_384_step"] = les_step(3, 6)
        print("les step done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nu1, nu2)
        print("cpu solve baseline done", file=sys.stderr, flush=True)
        if not args.no_serve:
            out["cpu_baseline"]["serving"] = cpu_serving_baseline(count=args.requests)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
