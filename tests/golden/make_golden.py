"""Writes the oracle fixtures of the two time-stepping configurations, run at
their decks' settings by the FP64 CPU oracle (oracle/, built by
`make -C oracle`).  Run from the repository root:

    python tests/golden/make_golden.py rbc256      # about 100 minutes on 8 cores
    python tests/golden/make_golden.py tgv_les64   # about 15 minutes on 8 cores

rbc256 -> tests/golden/rbc256_steps.json
    3-D Rayleigh-Benard convection, 256^3 cells, periodic x and y, no-slip
    walls in z, double-sided tanh stretching beta = 1.5 in z, Ra = 1e7, Pr = 1
    with rbcType 2 (nu = sqrt(Pr/Ra), kappa = 1/sqrt(Pr Ra), Fb = 1; timestep.cc:70-98,
    buoyantForce.cc:48-60), lsRK3 with Jacobi CN (cnTolerance 1e-6),
    dScheme 2, Morinishi non-linear term, dt = 1e-3, V(3,6) pressure solves
    of at most 20 cycles to 1e-6 with "Solve Coarsest" off (with it on the
    reference algorithm's pressure solve diverges on this grid).  Initial
    state u = 0, P = 1, T = 1 - z + U(-1e-3, 1e-3) (seed 2), 100 steps as the
    deck asks.  Per step: the tseries quantities (KE, max|div|, NuVol, NuWall;
    tseries.cc:209-490), the Jacobi counts and the pressure solve's cycles and
    residual; after step 3 and after the last step, per field: the SHA-256 of
    the core's bytes, L2 (rms), L-inf, sum.

tgv_les64 -> tests/golden/tgv_les64_history.json
    Decaying Taylor-Green vortex in a periodic (2 pi)^3 box at 64^3, Re =
    1600 (nu = 1/1600), the stretched-spiral-vortex sub-grid model, lsRK3
    with Jacobi CN (1e-6), dScheme 2, Morinishi, dt = 1e-3, 500 steps,
    V(3,6) pressure solves of at most 20 cycles to 1e-6.  Initial state: the
    Taylor-Green field of taylorGreen.cc:75-98 plus seeded U(-1, 1) noise of
    1 % of the peak velocity (seed 5) on every velocity component.  Per step:
    the run time, KE, max|div| and the three sub-grid diagnostics
    (spiral.cc:186-206: sum |K| dV, dissipation dV, nu_t dV, each divided by
    the volume as tseries does).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def field_summary(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return dict(sha256=hashlib.sha256(a.tobytes()).hexdigest(), l2=float(np.sqrt(np.mean(a * a))),
                linf=float(np.abs(a).max()), sum=float(a.sum()))


# ---- config 2: 256^3 Rayleigh-Benard ----------------------------------------
RBC = dict(n=256, pad=2, steps=100, early=3, ra=1e7, pr=1.0, dt=1e-3, seed=2)


def rbc_mg():
    return dict(vc_count=20, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2, pre=3, post=6)


def rbc_initial(z, n, pad, seed):
    """T = 1 - z + U(-1e-3, 1e-3) over the core (temperatureProfiles.cc:55-76
    with the seeded perturbation), u = 0, P = 1, Pp = 0."""
    sh = (n + 2 * pad,) * 3
    rng = np.random.default_rng(seed)
    T = np.zeros(sh)
    T[:] = 1.0 - z[None, None, :]
    c = (slice(pad, -pad),) * 3
    T[c] += rng.uniform(-1e-3, 1e-3, (n, n, n))
    zero = np.zeros(sh)
    return dict(vx=zero, vy=zero.copy(), vz=zero.copy(), p=np.ones(sh), t=T, pp=zero.copy())


def make_rbc256(threads):
    c = RBC
    n, pad = c["n"], c["pad"]
    ph = O.nse_physics(5, ra=c["ra"], pr=c["pr"], rbc_type=2, g=(0.0, 0.0, -1.0))
    nu, kappa = ph["nu"], ph["kappa"]
    g = O.grid_desc(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), pad=pad)
    d = O.nse_desc(g, mg=O.mg_params(bc=O.BC_NEUMANN, **rbc_mg()), dscheme=2, nlscheme=3, has_T=1, nu=nu,
                   kappa=kappa, dt=c["dt"], cn_tol=1e-6, buoy=ph["buoy"])
    s = O.Nse(d)
    z = O.field_metrics(g, 2)["x"]
    s.set(**rbc_initial(z, n, pad, c["seed"]))
    hist = []
    core = (slice(pad, -pad),) * 3
    early = None
    for step in range(c["steps"]):
        t0 = time.time()
        r = s.step()
        assert r == 0, r
        dg = s.diag()
        mg = s.mg_stats()
        hist.append(dict(step=step + 1, ke=dg["ke"], divmax=dg["divmax"], nuvol=dg["nuvol"], nuwall=dg["nuwall"],
                         jacobi_iters=list(s.jacobi_iters()), mg_cycles=mg["cycles"], mg_residual=mg["residual"]))
        print(f"rbc256 step {step + 1}: {time.time() - t0:.1f} s, {hist[-1]}", flush=True)
        if step + 1 == c["early"]:
            f = s.get()
            early = {k: field_summary(f[k][core]) for k in ("vx", "vy", "vz", "p", "t")}
    f = s.get()
    fields = {k: field_summary(f[k][core]) for k in ("vx", "vy", "vz", "p", "t")}
    return dict(config=dict(RBC, nu=nu, kappa=kappa, mg=rbc_mg(), threads=threads), history=hist,
                fields_early=early, fields=fields)


# ---- config 5: Taylor-Green LES at 64^3 -------------------------------------
TGV = dict(n=64, pad=2, steps=500, re=1600.0, dt=1e-3, noise=0.01, seed=5)


def tgv_mg():
    return dict(vc_count=20, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2, pre=3, post=6)


def tgv_initial(n, pad, noise, seed):
    L = 2 * np.pi
    x = (np.arange(-pad, n + pad) + 0.5) * (L / n)
    sh = (n + 2 * pad,) * 3
    rng = np.random.default_rng(seed)
    sx, cx = np.sin(x), np.cos(x)
    u = sx[:, None, None] * (cx[:, None] * cx[None, :])[None, :, :]
    u = u + noise * rng.uniform(-1, 1, sh)
    v = -(cx[:, None, None] * (sx[:, None] * cx[None, :])[None, :, :])
    v = v + noise * rng.uniform(-1, 1, sh)
    w = noise * rng.uniform(-1, 1, sh)
    zero = np.zeros(sh)
    return dict(vx=u, vy=v, vz=w, p=zero, t=zero.copy(), pp=zero.copy())


def make_tgv_les64(threads):
    c = TGV
    n, pad = c["n"], c["pad"]
    L = 2 * np.pi
    g = O.grid_desc(n, periodic=(1, 1, 1), length=(L, L, L), pad=pad)
    nu = O.nse_physics(2, re=c["re"])["nu"]
    d = O.nse_desc(g, mg=O.mg_params(bc=O.BC_NEUMANN, **tgv_mg()), dscheme=2, nlscheme=3, has_T=0,
                   nu=nu, kappa=nu, dt=c["dt"], cn_tol=1e-6, les=1)
    s = O.Nse(d)
    s.set(**tgv_initial(n, pad, c["noise"], c["seed"]))
    vol = L ** 3
    hist = []
    t0 = time.time()
    for step in range(c["steps"]):
        r = s.step()
        assert r == 0, r
        dg = s.diag()
        sg = s.sgs_diag()
        hist.append(dict(step=step + 1, time=s.time, ke=dg["ke"], divmax=dg["divmax"], sgs_ke=sg[0] / vol,
                         sgs_dissipation=sg[1] / vol, nu_sgs=sg[2] / vol))
        if (step + 1) % 25 == 0:
            print(f"tgv_les64 step {step + 1}: {time.time() - t0:.0f} s, ke {dg['ke']:.9f}", flush=True)
    f = s.get()
    core = (slice(pad, -pad),) * 3
    fields = {k: field_summary(f[k][core]) for k in ("vx", "vy", "vz", "p")}
    return dict(config=dict(TGV, nu=nu, length=L, mg=tgv_mg(), threads=threads), history=hist,
                fields=fields)


def main():
    which = sys.argv[1:] or ["rbc256", "tgv_les64"]
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    for w in which:
        if w == "rbc256":
            out, name = make_rbc256(threads), "rbc256_steps.json"
        elif w == "tgv_les64":
            out, name = make_tgv_les64(threads), "tgv_les64_history.json"
        else:
            raise SystemExit(f"unknown fixture {w}")
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(out, fh, indent=1)
        print(f"wrote {name}", flush=True)


if __name__ == "__main__":
    main()
