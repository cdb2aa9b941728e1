"""Client side of the batched request loop: the Poisson request mix of the
serving workload and an open/closed-loop driver for saras_server_*.

Request mix: 75 % 3-D cubes with n drawn log-uniformly in [32, 256], 25 % 2-D
squares with n log-uniform in [64, 2048], each rounded to the nearest
multigrid-friendly extent (below); boundary conditions uniformly Dirichlet
(homogeneous, every face), Neumann (every face) or periodic; RHS = 8 seeded
Fourier modes (sine modes for Dirichlet, cosine modes for Neumann, zero-mean
complex modes for periodic) with amplitudes ~ U(0.5, 1) |k|^-2; tolerance
1e-8 (relative L-inf residual).

Multigrid-friendly extents: multiples of 16 whose quotient by 16 has an odd
part of at most 7 (32, 48, 64, 80, 96, 112, 128, 160, 192, 224, 256, 320, ...).
The reference's depth rule (grid::setVCDepth, grid.cc:678-742) halves an axis
while it stays even, so such an extent coarsens to at most 7 cells (14 with
walls on every face, where the rule stops one level earlier).  A multiple of
16 with a larger odd part (208 = 13 * 16, 2000 = 125 * 16) leaves a coarsest
level of up to 250 cells per axis that no V-cycle budget solves.
"""
import math
import time

import numpy as np

from . import (BC_DIRICHLET, BC_NEUMANN, PREC_FP64, PREC_MIXED, Server, grid_desc, mg_params)

KINDS = ("dirichlet", "neumann", "periodic")


def odd_part(m):
    while m % 2 == 0:
        m //= 2
    return m


def mg_extents(lo, hi):
    """Multigrid-friendly extents in [lo, hi]: multiples of 16 whose quotient
    by 16 has an odd part of at most 7."""
    return [m for m in range(16 * ((lo + 15) // 16), hi + 1, 16) if odd_part(m // 16) <= 7]


def nearest_extent(v, lo, hi):
    """The multigrid-friendly extent nearest to v (the smaller one on a tie)."""
    return min(mg_extents(lo, hi), key=lambda m: (abs(m - v), m))


def draw_requests(count, seed=4, frac3d=0.75, n3=(32, 256), n2=(64, 2048)):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        d3 = rng.random() < frac3d
        lo, hi = n3 if d3 else n2
        n = nearest_extent(math.exp(rng.uniform(math.log(lo), math.log(hi))), lo, hi)
        kind = KINDS[int(rng.integers(0, 3))]
        out.append(dict(tag=t, dim=3 if d3 else 2, n=n, kind=kind, rhs_seed=int(rng.integers(0, 2**31))))
    return out


def request_grid(req):
    dim, n, kind = req["dim"], req["n"], req["kind"]
    per = (1, 1, 1) if kind == "periodic" else (0, 0, 0)
    if dim == 3:
        return grid_desc((n, n, n), dim=3, periodic=per, pad=1)
    return grid_desc((n, 1, n), dim=2, periodic=(per[0], 0, per[2]), pad=1)


def request_params(req, precision, tol=1e-8):
    """V(2,2) with "Solve Coarsest" on (the reference deck's default) to an
    absolute coarse tolerance of 1e-10, at most 100 V-cycles.  On the
    multigrid-friendly extents every request then converges in about 8-11
    cycles; smoothing the coarsest level instead takes up to 80 cycles where
    it keeps 10-14 cells per axis.  A request still above tolerance after 100
    cycles would complete as NOT_CONVERGED, as the reference's mgSolve returns
    silently.  Mixed precision: up to 10 FP32 cycles per refinement (the
    solver stops each refinement at its tolerance by the measured convergence
    factor)."""
    bc = BC_DIRICHLET if req["kind"] == "dirichlet" else BC_NEUMANN
    return mg_params(vc_count=100, vc_tol=tol, solve_flag=1, mg_tol=1e-10, pre=2, post=2, res_type=0, bc=bc,
                     precision=precision, inner_cycles=10, max_refine=100)


def request_rhs(req, modes=8):
    """Padded (pad 1) RHS of the request: separable Fourier modes on cell centres."""
    dim, n, kind = req["dim"], req["n"], req["kind"]
    rng = np.random.default_rng(req["rhs_seed"])
    xc = (np.arange(n) + 0.5) / n
    axes = 3 if dim == 3 else 2
    acc = np.zeros((n,) * axes)
    for _ in range(modes):
        if kind == "dirichlet":
            k = rng.integers(1, 9, axes)
            f = [np.sin(np.pi * kk * xc) for kk in k]
            amp = rng.uniform(0.5, 1.0) / float(k @ k)
        elif kind == "neumann":
            while True:
                k = rng.integers(0, 9, axes)
                if k.max() >= 1:
                    break
            f = [np.cos(np.pi * kk * xc) for kk in k]
            amp = rng.uniform(0.5, 1.0) / float(k @ k)
        else:
            while True:
                k = rng.integers(-4, 5, axes)
                if np.abs(k).max() >= 1:
                    break
            phi = rng.uniform(0, 2 * np.pi)
            f = [np.exp(2j * np.pi * kk * xc) for kk in k]
            f[0] = f[0] * np.exp(1j * phi)
            amp = rng.uniform(0.5, 1.0) / float(k @ k)
        if axes == 3:
            m = f[0][:, None, None] * (f[1][:, None] * f[2][None, :])[None, :, :]
        else:
            m = f[0][:, None] * f[1][None, :]
        acc += amp * np.real(m)
    if kind != "dirichlet":
        acc -= acc.mean()
    b = np.zeros(tuple(s + 2 for s in acc.shape))
    b[(slice(1, -1),) * axes] = acc
    return b


class RhsCache:
    """RHS arrays by (dim, n, kind, seed)."""

    def __init__(self):
        self.c = {}

    def get(self, req):
        key = (req["dim"], req["n"], req["kind"], req["rhs_seed"])
        if key not in self.c:
            self.c[key] = request_rhs(req)
        return self.c[key]


def run_stream(server, reqs, rhs, precision, rate=None, collect=None):
    """Submit `reqs` to `server`: all at once (rate None, closed loop) or as a
    Poisson process of `rate` requests/s (seeded).  Returns per-request
    records; `collect` (a set of tags) keeps those solutions."""
    rng = np.random.default_rng(12345)
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
    t_arr = np.cumsum(gaps) - gaps[0]
    recs, sols = {}, {}
    xs = {}
    t0 = server.clock()
    for q, req in enumerate(reqs):
        if rate:
            wait = t0 + t_arr[q] - server.clock()
            if wait > 0:
                time.sleep(wait)
        g = request_grid(req)
        p = request_params(req, precision)
        b = rhs.get(req)
        x = np.zeros_like(b)
        xs[req["tag"]] = x
        server.submit(g, p, b, x, req["tag"])
        for c in server.poll(wait_ms=0):
            recs[c["tag"]] = c
    server.drain()
    for c in server.poll(max_n=len(reqs) + 1):
        recs[c["tag"]] = c
    for tag, x in xs.items():
        if collect is not None and tag in collect:
            sols[tag] = x
    return recs, sols


def summarize(recs):
    """solves_per_s counts converged solves only (a NOT_CONVERGED completion is
    not a solve); requests_per_s is the raw completion rate beside it."""
    lat = np.array([r["t_done"] - r["t_submit"] for r in recs.values()])
    t_first = min(r["t_submit"] for r in recs.values())
    t_last = max(r["t_done"] for r in recs.values())
    ok = sum(1 for r in recs.values() if r["status"] in (0,) and r["stats"]["converged"])
    failed = [r for r in recs.values() if r["status"] not in (0, 1)]
    if failed:
        raise RuntimeError(f"{len(failed)} requests failed, first: {failed[0]}")
    span = t_last - t_first
    return dict(requests=len(recs), converged=ok, seconds=span, solves_per_s=ok / span,
                requests_per_s=len(recs) / span, p50_ms=float(np.percentile(lat, 50) * 1e3),
                p99_ms=float(np.percentile(lat, 99) * 1e3), mean_ms=float(lat.mean() * 1e3))


# ---- short Navier-Stokes requests -------------------------------------------
#
# Two families of short runs of uneven sizes, each `steps` lsRK3 steps:
#   rbc: Rayleigh-Benard boxes n^3 (periodic x, y; no-slip z walls; z-tanh
#        beta = 1.5; Ra = 1e7, Pr = 1, rbcType 2; dScheme 2, Morinishi),
#        T = 1 - z + U(-1e-3, 1e-3), u = 0, P = 1;
#   tg:  Taylor-Green boxes n^3 (periodic (2 pi)^3, Re = 1600, dScheme 2,
#        Morinishi), the taylorGreen.cc field plus 1 % seeded noise.
# n is drawn log-uniformly in [16, 64] and rounded to a multiple of 8.
NSE_KINDS = ("rbc", "tg")


def draw_nse_requests(count, seed=6, n=(16, 64), steps=(2, 5), tag0=100000):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        kind = NSE_KINDS[int(rng.integers(0, 2))]
        m = int(round(math.exp(rng.uniform(math.log(n[0]), math.log(n[1]))) / 8.0)) * 8
        m = min(max(m, n[0]), n[1])
        out.append(dict(tag=tag0 + t, kind=kind, n=m, steps=int(rng.integers(steps[0], steps[1] + 1)),
                        seed=int(rng.integers(0, 2**31))))
    return out


def nse_request_setup(req):
    """(grid, mg params, nse params, initial fields) of a short run."""
    from . import FBC_DIRICHLET, GRID_TANH_DOUBLE, level_metrics, nse_params, nse_physics

    n, pad = req["n"], 2
    sh = (n + 2 * pad,) * 3
    rng = np.random.default_rng(req["seed"])
    mg = mg_params(vc_count=20, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2, pre=3, post=6, bc=BC_NEUMANN,
                   precision=PREC_MIXED, inner_cycles=5, max_refine=10)
    c = (slice(pad, -pad),) * 3
    if req["kind"] == "rbc":
        g = grid_desc(n, periodic=(1, 1, 0), gtype=(0, 0, GRID_TANH_DOUBLE), beta=(1.0, 1.0, 1.5), pad=pad)
        ph = nse_physics(5, ra=1e7, pr=1.0, rbc_type=2, g=(0.0, 0.0, -1.0))
        prm = nse_params(dscheme=2, nlscheme=3, has_T=1, nu=ph["nu"], kappa=ph["kappa"], dt=1e-3, cn_tol=1e-6,
                         buoy=ph["buoy"])
        z = level_metrics(g, 0, 2)["x"][1:-1]
        T = np.zeros(sh)
        T[c] = 1.0 - z[None, None, :] + rng.uniform(-1e-3, 1e-3, (n, n, n))
        f = dict(vx=np.zeros(sh), vy=np.zeros(sh), vz=np.zeros(sh), p=np.ones(sh), t=T)
    else:
        L = 2 * np.pi
        g = grid_desc(n, periodic=(1, 1, 1), length=(L, L, L), pad=pad)
        nu = nse_physics(2, re=1600.0)["nu"]
        prm = nse_params(dscheme=2, nlscheme=3, has_T=0, nu=nu, kappa=nu, dt=1e-3, cn_tol=1e-6,
                         vbc=[([FBC_DIRICHLET] * 6, [0.0] * 6)] * 3)
        x = (np.arange(-pad, n + pad) + 0.5) * (L / n)
        sx, cx = np.sin(x), np.cos(x)
        u = sx[:, None, None] * (cx[:, None] * cx[None, :])[None, :, :] + 0.01 * rng.uniform(-1, 1, sh)
        v = -(cx[:, None, None] * (sx[:, None] * cx[None, :])[None, :, :]) + 0.01 * rng.uniform(-1, 1, sh)
        w = 0.01 * rng.uniform(-1, 1, sh)
        f = dict(vx=np.ascontiguousarray(u), vy=np.ascontiguousarray(v), vz=w, p=np.zeros(sh), t=np.zeros(sh))
    return g, mg, prm, f


def run_mixed_stream(server, reqs, nse_reqs, rhs, precision):
    """Submit Poisson requests and short Navier-Stokes runs interleaved, all
    at once; returns (records, Poisson solutions, NSE final fields) by tag."""
    recs, sols, outs = {}, {}, {}
    order = sorted([("p", r) for r in reqs] + [("n", r) for r in nse_reqs], key=lambda kr: kr[1]["tag"] % 7)
    for kind, r in order:
        if kind == "p":
            b = rhs.get(r)
            x = np.zeros_like(b)
            sols[r["tag"]] = x
            server.submit(request_grid(r), request_params(r, precision), b, x, r["tag"])
        else:
            g, mg, prm, f = nse_request_setup(r)
            outs[r["tag"]] = f
            server.submit_nse(g, mg, prm, r["steps"], f, r["tag"])
        for c in server.poll(wait_ms=0):
            recs[c["tag"]] = c
    server.drain()
    for c in server.poll(max_n=len(reqs) + len(nse_reqs) + 1):
        recs[c["tag"]] = c
    return recs, sols, outs
