"""The configurations at their decks' sizes against the FP64 CPU oracle: the
two time-stepping ones through fixtures (tests/golden/make_golden.py writes
them), the 512^3 Poisson solve against the oracle run in the test.

rbc256: 3-D Rayleigh-Benard convection at its full 256^3 size, the deck's
100 lsRK3 steps of the FP64 path.  Bar: per step the Jacobi counts, the
pressure solve's cycle count and residual and max|div| are equal, KE, NuVol
and NuWall agree to 1e-12 (sums over the grid in another order); after step
3 and after the last step every field's core is bit-identical (SHA-256 of its
bytes).  The same run with the mixed-precision pressure solve reports its
per-step and final-field differences from the FP64 run, bounded by
RBC_MIXED_HIST_RTOL and RBC_MIXED_FIELD_RTOL.

tgv_les64: the decaying Taylor-Green vortex with the spiral LES model at
64^3, 500 steps.  The sub-grid stress evaluates pow, exp, sin, acos and cos,
whose device results may differ from the host C library's in the last bit,
so the histories are held to a relative tolerance: KE and the sub-grid
diagnostics per step to LES_HIST_RTOL of their values, the dissipation
eps = -dKE/dt (centred differences of the KE history) to LES_EPS_RTOL of its
largest value, max|div| to LES_HIST_RTOL of its largest value.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import saras_amd as S

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LES_HIST_RTOL = 1e-8
LES_EPS_RTOL = 1e-6
RBC_MIXED_HIST_RTOL = 1e-4
RBC_MIXED_FIELD_RTOL = 1e-4
MIXED_512_RTOL = 1e-12


def load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        pytest.fail(f"fixture {name} is missing: run tests/golden/make_golden.py")
    with open(path) as fh:
        return json.load(fh)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def rbc256_run(G, precision):
    """Runs the fixture's Rayleigh-Benard deck with the given pressure-solve
    precision; returns the per-step records and the core of every field after
    the fixture's early step and after the last one."""
    c = G["config"]
    n, pad = c["n"], c["pad"]
    ph = S.nse_physics(5, ra=c["ra"], pr=c["pr"], rbc_type=2, g=(0.0, 0.0, -1.0))
    assert (ph["nu"], ph["kappa"]) == (c["nu"], c["kappa"])
    g = S.grid_desc(n, periodic=(1, 1, 0), gtype=(0, 0, S.GRID_TANH_DOUBLE), beta=(1.0, 1.0, 1.5), pad=pad)
    # the mixed solver checks the FP64 residual every 5 FP32 cycles: these
    # solves all run to the 20-cycle cap (their residual plateaus near 1e-4)
    mg = S.mg_params(bc=S.BC_NEUMANN, precision=precision, inner_cycles=5, max_refine=10, **c["mg"])
    prm = S.nse_params(dscheme=2, nlscheme=3, has_T=1, nu=ph["nu"], kappa=ph["kappa"], dt=c["dt"], cn_tol=1e-6,
                       buoy=ph["buoy"])
    s = S.Nse(g, mg, prm)
    sh = s.shape
    core = (slice(pad, -pad),) * 3
    z = S.level_metrics(g, 0, 2)["x"][1:-1]
    rng = np.random.default_rng(c["seed"])
    T = np.zeros(sh)
    T[core] = 1.0 - z[None, None, :]
    T[core] += rng.uniform(-1e-3, 1e-3, (n, n, n))
    zero = np.zeros(sh)
    s.set(vx=zero, vy=zero, vz=zero, p=np.ones(sh), t=T, pp=zero)
    del T
    early_step = c.get("early")
    recs, early = [], None
    for step in range(c["steps"]):
        r, st = s.step()
        assert r == S.OK, (precision, step + 1, r)
        d = s.diag()
        recs.append(dict(step=step + 1, jacobi_iters=st["jacobi_iters"], mg_cycles=st["mg"]["cycles"],
                         mg_residual=st["mg"]["residual"], ke=d["ke"], divmax=d["divmax"], nuvol=d["nuvol"],
                         nuwall=d["nuwall"]))
        if (step + 1) % 25 == 0:
            print(f"rbc256 precision {precision} step {step + 1}: ke {d['ke']:.15e}", flush=True)
        if step + 1 == early_step:
            early = {k: v[core].copy() for k, v in s.get().items() if k != "pp"}
    final = {k: v[core].copy() for k, v in s.get().items() if k != "pp"}
    s.close()
    return recs, early, final


@pytest.fixture(scope="module")
def rbc256():
    G = load("rbc256_steps.json")
    return G, rbc256_run(G, S.PREC_FP64)


def test_rbc256_steps_match_oracle_fixture(rbc256):
    G, (recs, early, final) = rbc256
    assert len(recs) == len(G["history"])
    for a, h in zip(recs, G["history"]):
        assert a["jacobi_iters"] == h["jacobi_iters"], (h, a)
        assert a["mg_cycles"] == h["mg_cycles"], (h, a)
        assert a["mg_residual"] == h["mg_residual"], (h, a)
        assert a["divmax"] == h["divmax"], (h, a)
        for k in ("ke", "nuvol", "nuwall"):
            assert abs(a[k] - h[k]) <= 1e-12 * abs(h[k]), (h["step"], k, a[k], h[k])
    if G.get("fields_early") is not None:
        for k, want in G["fields_early"].items():
            assert sha(early[k]) == want["sha256"], ("early", k, want)
    for k, want in G["fields"].items():
        got = final[k]
        assert sha(got) == want["sha256"], (k, float(np.abs(got).max()), want)


def field_diffs(a, b):
    """Per field: L2 (rms) and L-inf of a - b over the core relative to b's.
    The pressure (Neumann walls, periodic x and y) is defined up to a
    constant and is compared modulo its mean."""
    out = {}
    for k in b:
        x, y = a[k], b[k]
        if k == "p":
            x, y = x - x.mean(), y - y.mean()
        d = x - y
        out[k] = dict(l2_rel=float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(y * y))),
                      linf_rel=float(np.abs(d).max() / np.abs(y).max()))
    return out


def test_rbc256_mixed_pressure_history_against_fp64(rbc256):
    """The same run with the mixed-precision pressure solve.  Its solves stop
    at the same 20-cycle cap on the residual plateau as the FP64 ones, at a
    different point of the plateau, so the fields differ by what those
    residuals leave; the test reports the per-step and final differences from
    the FP64 run (which equals the oracle's bit for bit) and bounds them."""
    G, (recs64, _, final64) = rbc256
    recs, _, final = rbc256_run(G, S.PREC_MIXED)
    worst = {}
    for a, b in zip(recs, recs64):
        for k in ("ke", "nuvol", "nuwall"):
            worst[k] = max(worst.get(k, 0.0), abs(a[k] - b[k]) / abs(b[k]))
    worst["divmax"] = max(abs(a["divmax"] - b["divmax"]) for a, b in zip(recs, recs64)) / max(
        b["divmax"] for b in recs64)
    diffs = field_diffs(final, final64)
    print(json.dumps(dict(steps=len(recs), history_rel=worst, final_fields=diffs)), flush=True)
    for k in ("ke", "nuvol", "nuwall"):
        assert worst[k] <= RBC_MIXED_HIST_RTOL, (k, worst)
    for k, v in diffs.items():
        assert v["l2_rel"] <= RBC_MIXED_FIELD_RTOL, (k, diffs)


def test_tgv_les64_history_matches_oracle_fixture():
    G = load("tgv_les64_history.json")
    c = G["config"]
    n, pad, L = c["n"], c["pad"], c["length"]
    assert S.nse_physics(2, re=c["re"])["nu"] == c["nu"]
    g = S.grid_desc(n, periodic=(1, 1, 1), length=(L, L, L), pad=pad)
    mg = S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64, **c["mg"])
    prm = S.nse_params(dscheme=2, nlscheme=3, has_T=0, nu=c["nu"], kappa=c["nu"], dt=c["dt"], cn_tol=1e-6, les=1,
                       vbc=[([S.FBC_DIRICHLET] * 6, [0.0] * 6)] * 3)
    s = S.Nse(g, mg, prm)
    sh = s.shape
    x = (np.arange(-pad, n + pad) + 0.5) * (L / n)
    rng = np.random.default_rng(c["seed"])
    sx, cx = np.sin(x), np.cos(x)
    u = sx[:, None, None] * (cx[:, None] * cx[None, :])[None, :, :] + c["noise"] * rng.uniform(-1, 1, sh)
    v = -(cx[:, None, None] * (sx[:, None] * cx[None, :])[None, :, :]) + c["noise"] * rng.uniform(-1, 1, sh)
    w = c["noise"] * rng.uniform(-1, 1, sh)
    zero = np.zeros(sh)
    s.set(vx=u, vy=v, vz=w, p=zero, t=zero, pp=zero)
    vol = L ** 3
    hist = []
    for step in range(c["steps"]):
        r, st = s.step()
        assert r == S.OK, step
        d = s.diag()
        sg, t = s.sgs()
        hist.append(dict(time=t, ke=d["ke"], divmax=d["divmax"], sgs_ke=sg["sgke"] / vol,
                         sgs_dissipation=sg["disp"] / vol, nu_sgs=sg["nusg"] / vol))
        if (step + 1) % 100 == 0:
            print(f"tgv_les64 step {step + 1}: ke {d['ke']:.12f}", flush=True)
    s.close()
    ref = G["history"]
    assert len(hist) == len(ref)
    worst = {}
    for a, b in zip(hist, ref):
        assert a["time"] == b["time"]
        for k in ("ke", "sgs_ke", "sgs_dissipation", "nu_sgs"):
            rel = abs(a[k] - b[k]) / max(abs(b[k]), 1e-300)
            worst[k] = max(worst.get(k, 0.0), rel)
    dmax = max(b["divmax"] for b in ref)
    worst["divmax"] = max(abs(a["divmax"] - b["divmax"]) for a, b in zip(hist, ref)) / dmax
    # dissipation eps = -dKE/dt by centred differences of the two histories
    ke_g = np.array([h["ke"] for h in hist])
    ke_o = np.array([h["ke"] for h in ref])
    dt = c["dt"]
    eps_g = -(ke_g[2:] - ke_g[:-2]) / (2 * dt)
    eps_o = -(ke_o[2:] - ke_o[:-2]) / (2 * dt)
    worst["eps"] = float(np.abs(eps_g - eps_o).max() / np.abs(eps_o).max())
    print(json.dumps(worst), flush=True)
    for k in ("ke", "sgs_ke", "sgs_dissipation", "nu_sgs", "divmax"):
        assert worst[k] <= LES_HIST_RTOL, (k, worst)
    assert worst["eps"] <= LES_EPS_RTOL, worst


def test_config3_512_periodic_fourier_solves_match_oracle():
    """Config 3 at its deck's size: the 512^3 periodic Poisson problem with
    the 64-mode Fourier right-hand side (bench.fourier_rhs), V(2,2), solved to
    a relative L-inf residual of 1e-10.  The FP64 path equals the oracle's
    FP64 solve bit for bit (cycle count, residual, every value of the returned
    array); the mixed-precision solve reaches the tolerance by the oracle's
    own computeError and differs from the oracle's solution by at most
    MIXED_512_RTOL of its largest value (modulo the mean: the periodic
    problem's solution is defined up to a constant)."""
    import oracle as O
    from bench import fourier_rhs

    n, tol = 512, 1e-10
    kw = dict(vc_count=100, vc_tol=tol, mg_tol=1e-2, pre=2, post=2)
    b = fourier_rhs(n)
    mo = O.Multigrid(O.grid_desc((n, n, n), periodic=(1, 1, 1)), O.mg_params(bc=O.BC_NEUMANN, **kw))
    xo, so = mo.solve(np.zeros_like(b), b, 1)
    print(f"oracle 512^3: {so['cycles']} cycles, residual {so['residual']:.3e}", flush=True)
    g = S.grid_desc((n, n, n), periodic=(1, 1, 1))
    ms = S.Multigrid(g, S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64, **kw))
    xs, ss, status = ms.solve(np.zeros_like(b), b)
    ms.close()
    assert status == S.OK and ss["cycles"] == so["cycles"], (ss, so)
    assert ss["residual"] == so["residual"], (ss, so)
    assert np.array_equal(xs, xo), float(np.abs(xs - xo).max())
    del xs
    mm = S.Multigrid(g, S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_MIXED, inner_cycles=4, max_refine=40, **kw))
    xm, sm, status = mm.solve(np.zeros_like(b), b)
    mm.close()
    assert status == S.OK and sm["converged"], sm
    mo.set_level(0, xm, b)
    mo.impose_bc(0, 0)
    r = mo.error(0)
    assert r < 1.5 * tol, (r, sm)
    c_m, c_o = xm[1:-1, 1:-1, 1:-1], xo[1:-1, 1:-1, 1:-1]
    d = np.abs((c_m - c_m.mean()) - (c_o - c_o.mean())).max() / np.abs(c_o - c_o.mean()).max()
    print(json.dumps(dict(fp32_cycles=sm["cycles"], refinements=sm["refinements"], oracle_residual=r,
                          mixed_vs_oracle_linf_rel=float(d))), flush=True)
    assert d <= MIXED_512_RTOL, d
