"""GPU parity of the FP64 Navier-Stokes/Boussinesq time step against the FP64
CPU oracle.

Bar: bit-exact for every field value the reference defines (core cells and
the ghost cells its imposeBCs/syncFaces write: periodic pads and the first
wall layer), for the stage RHS and non-linear terms, the Jacobi iteration
counts and the pressure solves.  Diagnostics that sum over the grid reduce in
a different order and are held to 1e-12 relative; the maximum |div V| is
order-independent and bit-exact.
"""
import numpy as np
import pytest

import oracle as O
import saras_amd as S

pytestmark = pytest.mark.gpu


def mg_kw(**kw):
    d = dict(vc_count=8, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2, pre=3, post=6)
    d.update(kw)
    return d


def pair(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.5), pad=2, mg=None, gpu_kw=None, **kw):
    """The oracle and the library on the same deck; gpu_kw: further
    parameters of the library's handle only."""
    mg = mg or mg_kw()
    go = O.grid_desc(n, periodic=periodic, gtype=gtype, beta=beta, pad=pad)
    gs = S.grid_desc(n, periodic=periodic, gtype=gtype, beta=beta, pad=pad)
    so = O.Nse(O.nse_desc(go, mg=O.mg_params(bc=O.BC_NEUMANN, **mg), **kw))
    ss = S.Nse(gs, S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64, **mg), S.nse_params(**kw, **(gpu_kw or {})))
    return so, ss


def defined_mask(shape, n, pad, periodic):
    """Cells with at most one index outside the core, that index inside the
    periodic pads or on the first wall layer."""
    m = np.zeros(shape, dtype=bool)
    idx = np.indices(shape) - pad
    out = [(idx[a] < 0) | (idx[a] >= n[a]) for a in range(3)]
    nout = sum(o.astype(int) for o in out)
    m |= nout == 0
    for a in range(3):
        ok = out[a] & (nout == 1)
        if not periodic[a]:
            ok &= (idx[a] >= -1) & (idx[a] <= n[a])
        m |= ok
    return m


def rbc_init(shape, n, pad, z, seed=2, vel=0.0):
    rng = np.random.default_rng(seed)
    T = np.zeros(shape)
    T[:] = 1.0 - z[None, None, :]
    c = (slice(pad, -pad),) * 3
    T[c] += rng.uniform(-1e-3, 1e-3, n)
    V = [np.zeros(shape) for _ in range(3)]
    if vel:
        for v in V:
            v[c] = vel * rng.standard_normal(n)
    return dict(vx=V[0], vy=V[1], vz=V[2], p=np.ones(shape), t=T, pp=np.zeros(shape))


def assert_fields_equal(so, ss, names=("vx", "vy", "vz", "p", "t", "pp"), periodic=(1, 1, 0)):
    fo = so.get()
    fs = ss.get()
    g = so.d.g
    n = tuple(g.n[a] for a in range(3))
    m = defined_mask(so.shape, n, g.pad, periodic)
    for k in names:
        a, b = fo[k][m], fs[k][m]
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, (k, bad.size, np.abs(a - b).max())


GRIDS = [
    dict(n=(16, 16, 16), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.5)),
    dict(n=(12, 12, 12), periodic=(0, 0, 0), gtype=(2, 1, 2), beta=(1.3, 0.9, 1.2)),
    dict(n=(16, 8, 12), periodic=(1, 0, 1), gtype=(0, 2, 1), beta=(1.0, 1.1, 0.7)),
]
GIDS = ["ppn-16-z1.5", "nnn-12-stretched", "pnp-16x8x12"]


@pytest.mark.parametrize("gk", GRIDS, ids=GIDS)
@pytest.mark.parametrize("dscheme", [1, 2])
def test_field_ops_bitexact(gk, dscheme):
    pad = 2
    so, ss = pair(pad=pad, dscheme=dscheme, nlscheme=3, **gk)
    rng = np.random.default_rng(11)
    F = rng.standard_normal(so.shape)
    V = dict(vx=rng.standard_normal(so.shape), vy=rng.standard_normal(so.shape), vz=rng.standard_normal(so.shape),
             p=np.zeros(so.shape), t=np.zeros(so.shape), pp=np.zeros(so.shape))
    so.set(**V)
    ss.set(**V)
    c = (slice(pad, -pad),) * 3
    for op in range(8):
        a = so.op(op, F)[c]
        b = ss.op(op, F)[c]
        assert np.array_equal(a, b), (op, np.abs(a - b).max())


@pytest.mark.parametrize("gk", GRIDS, ids=GIDS)
@pytest.mark.parametrize("dscheme,nlscheme", [(1, 1), (2, 1), (1, 2), (2, 2), (2, 3)])
def test_vector_field_ops_bitexact(gk, dscheme, nlscheme):
    """vfield::divergence, sfield::gradient, vfield::computeDiff and
    vfield::computeNLin (central, hybrid upwind, Morinishi) over the core of
    caller arrays with arbitrary ghost values, accumulating into non-zero
    outputs where the reference accumulates."""
    pad = 2
    so, ss = pair(pad=pad, dscheme=dscheme, nlscheme=nlscheme, nu=2e-3, **gk)
    rng = np.random.default_rng(17)
    f = [rng.standard_normal(so.shape) for _ in range(3)]
    acc = [rng.standard_normal(so.shape) for _ in range(3)]
    c = (slice(pad, -pad),) * 3
    for op in range(4):
        out = acc if op >= 2 else None
        a = so.vec_op(op, *f, out=out)
        b = ss.vec_op(op, *f, out=out)
        for q in range(1 if op == 0 else 3):
            assert np.array_equal(a[q][c], b[q][c]), (op, q, np.abs(a[q][c] - b[q][c]).max())


@pytest.mark.parametrize("nlscheme", [1, 3])
@pytest.mark.parametrize("gk", GRIDS[:2], ids=GIDS[:2])
def test_stage_rhs_bitexact(gk, nlscheme):
    pad = 2
    periodic = gk["periodic"]
    so, ss = pair(pad=pad, dscheme=2, nlscheme=nlscheme, nu=3e-3, kappa=2e-3, dt=1e-3, **gk)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, tuple(gk["n"]), pad, z, vel=0.1)
    f["p"] = np.random.default_rng(4).standard_normal(so.shape)
    so.set(**f)
    ss.set(**f)
    assert_fields_equal(so, ss, periodic=periodic)
    c = (slice(pad, -pad),) * 3
    for stage in (0, 1, 2):
        ro, rs = so.rhs(stage), ss.rhs(stage)
        for k in ro:
            assert np.array_equal(ro[k][c], rs[k][c]), (stage, k, np.abs(ro[k][c] - rs[k][c]).max())


def test_rbc_steps_bitexact():
    """16^3 Rayleigh-Benard slice of the 256^3 run (PPN, z-tanh 1.5, Ra 1e7,
    Pr 1, dScheme 2, Morinishi, V(3,6) pressure solves): three full steps."""
    pad = 2
    nu = kappa = (1.0 / 1e7) ** 0.5
    so, ss = pair((16, 16, 16), pad=pad, nu=nu, kappa=kappa, dt=1e-3, cn_tol=1e-6)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, (16, 16, 16), pad, z)
    so.set(**f)
    ss.set(**f)
    for step in range(3):
        ro = so.step()
        rs, st = ss.step()
        assert ro == rs == 0
        assert so.jacobi_iters() == st["jacobi_iters"], step
        assert so.mg_stats()["cycles"] == st["mg"]["cycles"]
        assert so.mg_stats()["residual"] == st["mg"]["residual"]
        assert_fields_equal(so, ss)
    do, ds = so.diag(), ss.diag()
    assert do["divmax"] == ds["divmax"]
    for k in ("ke", "nuvol", "nuwall"):
        assert abs(do[k] - ds[k]) <= 1e-12 * abs(do[k]), k
    # the time-series quantities (tseries.cc:295-490) and probe values (probes.cc:185-205)
    td = ss.tsdata()
    assert td["divmax"] == do["divmax"] and abs(td["ke"] - do["ke"]) <= 1e-12 * do["ke"]
    te = so.thermal_energy()
    assert te > 0 and abs(td["te"] - te) <= 1e-12 * te
    pts = np.array([(0, 0, 0), (15, 7, 3), (4, 15, 15)])
    pv = ss.probe(pts)
    fo = so.get()
    for q, (i, j, k) in enumerate(pts):
        for f, name in enumerate(("vx", "vy", "vz", "p", "t")):
            assert pv[q, f] == fo[name][i + pad, j + pad, k + pad], (q, name)
    with pytest.raises(S.SarasError):
        ss.probe([(16, 0, 0)])


def test_rbc_rows_of_192_cells_bitexact():
    """Rows of 192 cells run in blocks of 192 threads (no idle lanes, as at
    384^3): two steps of a periodic 8 x 8 x 192 box with a seeded velocity
    field (Morinishi terms, buoyancy), bit-exact."""
    pad = 2
    n = (8, 8, 192)
    nu = kappa = 1e-3
    so, ss = pair(n, periodic=(1, 1, 1), gtype=(0, 0, 0), pad=pad, nu=nu, kappa=kappa, dt=1e-3, cn_tol=1e-6)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, n, pad, z, vel=0.1)
    so.set(**f)
    ss.set(**f)
    for step in range(2):
        ro = so.step()
        rs, st = ss.step()
        assert ro == rs == 0
        assert so.jacobi_iters() == st["jacobi_iters"], step
        assert so.mg_stats()["residual"] == st["mg"]["residual"]
        assert_fields_equal(so, ss, periodic=(1, 1, 1))


def test_walls_everywhere_mixed_bcs_bitexact():
    """Closed box, every wall kind: Dirichlet and Neumann values on V and T,
    a velocity that is not divergence free, two steps."""
    pad = 2
    vbc = [([S.FBC_DIRICHLET, S.FBC_NEUMANN, S.FBC_DIRICHLET, S.FBC_DIRICHLET, S.FBC_NEUMANN, S.FBC_DIRICHLET],
            [0.1, -0.2, 0.0, 0.3, 0.05, -0.1])] * 3
    tbc = ([S.FBC_NEUMANN, S.FBC_DIRICHLET, S.FBC_NEUMANN, S.FBC_NEUMANN, S.FBC_DIRICHLET, S.FBC_DIRICHLET],
           [0.2, 0.4, 0.0, -0.3, 1.0, 0.0])
    kw = dict(periodic=(0, 0, 0), gtype=(2, 1, 2), beta=(1.2, 0.8, 1.5), pad=pad, nu=4e-3, kappa=3e-3, dt=2e-3,
              cn_tol=1e-9, vbc=vbc, tbc=tbc, buoy=(0.2, 0.0, -0.9))
    so, ss = pair((12, 12, 12), **kw)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, (12, 12, 12), pad, z, vel=0.05, seed=7)
    so.set(**f)
    ss.set(**f)
    assert_fields_equal(so, ss, periodic=(0, 0, 0))
    for _ in range(2):
        assert so.step() == 0
        rs, st = ss.step()
        assert rs == 0
        assert so.jacobi_iters() == st["jacobi_iters"]
        assert_fields_equal(so, ss, periodic=(0, 0, 0))


def test_taylor_green_hydro_bitexact():
    """Periodic box, no scalar (lsRK3_d3::timeAdvance(V, P)), dScheme 2."""
    pad = 2
    so, ss = pair((16, 16, 16), periodic=(1, 1, 1), gtype=(0, 0, 0), pad=pad, has_T=0, nu=1e-2, dt=2e-3,
                  cn_tol=1e-10, mg=mg_kw(vc_count=10, vc_tol=1e-10, pre=3, post=3))
    x = (np.arange(-pad, 16 + pad) + 0.5) / 16
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    tp = 2 * np.pi
    f = dict(vx=np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z),
             vy=-np.cos(tp * X) * np.sin(tp * Y) * np.cos(tp * Z),
             vz=np.zeros(so.shape), p=np.zeros(so.shape), t=np.zeros(so.shape), pp=np.zeros(so.shape))
    so.set(**f)
    ss.set(**f)
    for _ in range(2):
        assert so.step() == 0
        assert ss.step()[0] == 0
    assert_fields_equal(so, ss, names=("vx", "vy", "vz", "p", "pp"), periodic=(1, 1, 1))


def test_jacobi_cap_reports_divergence():
    """cn_tol = 0 can never be met: both sides stop after maxIterations and
    report the reference's abort as JACOBI_DIVERGED (predictors.cc:133-141)."""
    pad = 2
    so, ss = pair((4, 4, 4), pad=pad, cn_tol=0.0, nu=1e-2, kappa=1e-2, dt=1e-3)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, (4, 4, 4), pad, z)
    so.set(**f)
    ss.set(**f)
    assert so.stage(0) == 2
    rs, st = ss.stage(0)
    assert rs == S.JACOBI_DIVERGED
    assert st["jacobi_iters"][0] == -1


def test_rbc_256_one_step_properties():
    """256^3 Rayleigh-Benard step at full size: finite, the temperature stays
    inside its boundary values up to the seeded perturbation, the pressure
    solve ran its cycles, and a second handle fed the same inputs reproduces
    the step bit for bit."""
    n, pad = 256, 2
    g = S.grid_desc(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), pad=pad)
    nu = (1.0 / 1e7) ** 0.5
    prm = S.nse_params(nu=nu, kappa=nu, dt=1e-3, cn_tol=1e-6)
    mg = S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64, **mg_kw(vc_count=20))
    zg = O.grid_desc(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), pad=pad)
    z = O.field_metrics(zg, 2)["x"]
    f = rbc_init((n + 2 * pad,) * 3, (n, n, n), pad, z)
    out = []
    for _ in range(2):
        s = S.Nse(g, mg, prm)
        s.set(**f)
        r, st = s.step()
        assert r == S.OK
        assert st["mg"]["cycles"] >= 1
        got = s.get()
        out.append(got)
        c = (slice(pad, -pad),) * 3
        assert np.isfinite(got["vz"][c]).all() and np.isfinite(got["p"][c]).all()
        assert got["t"][c].min() > -2e-3 and got["t"][c].max() < 1 + 2e-3
        dg = s.diag()
        assert np.isfinite(dg["ke"]) and dg["ke"] < 1e-6
        s.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_device_pointers_roundtrip():
    pad = 2
    so, ss = pair((16, 16, 16), pad=pad)
    rng = np.random.default_rng(1)
    arrs = {k: rng.standard_normal(ss.shape) for k in ss.NAMES}
    devs = {k: S.DeviceArray.from_host(v) for k, v in arrs.items()}
    S.check(S.lib().saras_nse_set_fields(ss.h, *[devs[k].ptr for k in ss.NAMES], S.PTR_DEVICE))
    outs = {k: S.DeviceArray.from_host(np.zeros(ss.shape)) for k in ss.NAMES}
    S.check(S.lib().saras_nse_get_fields(ss.h, *[outs[k].ptr for k in ss.NAMES], S.PTR_DEVICE))
    c = (slice(pad, -pad),) * 3
    for k in ss.NAMES:
        assert np.array_equal(outs[k].to_host(ss.shape)[c], arrs[k][c]), k


@pytest.mark.parametrize("nlscheme", [1, 3])
def test_euler_cn_bitexact(nlscheme):
    """eulerCN_d3 (Integration Scheme 1): the explicit RHS and three full
    steps of a closed-box Boussinesq run agree bit for bit."""
    pad = 2
    kw = dict(periodic=(1, 0, 0), gtype=(0, 1, 2), beta=(1.0, 0.9, 1.5), pad=pad, nu=4e-3, kappa=3e-3, dt=2e-3,
              cn_tol=1e-9, nlscheme=nlscheme, scheme=1, buoy=(0.0, 0.3, -1.0))
    so, ss = pair((12, 12, 12), **kw)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, (12, 12, 12), pad, z, vel=0.05, seed=8)
    so.set(**f)
    ss.set(**f)
    c = (slice(pad, -pad),) * 3
    ro, rs = so.rhs(0), ss.rhs(0)
    for k in ("rx", "ry", "rz", "rt"):
        assert np.array_equal(ro[k][c], rs[k][c]), k
    so.set(**f)
    ss.set(**f)
    for _ in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r == 0
        assert so.jacobi_iters() == st["jacobi_iters"]
        assert_fields_equal(so, ss, periodic=(1, 0, 0))


@pytest.mark.parametrize("dscheme", [1, 2])
def test_upwind_scheme_bitexact(dscheme):
    """Hybrid upwind non-linear term (Non-Linear Term 2): Peclet numbers on
    both sides of the limit, both biasing directions; the scalar operator,
    the stage RHS and two RK3 steps agree bit for bit."""
    pad = 2
    kw = dict(periodic=(1, 0, 0), gtype=(0, 2, 1), beta=(1.0, 1.2, 0.9), pad=pad, nu=0.03, kappa=0.02, dt=1e-3,
              cn_tol=1e-9, nlscheme=2, dscheme=dscheme, central_bias=0.7, pe_limit=2.0)
    so, ss = pair((12, 12, 12), **kw)
    rng = np.random.default_rng(19)
    sh = so.shape
    f = dict(vx=rng.standard_normal(sh), vy=rng.standard_normal(sh), vz=rng.standard_normal(sh),
             p=rng.standard_normal(sh), t=rng.standard_normal(sh), pp=np.zeros(sh))
    so.set(**f)
    ss.set(**f)
    c = (slice(pad, -pad),) * 3
    F = rng.standard_normal(sh)
    assert np.array_equal(so.op(7, F)[c], ss.op(7, F)[c])
    ro, rs = so.rhs(0), ss.rhs(0)
    for k in ro:
        assert np.array_equal(ro[k][c], rs[k][c]), k
    so.set(**f)
    ss.set(**f)
    for _ in range(2):
        assert so.step() == 0
        assert ss.step()[0] == 0
        assert_fields_equal(so, ss, periodic=(1, 0, 0))


def test_cfl_time_step_matches_oracle():
    pad = 2
    so, ss = pair((16, 16, 16), pad=pad)
    rng = np.random.default_rng(20)
    sh = so.shape
    f = dict(vx=rng.standard_normal(sh), vy=rng.standard_normal(sh), vz=rng.standard_normal(sh),
             p=np.zeros(sh), t=np.zeros(sh), pp=np.zeros(sh))
    so.set(**f)
    ss.set(**f)
    assert so.cfl(0.1, 1.0) == ss.cfl(0.1, 1.0)
    assert so.cfl(0.1, 1e-6) == ss.cfl(0.1, 1e-6) == 1e-6
    # the next step uses the capped dt on both sides
    assert so.step() == 0
    assert ss.step()[0] == 0
    assert_fields_equal(so, ss)


def tg_les_pair(n=16, seed=5, noise=0.05, scheme=0, sf32=0):
    """Taylor-Green vortex with seeded noise in a periodic unit box, the
    spiral LES model on (lsRK3_d3 or eulerCN_d3), nu = 1/1600; sf32: the
    library's structure-function integral in single precision."""
    pad = 2
    so, ss = pair((n, n, n), periodic=(1, 1, 1), gtype=(0, 0, 0), pad=pad, has_T=0, nu=1.0 / 1600, dt=1e-3,
                  cn_tol=1e-10, scheme=scheme, les=1, mg=mg_kw(vc_count=10, vc_tol=1e-10, pre=3, post=3),
                  gpu_kw={"les_sf32": sf32})
    x = (np.arange(-pad, n + pad) + 0.5) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    tp = 2 * np.pi
    rng = np.random.default_rng(seed)
    sh = so.shape
    f = dict(vx=np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z) + noise * rng.uniform(-1, 1, sh),
             vy=-np.cos(tp * X) * np.sin(tp * Y) * np.cos(tp * Z) + noise * rng.uniform(-1, 1, sh),
             vz=noise * rng.uniform(-1, 1, sh), p=np.zeros(sh), t=np.zeros(sh), pp=np.zeros(sh))
    so.set(**f)
    ss.set(**f)
    return so, ss, pad


# The sub-grid stress evaluates pow, exp, sin, acos and cos, whose device
# implementations may differ from the host C library in the last bit; the
# LES terms are therefore held to a relative tolerance (1e-9 of the largest
# value) instead of bit equality.
LES_RTOL = 1e-9


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler_cn"])
def test_les_spiral_matches_oracle(scheme):
    """Steps before solTime > 5 tStp are bit-exact (no model); then the stage
    RHS with the spiral sub-grid stress divergence, its diagnostics and two
    more steps agree with the oracle to LES_RTOL."""
    so, ss, pad = tg_les_pair(scheme=scheme)
    for step in range(6):
        assert so.step() == 0
        assert ss.step()[0] == 0
    assert_fields_equal(so, ss, names=("vx", "vy", "vz", "p", "pp"), periodic=(1, 1, 1))
    assert so.time == ss.sgs()[1]
    c = (slice(pad, -pad),) * 3
    ro, rs = so.rhs(0), ss.rhs(0)
    keys = ("rx", "ry", "rz") + (("hx", "hy", "hz") if scheme == 0 else ())
    for k in keys:
        a, b = ro[k][c], rs[k][c]
        assert np.abs(a - b).max() <= LES_RTOL * np.abs(a).max(), (k, np.abs(a - b).max())
    do = so.sgs_diag()
    dg, t = ss.sgs()
    for q, k in enumerate(("sgke", "disp", "nusg")):
        assert abs(dg[k] - do[q]) <= LES_RTOL * abs(do[q]), (k, dg[k], do[q])
    f0 = so.get()
    ss.set(**f0)   # restart both from the same state for the steps
    for _ in range(2):
        assert so.step() == 0
        assert ss.step()[0] == 0
    fo, fs = so.get(), ss.get()
    for k in ("vx", "vy", "vz", "p"):
        a, b = fo[k][c], fs[k][c]
        assert np.abs(a - b).max() <= LES_RTOL * np.abs(a).max(), (k, np.abs(a - b).max())


# With the structure-function integral in single precision the sub-grid
# energy K carries about 1e-7 of relative rounding; its diagnostics and the
# terms it enters are held to 1e-6.
LES_SF32_RTOL = 1e-6


def test_les_sf32_close_to_oracle():
    """les_sf32 = 1: the sub-grid diagnostics (sums of |K|, of the
    dissipation and of nu_t, which only the model sets) and the stage RHS agree
    with the oracle's double-precision model to LES_SF32_RTOL, and the
    diagnostics differ from the double-precision kernel's (the variant ran)."""
    so, ss, pad = tg_les_pair(sf32=1)
    _, sd, _ = tg_les_pair(sf32=0)
    for step in range(6):
        assert so.step() == 0
        assert ss.step()[0] == 0
        assert sd.step()[0] == 0
    c = (slice(pad, -pad),) * 3
    ro, rs = so.rhs(0), ss.rhs(0)
    for k in ("rx", "ry", "rz", "hx", "hy", "hz"):
        a, b = ro[k][c], rs[k][c]
        assert np.abs(a - b).max() <= LES_SF32_RTOL * np.abs(a).max(), (k, np.abs(a - b).max())
    do = so.sgs_diag()
    dg, _ = ss.sgs()
    dd, _ = sd.sgs()
    for q, k in enumerate(("sgke", "disp", "nusg")):
        assert abs(dg[k] - do[q]) <= LES_SF32_RTOL * abs(do[q]), (k, dg[k], do[q])
    assert any(dg[k] != dd[k] for k in ("sgke", "disp", "nusg")), (dg, dd)
    ss.close()
    sd.close()


def test_les_needs_periodic_box():
    with pytest.raises(S.SarasError) as e:
        pair((16, 16, 16), periodic=(1, 1, 0), les=1)
    assert e.value.status == S.BAD_SHAPE


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler-cn"])
def test_channel_constant_pgrad_bitexact(scheme):
    """Force type 5 (constantPGrad, force.h:148-153) on the channelTest wiring:
    x and y periodic, no-slip z walls, z tanh-stretched with beta 1, dScheme 1,
    central non-linear term, nu = 0.1, a perturbed flow; three steps of both
    integration schemes agree bit for bit."""
    pad = 1
    n = (12, 8, 16)
    so, ss = pair(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.0), pad=pad, has_T=0, dscheme=1,
                  nlscheme=1, nu=0.1, kappa=0.1, dt=2e-3, cn_tol=1e-9, scheme=scheme, mean_pgrad=1.0)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, n, pad, z, vel=0.3, seed=11)
    f["t"] = np.zeros(so.shape)
    so.set(**f)
    ss.set(**f)
    for _ in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r == 0
        assert so.jacobi_iters() == st["jacobi_iters"]
        assert_fields_equal(so, ss, names=("vx", "vy", "vz", "p", "pp"), periodic=(1, 1, 0))


def run_channel(gtype, nz, dt, steps, nu=0.1, G=1.0, L=2.0):
    n, pad = (4, 4, nz), 1
    g = S.grid_desc(n, length=(4.0, 2.0, L), periodic=(1, 1, 0), gtype=(0, 0, gtype), beta=(1.0, 1.0, 1.0), pad=pad)
    mg = S.mg_params(vc_count=5, vc_tol=1e-12, solve_flag=0, pre=2, post=2, bc=S.BC_NEUMANN, precision=S.PREC_FP64)
    s = S.Nse(g, mg, S.nse_params(dscheme=1, nlscheme=1, has_T=0, nu=nu, kappa=nu, dt=dt, cn_tol=1e-13, scheme=1,
                                  mean_pgrad=G))
    zero = np.zeros(s.shape)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=zero, pp=zero)
    for _ in range(steps):
        r, _st = s.step()
        assert r == 0
    f = s.get()
    s.close()
    z = S.level_metrics(g, 0, 2)["x"][1:-1]
    c = slice(pad, -pad)
    ux = f["vx"][c, c, c]
    assert np.abs(ux - ux[:1, :1, :]).max() == 0      # every column alike
    assert np.abs(f["vz"][c, c, c]).max() == 0
    return z, ux[0, 0]


def test_channel_poiseuille_steady():
    """The steady channel of the channelTest deck (Re 10, mean pressure
    gradient 1, Lz = 2, Euler-CN) on the GPU.  Uniform z: the discrete steady
    state is the parabola raised by G h^2 / (8 nu),
    G/(2 nu) (z (L - z) + h^2/4), to 1e-9 of its peak.  The deck's tanh z grid
    (beta 1): checkChannel.py's analytic profile
    (tau_w delta / 2 nu)(z/delta)(2 - z/delta) within 0.3 % of its peak (the
    oracle's steady state on that grid is 0.14 % away)."""
    nu, G, L = 0.1, 1.0, 2.0
    z, u = run_channel(0, 32, 0.1, 1100)
    h = L / 32
    ue = G / (2 * nu) * (z * (L - z) + h * h / 4)
    assert np.abs(u - ue).max() < 1e-9 * ue.max()
    z, u = run_channel(2, 32, 0.05, 1500)
    delta = L / 2
    ua = (delta * G) * delta / (2 * nu) * (z / delta) * (2 - z / delta)
    assert np.abs(u - ua).max() < 3e-3 * ua.max()


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler-cn"])
def test_plume_hot_plate_bitexact(scheme):
    """The plume problem's walls (probType 9): hotPlate at the bottom
    (hotPlate.cc), plumeOut at the top (plumeOut.cc) fed by the outflow Vz
    (zero gradient at the top, on the x walls), no-slip elsewhere; a stretched
    12^3 box with a flow whose top Vz has both signs.  Three steps agree bit
    for bit, ghost layers included."""
    pad = 2
    vz_t = [O.FBC_NEUMANN, O.FBC_NEUMANN, O.FBC_DIRICHLET, O.FBC_DIRICHLET, O.FBC_DIRICHLET, O.FBC_NEUMANN]
    vbc = [([O.FBC_DIRICHLET] * 6, [0.0] * 6)] * 2 + [(vz_t, [0.0] * 6)]
    tbc = ([O.FBC_NEUMANN] * 4 + [O.FBC_HOT_PLATE, O.FBC_PLUME_OUT], [0.0] * 6)
    n = (12, 12, 12)
    so, ss = pair(n, periodic=(0, 0, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.2), pad=pad, dscheme=2, nlscheme=1,
                  nu=5e-3, kappa=5e-3, dt=2e-3, cn_tol=1e-9, scheme=scheme, vbc=vbc, tbc=tbc, plate_radius=0.3)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, n, pad, z, vel=0.3, seed=17)
    so.set(**f)
    ss.set(**f)
    top = ss.get()["vz"][pad:-pad, pad:-pad, -pad]
    assert 0 < (top > 0).sum() < n[0] * n[1]
    assert_fields_equal(so, ss, periodic=(0, 0, 0))
    for _ in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r == 0
        assert so.jacobi_iters() == st["jacobi_iters"]
        assert_fields_equal(so, ss, periodic=(0, 0, 0))
    # TimeSeries.dat's wall Nusselt number of an open top: the bottom wall alone
    td, dg = ss.tsdata(), so.diag()
    assert abs(td["nuwall"] - dg["nuwall"]) <= 1e-12 * abs(dg["nuwall"])


def test_special_walls_rejected_where_the_reference_has_none():
    g = S.grid_desc((8, 8, 8), periodic=(0, 0, 0), pad=1)
    mg = S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64)
    bad = [dict(tbc=([S.FBC_NEUMANN] * 4 + [S.FBC_PLUME_OUT, S.FBC_NEUMANN], [0.0] * 6)),   # plume at the bottom
           dict(tbc=([S.FBC_NEUMANN] * 5 + [S.FBC_HOT_PLATE], [0.0] * 6)),                  # plate at the top
           dict(vbc=[([S.FBC_HOT_PLATE] * 6, [0.0] * 6)] * 3),                             # on velocity
           dict(has_T=0, tbc=([S.FBC_NEUMANN] * 4 + [S.FBC_HOT_PLATE, S.FBC_NEUMANN], [0.0] * 6))]
    for kw in bad:
        args = dict(dscheme=1, nlscheme=1, has_T=1, nu=1e-2, kappa=1e-2, dt=1e-3)
        args.update(kw)
        with pytest.raises(S.SarasError):
            S.Nse(g, mg, S.nse_params(**args))


def test_morinishi_second_order_steps_bitexact():
    """Morinishi's term with dScheme 1 (firstOrder: second-order differences
    everywhere, vfield.cc:289-462, sfield.cc:200-311) in a Boussinesq box with
    z walls: three RK3 steps agree bit for bit."""
    pad = 1
    n = (12, 8, 16)
    so, ss = pair(n, pad=pad, dscheme=1, nlscheme=3, nu=4e-3, kappa=3e-3, dt=2e-3, cn_tol=1e-9)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, n, pad, z, vel=0.1, seed=23)
    so.set(**f)
    ss.set(**f)
    for _ in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r == 0
        assert so.jacobi_iters() == st["jacobi_iters"]
        assert_fields_equal(so, ss)


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler-cn"])
def test_inflow_null_pressure_wall_bitexact(scheme):
    """The inflow problem (probType 3) with the drivers' wiring: Vx = 1 on
    x-lo, zero-gradient outflow on x-hi, no-slip y and z walls, P null on
    x-lo (nullBC, boundary.h:205-212: its ghosts keep the caller's values plus
    every correction's ghost).  The caller's P carries random ghosts; three
    steps on a stretched 16 x 8 x 8 box agree bit for bit, the x-lo ghosts of
    P included."""
    pad = 2
    w = S.nse_wiring(3)
    assert w["pbc"][0][0] == S.FBC_NULL
    n = (16, 8, 8)
    so, ss = pair(n, periodic=(0, 0, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.2), pad=pad, dscheme=2, nlscheme=3,
                  has_T=0, nu=2e-2, dt=2e-3, cn_tol=1e-10, scheme=scheme, mg=mg_kw(vc_count=6, vc_tol=1e-10),
                  **w)
    rng = np.random.default_rng(31)
    c = (slice(pad, -pad),) * 3
    f = {k: np.zeros(so.shape) for k in ("vx", "vy", "vz", "p", "t", "pp")}
    f["vx"][c] = 1.0 + 0.1 * rng.standard_normal(n)
    f["vy"][c] = 0.1 * rng.standard_normal(n)
    f["vz"][c] = 0.1 * rng.standard_normal(n)
    f["p"][:] = rng.uniform(-1.0, 1.0, so.shape)
    so.set(**f)
    ss.set(**f)
    names = ("vx", "vy", "vz", "p", "pp")
    assert_fields_equal(so, ss, names=names, periodic=(0, 0, 0))
    for _ in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r in (0, S.NOT_CONVERGED)
        assert so.jacobi_iters() == st["jacobi_iters"]
        assert_fields_equal(so, ss, names=names, periodic=(0, 0, 0))
    P = ss.get()["p"]
    ghost, edge = P[pad - 1, pad:-pad, pad:-pad], P[pad, pad:-pad, pad:-pad]
    assert not np.array_equal(ghost, edge)   # not the zero-gradient rule


def test_null_wall_only_for_pressure_on_wall_axes():
    g = S.grid_desc((8, 8, 8), periodic=(1, 0, 0), pad=1)
    mg = S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64)
    nul = [S.FBC_NULL] + [S.FBC_NEUMANN] * 5
    bad = [dict(pbc=(nul, [0.0] * 6)),                                       # x is periodic
           dict(vbc=[([S.FBC_NULL] + [S.FBC_DIRICHLET] * 5, [0.0] * 6)] * 3),   # on velocity
           dict(tbc=([S.FBC_NEUMANN] * 3 + [S.FBC_NULL] + [S.FBC_DIRICHLET] * 2, [0.0] * 6))]   # on T
    for kw in bad:
        args = dict(dscheme=1, nlscheme=1, has_T=1, nu=1e-2, kappa=1e-2, dt=1e-3)
        args.update(kw)
        with pytest.raises(S.SarasError):
            S.Nse(g, mg, S.nse_params(**args))
    # a y wall is a wall axis here
    S.Nse(g, mg, S.nse_params(dscheme=1, nlscheme=1, has_T=0, nu=1e-2, dt=1e-3,
                              pbc=([S.FBC_NEUMANN] * 2 + [S.FBC_NULL] + [S.FBC_NEUMANN] * 3, [0.0] * 6))).close()


@pytest.mark.parametrize("n", [(12, 8, 16), (8, 16, 12)], ids=["ny<nz", "ny>nz"])
def test_morinishi_z_walls_ny_ne_nz_bitexact(n):
    """Morinishi with dScheme 2 on a z-walled grid with n_y != n_z: the z edge
    test reads core.ubound(1) (vfield.cc:329, 434; sfield.cc:236, 295), so
    the last z cell takes the 4th-order stencil and reads the second ghost
    layer above the top wall, which keeps the caller's values (random here).
    Three RK3 steps of a Boussinesq box agree bit for bit with the oracle."""
    pad = 2
    so, ss = pair(n, pad=pad, dscheme=2, nlscheme=3, nu=4e-3, kappa=3e-3, dt=2e-3, cn_tol=1e-9)
    z = O.field_metrics(so.d.g, 2)["x"]
    f = rbc_init(so.shape, n, pad, z, vel=0.1, seed=29)
    rng = np.random.default_rng(30)
    for k in ("vx", "vy", "vz", "t"):
        f[k][:, :, -1] = rng.uniform(-0.5, 0.5, f[k].shape[:2])   # the layer above the first top ghost
    so.set(**f)
    ss.set(**f)
    assert_fields_equal(so, ss)
    for _ in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r == 0
        assert so.jacobi_iters() == st["jacobi_iters"]
        assert_fields_equal(so, ss)
    # the layer is read: other values there give another result
    so2 = pair(n, pad=pad, dscheme=2, nlscheme=3, nu=4e-3, kappa=3e-3, dt=2e-3, cn_tol=1e-9)[1]
    f2 = {k: v.copy() for k, v in f.items()}
    f2["vx"][:, :, -1] += 0.25
    so2.set(**f2)
    assert so2.step()[0] == 0
    so.set(**f)
    ss.set(**f)
    ss.step()
    assert not np.array_equal(so2.get()["vx"], ss.get()["vx"])



def test_sgs_bench_hook():
    """saras_nse_bench_sgs times the sub-grid stress kernel alone; a handle
    without the model refuses."""
    so, ss, _ = tg_les_pair(n=16)
    ms = ss.bench_sgs(2)
    assert ms > 0.0
    g = S.grid_desc((8, 8, 8), periodic=(1, 1, 1), pad=2)
    plain = S.Nse(g, S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64),
                  S.nse_params(dscheme=2, nlscheme=3, has_T=0, nu=1e-2, dt=1e-3))
    with pytest.raises(S.SarasError):
        plain.bench_sgs(1)
