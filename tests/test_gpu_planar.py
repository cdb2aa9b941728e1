"""Planar (2-D, x-z) runs of the time-step path: the reference's PLANAR
build (lsRK3_d2 / eulerCN_d2 with hydro_d2) restated as one y row whose y
terms are exact zeros.  Bit-exact against the oracle's PLANAR restatement
(fields with their defined ghost cells, Jacobi counts, pressure solves) for
both integration schemes, dScheme 1 and 2, the three non-linear terms, walls
and the plume's open top; also pinned by what the mathematics fixes: the 2-D
Taylor-Green vortex is an exact Navier-Stokes solution whose kinetic energy
decays as exp(-4 nu t) in a (2 pi)^2 box, and a 3-D run of the same flow,
uniform in y, evolves the same x-z fields to the solvers' tolerances."""
import numpy as np
import pytest

import oracle as O
import saras_amd as S

pytestmark = pytest.mark.gpu


def tg_fields(shape, n, pad, L, dim):
    x = (np.arange(-pad, n + pad) + 0.5) * (L / n)
    sx, cx = np.sin(x), np.cos(x)
    ny = shape[1]
    u = np.broadcast_to((sx[:, None] * cx[None, :])[:, None, :], shape).copy()
    w = np.broadcast_to((-cx[:, None] * sx[None, :])[:, None, :], shape).copy()
    zero = np.zeros(shape)
    return dict(vx=u, vy=zero.copy(), vz=w, p=zero.copy(), t=zero.copy(), pp=zero.copy())


def run_tg(dim, scheme, n=64, steps=40, nu=0.05, dt=2e-3, nlscheme=1):
    pad = 1
    L = 2 * np.pi
    ny = 1 if dim == 2 else 4
    g = S.grid_desc((n, ny, n), dim=dim, periodic=(1, 1, 1), length=(L, L if dim == 3 else 1.0, L), pad=pad)
    mg = S.mg_params(vc_count=30, vc_tol=1e-10, solve_flag=0, mg_tol=1e-2, pre=3, post=3, bc=S.BC_NEUMANN,
                     precision=S.PREC_FP64)
    prm = S.nse_params(dscheme=1, nlscheme=nlscheme, has_T=0, nu=nu, kappa=nu, dt=dt, cn_tol=1e-12, scheme=scheme,
                       vbc=[([S.FBC_DIRICHLET] * 6, [0.0] * 6)] * 3)
    s = S.Nse(g, mg, prm)
    s.set(**tg_fields(s.shape, n, pad, L, dim))
    ke = [s.diag()["ke"]]
    for _ in range(steps):
        r, st = s.step()
        S.check(r)
        ke.append(s.diag()["ke"])
    f = s.get()
    s.close()
    return np.array(ke), f


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler_cn"])
def test_planar_taylor_green_decay(scheme):
    nu, dt, steps = 0.05, 2e-3, 40
    ke, _ = run_tg(2, scheme, nu=nu, dt=dt, steps=steps)
    t = dt * np.arange(steps + 1)
    exact = ke[0] * np.exp(-4 * nu * t)
    assert abs(ke[0] - 0.25) < 1e-3          # <u^2 + w^2>/2 of the vortex
    # second-order stencils on 64^2 and first/third-order time stepping
    assert np.max(np.abs(ke / exact - 1)) < 2e-4, np.max(np.abs(ke / exact - 1))


@pytest.mark.parametrize("nlscheme", [1, 2, 3], ids=["central", "upwind", "morinishi"])
def test_planar_equals_y_uniform_3d(nlscheme):
    """A planar run against a 3-D run whose fields do not vary along y, for
    the central, hybrid-upwind and Morinishi non-linear terms."""
    ke2, f2 = run_tg(2, 0, n=32, steps=10, nlscheme=nlscheme)
    ke3, f3 = run_tg(3, 0, n=32, steps=10, nlscheme=nlscheme)
    assert np.max(np.abs(ke2 / ke3 - 1)) < 1e-9
    c2 = (slice(1, -1), 0, slice(1, -1))
    for k in ("vx", "vz"):
        a, b = f2[k][c2], f3[k][1:-1, 1, 1:-1]
        assert np.max(np.abs(a - b)) < 1e-8 * np.max(np.abs(b)), k
    # the 3-D pressure iterate is y-uniform only to rounding (red-black order)
    assert np.max(np.abs(f3["vy"][1:-1, 1:-1, 1:-1])) < 1e-12


def planar_pair(n, periodic, gtype=(0, 0, 0), beta=(1.0, 1.0, 1.0), pad=1, mg=None, **kw):
    mg = mg or dict(vc_count=12, vc_tol=1e-9, solve_flag=0, mg_tol=1e-2, pre=3, post=3)
    gk = dict(dim=2, periodic=periodic, gtype=gtype, beta=beta, pad=pad)
    so = O.Nse(O.nse_desc(O.grid_desc(n, **gk), mg=O.mg_params(bc=O.BC_NEUMANN, **mg), **kw))
    ss = S.Nse(S.grid_desc(n, **gk), S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64, **mg), S.nse_params(**kw))
    assert so.shape == ss.shape
    return so, ss


def planar_mask(shape, n, pad, periodic):
    """Core cells and the ghosts the reference's imposeBCs/syncFaces define on
    the single y row: x and z periodic pads, the first wall layer."""
    m = np.zeros(shape, dtype=bool)
    I, K = np.meshgrid(np.arange(shape[0]) - pad, np.arange(shape[2]) - pad, indexing="ij")
    ox, oz = (I < 0) | (I >= n[0]), (K < 0) | (K >= n[2])
    okx = ox & ~oz & (periodic[0] | ((I >= -1) & (I <= n[0])))
    okz = oz & ~ox & (periodic[2] | ((K >= -1) & (K <= n[2])))
    m[:, 0, :] = (~ox & ~oz) | okx | okz
    return m


def assert_planar_equal(so, ss, n, pad, periodic, names=("vx", "vz", "p", "t", "pp")):
    fo, fs = so.get(), ss.get()
    m = planar_mask(so.shape, n, pad, periodic)
    for k in names:
        bad = np.flatnonzero(fo[k][m] != fs[k][m])
        assert bad.size == 0, (k, bad.size, np.abs(fo[k][m] - fs[k][m]).max())
    assert np.abs(fs["vy"]).max() == 0


PLANAR_CASES = [
    # (id, n, periodic, gtype, beta, pad, dscheme, nlscheme, scheme)
    ("rbc-central-rk3", (24, 1, 16), (1, 0, 0), (0, 0, 2), (1.0, 1.0, 1.2), 1, 1, 1, 0),
    ("rbc-upwind-rk3", (24, 1, 16), (1, 0, 0), (0, 0, 2), (1.0, 1.0, 1.2), 1, 1, 2, 0),
    ("rbc-morinishi-rk3", (24, 1, 16), (1, 0, 0), (0, 0, 2), (1.0, 1.0, 1.2), 1, 1, 3, 0),
    ("rbc-morinishi-euler", (24, 1, 16), (1, 0, 0), (0, 0, 2), (1.0, 1.0, 1.2), 1, 1, 3, 1),
    ("box-d2-central-rk3", (16, 1, 16), (0, 0, 0), (2, 0, 1), (1.1, 1.0, 0.9), 2, 2, 1, 0),
    ("box-d2-upwind-euler", (16, 1, 16), (0, 0, 0), (2, 0, 1), (1.1, 1.0, 0.9), 2, 2, 2, 1),
    ("periodic-d2-morinishi-rk3", (16, 1, 24), (1, 0, 1), (0, 0, 0), (1.0, 1.0, 1.0), 2, 2, 3, 0),
    # z walls with dScheme 2: the z edge test reads core.ubound(1) = 0, so the
    # last z cell reads the second ghost layer above the top wall
    ("rbc-d2-morinishi-rk3", (24, 1, 16), (1, 0, 0), (0, 0, 2), (1.0, 1.0, 1.2), 2, 2, 3, 0),
]


@pytest.mark.parametrize("case", PLANAR_CASES, ids=[c[0] for c in PLANAR_CASES])
def test_planar_steps_bitexact(case):
    """Three planar steps against the oracle's PLANAR restatement (lsRK3_d2 /
    eulerCN_d2, vfield/sfield PLANAR branches, predictors.cc PLANAR Jacobi):
    a buoyant scalar with a warm bottom and seeded perturbations."""
    _, n, periodic, gtype, beta, pad, dscheme, nlscheme, scheme = case
    so, ss = planar_pair(n, periodic, gtype, beta, pad, dscheme=dscheme, nlscheme=nlscheme, scheme=scheme, has_T=1,
                         nu=2e-3, kappa=3e-3, dt=2e-3, cn_tol=1e-9, buoy=(0.0, 0.0, -1.0))
    rng = np.random.default_rng(21)
    sh = so.shape
    c = (slice(pad, -pad), 0, slice(pad, -pad))
    f = {k: np.zeros(sh) for k in ("vx", "vy", "vz", "p", "t", "pp")}
    f["t"][c] = 0.5 + 0.05 * rng.standard_normal((n[0], n[2]))
    f["vx"][c] = 0.1 * rng.standard_normal((n[0], n[2]))
    f["vz"][c] = 0.1 * rng.standard_normal((n[0], n[2]))
    if pad == 2 and not periodic[2]:
        for k in ("vx", "vz", "t"):   # the second layer above the top wall, kept as given
            f[k][:, 0, -1] = rng.uniform(-0.5, 0.5, sh[0])
    so.set(**f)
    ss.set(**f)
    assert_planar_equal(so, ss, n, pad, periodic)
    for step in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r == S.OK
        assert so.jacobi_iters() == st["jacobi_iters"], step
        assert so.mg_stats()["cycles"] == st["mg"]["cycles"]
        assert so.mg_stats()["residual"] == st["mg"]["residual"]
        assert_planar_equal(so, ss, n, pad, periodic)
    do, ds = so.diag(), ss.diag()
    assert do["divmax"] == ds["divmax"]
    for k in ("ke", "nuvol", "nuwall"):
        assert abs(do[k] - ds[k]) <= 1e-12 * abs(do[k]), (k, do[k], ds[k])


def test_planar_plume_bitexact():
    """The planar plume (probType 9): plumeOut top wall of T and the Dirichlet
    top of the pressure-correction solve (poisson_d2.cc:641-647), three steps
    bit for bit with the oracle, ghosts of Pp included."""
    n, pad = (24, 1, 24), 1
    D, N = S.FBC_DIRICHLET, S.FBC_NEUMANN
    vbc = [([D] * 6, [0.0] * 6), ([D] * 6, [0.0] * 6), ([D] * 5 + [N], [0.0] * 6)]
    tbc = ([N] * 4 + [D, S.FBC_PLUME_OUT], [0.0] * 4 + [1.0, 0.0])
    so, ss = planar_pair(n, (1, 0, 0), (0, 0, 1), (1.0, 1.0, 0.8), pad, dscheme=1, nlscheme=1, has_T=1, nu=1e-2,
                         kappa=1e-2, dt=1e-3, cn_tol=1e-10, buoy=(0.0, 0.0, -1.0), vbc=vbc, tbc=tbc)
    rng = np.random.default_rng(2)
    sh = so.shape
    f = {k: np.zeros(sh) for k in ("vx", "vy", "vz", "p", "t", "pp")}
    f["t"][pad:-pad, 0, pad:-pad] = rng.uniform(0.0, 0.1, (n[0], n[2]))
    so.set(**f)
    ss.set(**f)
    for step in range(3):
        assert so.step() == 0
        r, st = ss.step()
        assert r == S.OK
        assert so.jacobi_iters() == st["jacobi_iters"], step
        assert_planar_equal(so, ss, n, pad, (1, 0, 0))


def test_planar_rejects_unsupported_schemes():
    mg = S.mg_params(precision=S.PREC_FP64, bc=S.BC_NEUMANN)
    gp = S.grid_desc((16, 1, 16), dim=2, periodic=(1, 1, 1), pad=2)
    with pytest.raises(S.SarasError):   # the LES model is 3-D only
        S.Nse(gp, mg, S.nse_params(dscheme=2, nlscheme=3, has_T=0, nu=0.01, kappa=0.01, dt=1e-3, les=1))
    gw = S.grid_desc((16, 1, 16), dim=2, periodic=(1, 1, 0), pad=1)
    plate = ([S.FBC_NEUMANN] * 4 + [S.FBC_HOT_PLATE, S.FBC_PLUME_OUT], [0.0] * 6)
    with pytest.raises(S.SarasError):   # PLANAR builds have no heating plate
        S.Nse(gw, mg, S.nse_params(dscheme=1, nlscheme=1, has_T=1, nu=0.01, kappa=0.01, dt=1e-3, tbc=plate))


def test_planar_plume_pressure_top_wall():
    """A planar plume (probType 9): T's top wall is plumeOut and the pressure
    correction is solved with a Dirichlet top wall (poisson_d2.cc:641-647).
    Fluid driven by the warm bottom leaves through the top: the steps stay
    finite, and Pp's ghosts are -Pp(n-1) at the top and Pp(0) at the bottom."""
    n, pad = 32, 1
    g = S.grid_desc((n, 1, n), dim=2, periodic=(1, 1, 0), pad=pad)
    mg = S.mg_params(vc_count=30, vc_tol=1e-10, solve_flag=0, pre=3, post=3, bc=S.BC_NEUMANN, precision=S.PREC_FP64)
    D, N = S.FBC_DIRICHLET, S.FBC_NEUMANN
    vbc = [([D] * 6, [0.0] * 6), ([D] * 6, [0.0] * 6), ([D] * 5 + [N], [0.0] * 6)]
    tbc = ([N] * 4 + [D, S.FBC_PLUME_OUT], [0.0] * 4 + [1.0, 0.0])
    s = S.Nse(g, mg, S.nse_params(dscheme=1, nlscheme=1, has_T=1, nu=1e-2, kappa=1e-2, dt=1e-3, cn_tol=1e-10,
                                  buoy=(0.0, 0.0, -1.0), vbc=vbc, tbc=tbc))
    rng = np.random.default_rng(2)
    zero = np.zeros(s.shape)
    t = zero.copy()
    t[pad:-pad, 0, pad:-pad] = rng.uniform(0.0, 0.1, (n, n))
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=t, pp=zero)
    for _ in range(5):
        r, st = s.step()
        S.check(r)
    f = s.get()
    for k in ("vx", "vz", "p", "t", "pp"):
        assert np.all(np.isfinite(f[k])), k
    assert np.abs(f["vz"][pad:-pad, 0, -pad - 1]).max() > 0           # flow through the open top
    pp = f["pp"][pad:-pad, 0, :]
    np.testing.assert_array_equal(pp[:, -1], -pp[:, -2])            # Dirichlet top of the Pp solve
    np.testing.assert_array_equal(pp[:, 0], pp[:, 1])               # Neumann bottom
