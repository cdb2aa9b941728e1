"""Pins of the oracle's multigrid restatement (lib/poisson/*) by brute force on
tiny grids, closed-form solutions and the reference's own r^2/6 mgTest case."""
import numpy as np
import pytest

import oracle as O
from dense_ops import assemble, axis_coeffs, wall_rhs_shift

KIND = {O.BC_NEUMANN: "N", O.BC_DIRICHLET: "D", O.BC_DIRICHLET_TEST: "D"}


def make_case(n, dim, per, gtype=(0, 0, 0), beta=(1.0, 1.0, 1.0), bc=O.BC_NEUMANN, L=(1.0, 1.0, 1.0), **kw):
    g = O.grid_desc(n, dim=dim, periodic=per, gtype=gtype, beta=beta, length=L)
    p = O.mg_params(bc=bc, **kw)
    return g, p


def dense_level(g, level, bc):
    axes = (0, 2) if g.dim == 2 else (0, 1, 2)
    coeffs, kinds, shape = [], [], []
    for a in axes:
        m = O.level_metrics(g, level, a)
        h = (1 << level) / g.n[a]
        coeffs.append(axis_coeffs(m["xix2"], m["xixx"], h))
        kinds.append("P" if g.periodic[a] else KIND[bc])
        shape.append(g.n[a] >> level)
    return assemble(shape, coeffs, kinds), tuple(shape), coeffs, kinds


def pad_core(c):
    return np.pad(c, 1)


CASES = [
    ((8, 8, 8), 3, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET),
    ((8, 8, 8), 3, (0, 0, 0), (2, 2, 2), O.BC_DIRICHLET),
    ((8, 8, 8), 3, (1, 1, 0), (0, 0, 2), O.BC_NEUMANN),
    ((8, 4, 8), 3, (1, 1, 1), (0, 0, 0), O.BC_NEUMANN),
    ((6, 6, 6), 3, (1, 0, 1), (0, 1, 0), O.BC_NEUMANN),
    ((3, 3, 3), 3, (1, 1, 1), (0, 0, 0), O.BC_NEUMANN),
    ((16, 1, 8), 2, (0, 0, 0), (2, 0, 0), O.BC_DIRICHLET),
    ((16, 1, 16), 2, (1, 0, 0), (0, 0, 2), O.BC_NEUMANN),
    ((5, 1, 7), 2, (1, 0, 1), (0, 0, 0), O.BC_NEUMANN),
]


def rand_in_range(A, shape, rng):
    x = rng.standard_normal(shape)
    return x, (A @ x.ravel()).reshape(shape)


@pytest.mark.parametrize("n,dim,per,gt,bc", CASES)
def test_smoother_fixed_point(n, dim, per, gt, bc):
    """The exact discrete solution is a fixed point of one RB-SOR sweep: any
    wrong sign, coefficient or neighbour index in the smoother breaks it."""
    g, p = make_case(n, dim, per, gtype=gt, beta=(1.3, 1.1, 1.5), bc=bc, sor=1.3)
    mg = O.Multigrid(g, p)
    A, shape, _, _ = dense_level(g, 0, bc)
    rng = np.random.default_rng(1)
    x, b = rand_in_range(A, shape, rng)
    mg.set_level(0, pad_core(x.reshape(mg.padded_shape(0)[0] - 2, -1, mg.padded_shape(0)[-1] - 2)).reshape(mg.padded_shape(0))
                 if False else pad_core(x), pad_core(b))
    mg.smooth(0, 1, zero_bc=1)
    lhs, _ = mg.get_level(0)
    c = lhs[(slice(1, -1),) * lhs.ndim]
    np.testing.assert_allclose(c, x, rtol=0, atol=1e-11 * np.abs(x).max())


def colour_masks(shape):
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    par = sum(grids) % 2
    return par == 0, par == 1


def neighbour(x, ax, off, kind):
    """Value of the neighbour at offset off along ax with ghost rule kind."""
    n = x.shape[ax]
    y = np.roll(x, -off, axis=ax)
    if kind == "P":
        return y
    edge = [slice(None)] * x.ndim
    edge[ax] = n - 1 if off == +1 else 0
    own = x[tuple(edge)]
    y[tuple(edge)] = own if kind == "N" else -own
    return y


def half_sweep_jacobi(x, b, coeffs, kinds, mask, w):
    """Characterisation of one colour pass of the reference: wall ghosts are
    images of the cell itself and periodic ghosts are refreshed before every
    colour pass (poisson_d3.cc:100,173), so every cell of the colour is
    updated from values as they were at the start of the pass."""
    num = np.zeros_like(x)
    den = np.zeros_like(x)
    for ax, ((c2, c1), kind) in enumerate(zip(coeffs, kinds)):
        shp = [1] * x.ndim
        shp[ax] = -1
        c2b, c1b = c2.reshape(shp), c1.reshape(shp)
        xp, xm = neighbour(x, ax, +1, kind), neighbour(x, ax, -1, kind)
        num = num + c2b * (xp + xm) + c1b * (xp - xm)
        den = den + c2b
    new = (1 - w) * x + (num - b) * w / (2.0 * den)
    return np.where(mask, new, x)


@pytest.mark.parametrize("n,dim,per,gt,bc", CASES)
def test_sweep_matches_colour_jacobi_characterisation(n, dim, per, gt, bc):
    g, p = make_case(n, dim, per, gtype=gt, beta=(1.3, 1.1, 1.5), bc=bc, sor=0.9)
    mg = O.Multigrid(g, p)
    _, shape, coeffs, kinds = dense_level(g, 0, bc)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape)
    b = rng.standard_normal(shape)
    mg.set_level(0, pad_core(x), pad_core(b))
    mg.smooth(0, 2, zero_bc=1)
    lhs, _ = mg.get_level(0)
    red, black = colour_masks(shape)
    y = x.copy()
    for _ in range(2):
        y = half_sweep_jacobi(y, b, coeffs, kinds, red, 0.9)
        y = half_sweep_jacobi(y, b, coeffs, kinds, black, 0.9)
    np.testing.assert_allclose(lhs[(slice(1, -1),) * lhs.ndim], y, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n,dim,per,gt,bc", CASES)
def test_vcycle_solve_matches_dense(n, dim, per, gt, bc):
    # moderate stretching: with beta = 1.5 on 8-16 cells and Neumann walls the
    # reference V-cycle itself diverges (coarse solve capped, restriction
    # weights of poisson_d3.cc:494-508), see test_strong_stretch_singular_diverges
    g, p = make_case(n, dim, per, gtype=gt, beta=(0.6, 0.5, 0.7), bc=bc, vc_count=60, vc_tol=1e-12,
                     mg_tol=1e-13, pre=2, post=2)
    mg = O.Multigrid(g, p)
    A, shape, _, _ = dense_level(g, 0, bc)
    rng = np.random.default_rng(3)
    xt, b = rand_in_range(A, shape, rng)
    x0 = np.zeros(O.field_shape(g, 1))
    bf = np.pad(b, 1)
    x, st = mg.solve(x0, bf, 1)
    xc = O.core(x, 1)
    r = b.ravel() - A @ xc.ravel()
    assert np.abs(r).max() / np.abs(b).max() < 1e-11
    assert st["residual"] < 1e-12
    # independent check of computeError's L-infinity value
    assert st["residual"] == pytest.approx(np.abs(r).max() / np.abs(b).max(), rel=1e-6, abs=1e-15)
    if all(per[a] for a in ((0, 2) if dim == 2 else range(3))) or bc == O.BC_NEUMANN:
        if all(t == 0 for t in gt):
            d = xc - xt
            np.testing.assert_allclose(d, d.mean(), atol=1e-9 * np.abs(xt).max())
    else:
        np.testing.assert_allclose(xc, xt, atol=1e-9 * np.abs(xt).max())


@pytest.mark.parametrize("res_type", [0, 1, 2])
def test_compute_error_norms(res_type):
    g, p = make_case((8, 8, 8), 3, (0, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.4), bc=O.BC_DIRICHLET)
    mg = O.Multigrid(g, p)
    A, shape, _, _ = dense_level(g, 0, O.BC_DIRICHLET)
    rng = np.random.default_rng(4)
    x, b = rng.standard_normal(shape), rng.standard_normal(shape)
    mg.set_level(0, pad_core(x), pad_core(b))
    mg.impose_bc(0, 0)
    r = b.ravel() - A @ x.ravel()
    bb = b.ravel()
    want = [np.abs(r).max() / np.abs(bb).max(), np.abs(r).sum() / np.abs(bb).sum(),
            np.sqrt((r ** 2).mean()) / np.sqrt((bb ** 2).mean())][res_type]
    assert mg.error(res_type) == pytest.approx(want, rel=1e-12)


def test_restrict_uniform_is_mean_and_prolong_is_injection():
    g, p = make_case((8, 8, 8), 3, (1, 1, 1))
    mg = O.Multigrid(g, p)
    rng = np.random.default_rng(5)
    A, shape, _, _ = dense_level(g, 0, O.BC_NEUMANN)
    x, b = rng.standard_normal(shape), rng.standard_normal(shape)
    mg.set_level(0, pad_core(x), pad_core(b))
    mg.impose_bc(0, 1)
    mg.residual_restrict(0)
    r = (b.ravel() - A @ x.ravel()).reshape(shape)
    _, rc = mg.get_level(1)
    want = r.reshape(4, 2, 4, 2, 4, 2).mean(axis=(1, 3, 5))
    np.testing.assert_allclose(rc[1:-1, 1:-1, 1:-1], want, rtol=1e-12, atol=1e-12)
    e = rng.standard_normal((4, 4, 4))
    mg.set_level(1, pad_core(e), None)
    mg.prolong_correct(0)
    lhs, _ = mg.get_level(0)
    inj = np.repeat(np.repeat(np.repeat(e, 2, 0), 2, 1), 2, 2)
    np.testing.assert_allclose(lhs[1:-1, 1:-1, 1:-1], inj + x, rtol=1e-15, atol=1e-15)


def test_restrict_stretched_weights():
    """poisson_d3.cc:494-508: child weights are separable products of the
    per-axis fractions (X - x_2i)/(x_2i+1 - x_2i) for the 2i child, and sum to
    one (a constant residual restricts to the same constant)."""
    g, p = make_case((8, 8, 8), 3, (1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.8))
    mg = O.Multigrid(g, p)
    mf, mc = O.level_metrics(g, 0, 2), O.level_metrics(g, 1, 2)
    zf, zc = mf["x"][1:-1], mc["x"][1:-1]
    fz = (zc - zf[0::2]) / (zf[1::2] - zf[0::2])          # weight of child 2k along z
    # residual = indicator of the even-z children: restricted value = fz
    for val in (1.0, 0.0):
        tmp = np.zeros((8, 8, 8))
        tmp[:, :, 0::2] = 1.0
        tmp[:, :, 1::2] = val
        # make the residual equal tmp: x = 0, b = tmp
        mg.set_level(0, np.zeros((10, 10, 10)), pad_core(tmp))
        mg.impose_bc(0, 1)
        mg.residual_restrict(0)
        _, rc = mg.get_level(1)
        want = fz + val * (1 - fz) if val else fz
        np.testing.assert_allclose(rc[1:-1, 1:-1, 1:-1], np.broadcast_to(want, (4, 4, 4)), rtol=1e-13)
    assert np.all((fz > 0) & (fz < 1))


def test_dirichlet_sine_second_order():
    """u = sin(pi x) sin(pi y) sin(pi z), b = -3 pi^2 u on [0,1]^3 with
    homogeneous Dirichlet walls (the config-1 problem): the discrete solution
    converges to u at second order."""
    errs = []
    for nn in (16, 32):
        g, p = make_case((nn,) * 3, 3, (0, 0, 0), bc=O.BC_DIRICHLET, vc_count=40, vc_tol=1e-11,
                         mg_tol=1e-10, pre=2, post=2)
        mg = O.Multigrid(g, p)
        xc = (np.arange(nn) + 0.5) / nn
        X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
        u = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
        b = np.pad(-3 * np.pi ** 2 * u, 1)
        x, st = mg.solve(np.zeros_like(b), b, 1)
        assert st["residual"] < 1e-11
        errs.append(np.abs(O.core(x, 1) - u).max())
    assert 3.5 < errs[0] / errs[1] < 4.5


def test_periodic_fourier_mode_exact_eigenvalue():
    n = 16
    g, p = make_case((n, n, n), 3, (1, 1, 1), vc_count=60, vc_tol=1e-13, mg_tol=1e-14, pre=2, post=2)
    mg = O.Multigrid(g, p)
    xc = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
    k = (1, 2, 3)
    u = np.cos(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z))
    h = 1.0 / n
    lam = sum(-4 * np.sin(np.pi * kk * h) ** 2 / h ** 2 for kk in k)
    x, st = mg.solve(np.zeros((n + 2,) * 3), np.pad(lam * u, 1), 1)
    xc_ = O.core(x, 1)
    np.testing.assert_allclose(xc_ - xc_.mean(), u, atol=1e-10)


@pytest.mark.parametrize("nn", [16, 32])
def test_reference_mgtest_r2_over_6(nn):
    """The reference's own Poisson self-test (poisson.cc:150-175,
    poisson_d3.cc:824-890): b = 1 with Dirichlet walls r^2/6 on the unit cube;
    the deviation from r^2/6 falls as h^2."""
    g, p = make_case((nn,) * 3, 3, (0, 0, 0), bc=O.BC_DIRICHLET_TEST, vc_count=20, vc_tol=1e-12,
                     mg_tol=1e-6, pre=5, post=5)
    mg = O.Multigrid(g, p)
    xc = (np.arange(nn) + 0.5) / nn
    X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
    pa = ((X - .5) ** 2 + (Y - .5) ** 2 + (Z - .5) ** 2) / 6.0
    x, st = mg.solve(np.zeros((nn + 2,) * 3), np.pad(np.ones((nn,) * 3), 1), 1)
    dev = np.abs(O.core(x, 1) - pa).max()
    assert dev < 0.1 / nn ** 2
    # dense solve of the same discrete problem (wall values moved to the RHS)
    if nn == 16:
        A, shape, coeffs, kinds = dense_level(g, 0, O.BC_DIRICHLET_TEST)
        c = (np.arange(nn) + 0.5) / nn
        wx = ((0.5) ** 2 + (c[:, None] - .5) ** 2 + (c[None, :] - .5) ** 2) / 6.0
        walls = [(wx, wx), (wx, wx), (wx, wx)]
        sh = wall_rhs_shift(shape, coeffs, kinds, walls)
        xd = np.linalg.solve(A, (np.ones(shape) - sh).ravel()).reshape(shape)
        np.testing.assert_allclose(O.core(x, 1), xd, atol=1e-9)


def test_coarse_cap_flag():
    g, p = make_case((16, 16, 16), 3, (1, 1, 1), mg_tol=1e-30, vc_count=1, pre=1, post=1)
    mg = O.Multigrid(g, p)
    rng = np.random.default_rng(7)
    b = np.pad(rng.standard_normal((16, 16, 16)), 1)
    _, st = mg.solve(np.zeros_like(b), b, 1)
    assert st["coarse_capped"] == 1
    assert st["coarse_iters_max"] == mg.max_count + 1


def test_2d_planar_dirichlet_sine():
    n = 64
    g, p = make_case((n, 1, n), 2, (0, 0, 0), bc=O.BC_DIRICHLET, vc_count=40, vc_tol=1e-10,
                     mg_tol=1e-10, pre=2, post=2)
    mg = O.Multigrid(g, p)
    xc = (np.arange(n) + 0.5) / n
    X, Z = np.meshgrid(xc, xc, indexing="ij")
    u = np.sin(np.pi * X) * np.sin(2 * np.pi * Z)
    b = np.pad(-5 * np.pi ** 2 * u, 1)
    x, st = mg.solve(np.zeros_like(b), b, 1)
    assert st["residual"] < 1e-10
    assert np.abs(O.core(x, 1) - u).max() < 4.0 / n ** 2


def test_strong_stretch_singular_diverges():
    """Recorded behaviour of the reference algorithm: 8^3, periodic x/y,
    Neumann z, double-sided tanh beta = 1.5 -> the residual grows cycle after
    cycle while the same grid with Dirichlet walls converges."""
    out = {}
    for bc in (O.BC_NEUMANN, O.BC_DIRICHLET):
        g, p = make_case((8, 8, 8), 3, (1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), bc=bc,
                         vc_count=8, vc_tol=1e-14, mg_tol=1e-2, pre=2, post=2)
        mg = O.Multigrid(g, p)
        A, shape, _, _ = dense_level(g, 0, bc)
        _, b = rand_in_range(A, shape, np.random.default_rng(3))
        _, st = mg.solve(np.zeros((10, 10, 10)), np.pad(b, 1), 1, history=True)
        out[bc] = st["history"]
    assert out[O.BC_NEUMANN][-1] > 10 * out[O.BC_NEUMANN][0]
    assert out[O.BC_DIRICHLET][-1] < 1e-3 * out[O.BC_DIRICHLET][0]


def test_residual_entry_matches_dense_operator():
    g, p = make_case((8, 8, 8), 3, (1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 0.6), bc=O.BC_NEUMANN)
    mg = O.Multigrid(g, p)
    A, shape, _, _ = dense_level(g, 0, O.BC_NEUMANN)
    x = np.random.default_rng(8).standard_normal(shape)
    ax = mg.apply(pad_core(x))
    np.testing.assert_allclose(ax[1:-1, 1:-1, 1:-1].ravel(), A @ x.ravel(), rtol=1e-12, atol=1e-10)
    assert np.all(ax[0] == 0) and np.all(ax[:, :, -1] == 0)


@pytest.mark.parametrize("gt", [(0, 0, 0), (0, 0, 2)], ids=["uniform", "tanh-z"])
def test_2d_neumann_top_dirichlet_matches_dense(gt):
    """multigrid_d2 of probType 9 (poisson_d2.cc:641-647): Neumann walls with a
    homogeneous Dirichlet top.  The problem is no longer singular, so the
    V-cycle solution equals the dense solve itself; a Neumann top, a wrong
    sign or the rule on the bottom wall misses it."""
    n = (16, 1, 16)
    g, p = make_case(n, 2, (1, 0, 0), gtype=gt, beta=(0.6, 0.5, 0.7), bc=O.BC_NEUMANN, vc_count=80,
                     vc_tol=1e-12, mg_tol=1e-13, pre=2, post=2, top_dirichlet=1)
    mg = O.Multigrid(g, p)
    A, shape, coeffs, kinds = dense_level(g, 0, O.BC_NEUMANN)
    A = assemble(shape, coeffs, [kinds[0], "ND"])
    rng = np.random.default_rng(8)
    xt, b = rand_in_range(A, shape, rng)
    x, st = mg.solve(np.zeros(O.field_shape(g, 1)), np.pad(b, 1), 1)
    xc = O.core(x, 1)
    np.testing.assert_allclose(xc, xt, atol=1e-9 * np.abs(xt).max())
    assert np.array_equal(x[1:-1, -1], -x[1:-1, -2])
