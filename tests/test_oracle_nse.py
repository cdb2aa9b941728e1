"""Pins of the oracle's Navier-Stokes restatement (lib/field, lib/boundary,
lib/timestep, lib/io/tseries.cc) by what the mathematics fixes:

* derivative stencils: exactness on polynomials of the computational
  coordinate xi (4th order interior, 2nd order edge cells), chain rule with the
  grid metrics, convergence order on a sinusoid;
* Morinishi non-linear term: exact skew-symmetric form on low-degree
  polynomial velocities and scalars;
* boundary rules: a linear field's ghost value equals the field at the mirror
  point of the wall (Dirichlet value / Neumann gradient);
* Jacobi Crank-Nicolson: equal to the dense implicit solve (walls included);
* RK3 stage: the exact amplification factor of a periodic Fourier mode under
  the heat equation, and Taylor-Green decay for the full non-linear step;
* diagnostics: conduction state Nu = 1, uniform-flow kinetic energy.
"""
import numpy as np
import pytest

import oracle as O

FD, FN = O.FBC_DIRICHLET, O.FBC_NEUMANN


def make(n, periodic=(1, 1, 1), gtype=(0, 0, 0), beta=(1.0, 1.0, 1.0), pad=2, **kw):
    g = O.grid_desc(n, periodic=periodic, gtype=gtype, beta=beta, pad=pad)
    kw.setdefault("mg", O.mg_params(vc_count=20, vc_tol=1e-12, solve_flag=0, pre=3, post=3))
    d = O.nse_desc(g, **kw)
    return g, d, O.Nse(d)


def xi_axis(n, pad):
    return (np.arange(-pad, n + pad) + 0.5) / n


def along(vals, axis, shape):
    s = [1, 1, 1]
    s[axis] = len(vals)
    return np.broadcast_to(np.asarray(vals).reshape(s), shape).copy()


def core(a, p):
    return a[p:-p, p:-p, p:-p]


# ---- derivative stencils --------------------------------------------------

POLY4 = np.polynomial.Polynomial([0.3, -1.2, 0.7, 2.1, -0.9])
POLY2 = np.polynomial.Polynomial([0.3, -1.2, 0.7])


@pytest.mark.parametrize("gtype", [(0, 0, 0), (2, 1, 2)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_derivatives_exact_on_xi_polynomials(gtype, axis):
    n, p = 12, 2
    g, d, s = make(n, periodic=(0, 0, 0), gtype=gtype, beta=(1.3, 0.9, 1.5))
    xi = xi_axis(n, p)
    m = O.field_metrics(g, axis)
    for poly, interior_only in ((POLY2, False), (POLY4, True)):
        F = along(poly(xi), axis, s.shape)
        d1 = core(s.op(axis, F), p)
        d2 = core(s.op(3 + axis, F), p)
        e1 = along((m["xi_x"] * poly.deriv(1)(xi))[p:-p], axis, d1.shape)
        e2 = along((m["xixx"] * poly.deriv(1)(xi) + m["xix2"] * poly.deriv(2)(xi))[p:-p], axis, d1.shape)
        sl = [slice(None)] * 3
        if interior_only:
            sl[axis] = slice(1, n - 1)
        sl = tuple(sl)
        scale = max(np.abs(e2).max(), 1.0)
        assert np.abs(d1[sl] - e1[sl]).max() < 1e-9 * scale
        assert np.abs(d2[sl] - e2[sl]).max() < 1e-9 * scale
        # derivatives across the other axes vanish
        for other in {0, 1, 2} - {axis}:
            assert np.abs(s.op(other, F)).max() < 1e-9 * scale
    # degree 4 is not exact at the 2nd-order edge cells (the override is real)
    F = along(POLY4(xi), axis, s.shape)
    d1 = core(s.op(axis, F), p)
    e1 = along((m["xi_x"] * POLY4.deriv(1)(xi))[p:-p], axis, d1.shape)
    edge = [slice(None)] * 3
    edge[axis] = 0
    assert np.abs(d1[tuple(edge)] - e1[tuple(edge)]).max() > 1e-6


def test_derivative_convergence_orders():
    """sin(2 pi x) on a periodic uniform axis: 4th order in the interior,
    2nd order at the first/last cell (the reference applies the 2nd-order
    override on every rank-edge cell, periodic axes included)."""
    errs_in, errs_edge = [], []
    for n in (16, 32, 64):
        g, d, s = make(n, periodic=(1, 1, 1))
        p = 2
        x = xi_axis(n, p)
        F = along(np.sin(2 * np.pi * x), 0, s.shape)
        d1 = core(s.op(0, F), p)[:, 0, 0]
        ex = 2 * np.pi * np.cos(2 * np.pi * x[p:-p])
        errs_in.append(np.abs(d1[1:-1] - ex[1:-1]).max())
        errs_edge.append(max(abs(d1[0] - ex[0]), abs(d1[-1] - ex[-1])))
    r_in = np.log2(np.array(errs_in[:-1]) / np.array(errs_in[1:]))
    r_ed = np.log2(np.array(errs_edge[:-1]) / np.array(errs_edge[1:]))
    assert np.all(r_in > 3.7) and np.all(r_ed > 1.8) and np.all(r_ed < 2.3)


def test_laplacian_is_sum_of_second_derivatives():
    n, p = 10, 2
    g, d, s = make(n, periodic=(1, 0, 0), gtype=(0, 1, 2), beta=(1.0, 1.1, 1.4))
    rng = np.random.default_rng(5)
    F = rng.standard_normal(s.shape)
    lap = s.op(6, F)
    ref = ((0.0 + s.op(3, F)) + s.op(4, F)) + s.op(5, F)
    assert np.array_equal(core(lap, p), core(ref, p))


# ---- Morinishi non-linear term ---------------------------------------------

def test_morinishi_scalar_skew_symmetric_form():
    """H = -1/2 [u_a dF/dx_a + d(u_a F)/dx_a] with u_a linear and F quadratic
    in xi along each axis (products of degree <= 3: exact at interior cells)."""
    n, p = 12, 2
    g, d, s = make(n, periodic=(0, 0, 0), gtype=(2, 0, 1), beta=(1.2, 1.0, 0.8), nlscheme=3)
    xi = [xi_axis(n, p)] * 3
    m = [O.field_metrics(g, a)["xi_x"] for a in range(3)]
    sh = s.shape
    X, Y, Z = (along(xi[a], a, sh) for a in range(3))
    u = [0.4 + 0.3 * Y, -0.2 + 0.5 * Z, 0.1 - 0.6 * X]        # u_a independent of x_a
    F = 0.2 + X - 0.7 * X ** 2 + 0.5 * Y ** 2 - 0.3 * Z + 0.9 * Z ** 2
    dF = [1 - 1.4 * X, 1.0 * Y, -0.3 + 1.8 * Z]
    s.set(vx=u[0], vy=u[1], vz=u[2], t=np.zeros(sh), p=np.zeros(sh), pp=np.zeros(sh))
    H = s.op(7, F)
    # BCs of the set() call overwrite the first ghost layer of V on the walls;
    # interior cells (>= 2 away from every wall) never read it
    exp = np.zeros(sh)
    for a in range(3):
        xa = along(m[a], a, sh)
        exp -= 0.5 * (u[a] * xa * dF[a] + xa * (u[a] * dF[a]))   # d(u_a F) = u_a dF since du_a/dx_a = 0
    sl = (slice(p + 2, -p - 2),) * 3
    assert np.abs(H[sl] - exp[sl]).max() < 1e-9


def test_morinishi_vector_rhs_nonlinear_term():
    """tempVF after the RHS assembly holds the Morinishi term of V
    (vfield.cc:271-462) on a periodic box: check against the analytic
    -1/2 [u_j du_i/dx_j + d(u_j u_i)/dx_j] for trigonometric velocities
    (4th-order interior, 2nd-order first/last cells -> tolerance by n)."""
    n, p = 32, 2
    g, d, s = make(n, periodic=(1, 1, 1), has_T=0, nu=0.0, kappa=0.0, nlscheme=3)
    x = xi_axis(n, p)
    sh = s.shape
    X, Y, Z = (along(x, a, sh) for a in range(3))
    tp = 2 * np.pi
    u = np.sin(tp * X) * np.cos(tp * Y)
    v = 0.5 * np.cos(tp * Z)
    w = np.sin(tp * Y)
    s.set(vx=u, vy=v, vz=w, p=np.zeros(sh), t=np.zeros(sh), pp=np.zeros(sh))
    r = s.rhs(0)
    ux, uy = tp * np.cos(tp * X) * np.cos(tp * Y), -tp * np.sin(tp * X) * np.sin(tp * Y)
    vz = -0.5 * tp * np.sin(tp * Z)
    wy = tp * np.cos(tp * Y)
    div = ux  # v_y = w_z = 0
    adv = [u * ux + v * uy, w * vz, v * wy]
    exp = [-(adv[0] + 0.5 * div * u), -(adv[1] + 0.5 * div * v), -(adv[2] + 0.5 * div * w)]
    for c, key in enumerate(("hx", "hy", "hz")):
        err = np.abs(core(r[key] - exp[c], p))
        assert err[2:-2, 2:-2, 2:-2].max() < 5e-3, (key, err.max())
        assert err.max() < 0.15


# ---- boundary rules ----------------------------------------------------------

@pytest.mark.parametrize("kind", [FD, FN])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_wall_ghost_is_mirror_value(kind, axis):
    n, p = 8, 2
    per = [1, 1, 1]
    per[axis] = 0
    L = 1.0
    g, d, s = make(n, periodic=per, gtype=(2, 2, 2), beta=(1.5, 1.5, 1.5))
    xc = O.field_metrics(g, axis)["x"]
    a0, slope = 0.7, -1.3
    F = along(a0 + slope * xc, axis, s.shape)
    if kind == FD:
        vals = [0.0] * 6
        vals[2 * axis], vals[2 * axis + 1] = a0, a0 + slope * L
    else:
        vals = [0.0] * 6
        vals[2 * axis] = vals[2 * axis + 1] = slope
    d2 = O.nse_desc(g, has_T=1, tbc=([kind] * 6, vals), mg=d.mg)
    s2 = O.Nse(d2)
    s2.set(vx=np.zeros(s.shape), vy=np.zeros(s.shape), vz=np.zeros(s.shape),
           p=np.zeros(s.shape), t=F, pp=np.zeros(s.shape))
    T = s2.get()["t"]
    x0, xn = xc[p], xc[-p - 1]
    lo = [slice(p, -p)] * 3
    hi = list(lo)
    lo[axis], hi[axis] = p - 1, -p
    assert np.allclose(T[tuple(lo)], a0 + slope * (-x0), atol=1e-13)
    assert np.allclose(T[tuple(hi)], a0 + slope * (2 * L - xn), atol=1e-13)


def test_periodic_pads_wrap_two_layers():
    n, p = 8, 2
    g, d, s = make(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5))
    rng = np.random.default_rng(3)
    F = rng.standard_normal(s.shape)
    s.set(vx=F, vy=F, vz=F, p=F, t=F, pp=np.zeros(s.shape))
    V = s.get()["vx"]
    c = slice(p, -p)
    assert np.array_equal(V[0:p, c, c], V[n:n + p, c, c])
    assert np.array_equal(V[n + p:, c, c], V[p:2 * p, c, c])
    assert np.array_equal(V[c, 0:p, c], V[c, n:n + p, c])


# ---- Jacobi Crank-Nicolson = dense implicit solve ---------------------------

def dense_cn(g, shape_core, coef, kinds_vals):
    """(I - c L2) T = rhs with per-wall Dirichlet/Neumann ghost rules -> A, shift"""
    import itertools
    N = int(np.prod(shape_core))
    A = np.eye(N)
    shift = np.zeros(N)
    idx = np.arange(N).reshape(shape_core)
    for flat, pos in enumerate(itertools.product(*[range(k) for k in shape_core])):
        for ax in range(3):
            n = shape_core[ax]
            m = O.field_metrics(g, ax)
            i = pos[ax]
            h = 1.0 / n
            c2, c1 = m["xix2"][i + g.pad] / h ** 2, m["xixx"][i + g.pad] * (0.5 / h)
            A[flat, flat] -= coef * (-2.0 * c2)
            for off, wgt in ((1, c2 + c1), (-1, c2 - c1)):
                j = i + off
                q = list(pos)
                if 0 <= j < n:
                    q[ax] = j
                    A[flat, idx[tuple(q)]] -= coef * wgt
                elif g.periodic[ax]:
                    q[ax] = j % n
                    A[flat, idx[tuple(q)]] -= coef * wgt
                else:
                    wall = 2 * ax + (1 if off > 0 else 0)
                    kind, val = kinds_vals[0][wall], kinds_vals[1][wall]
                    if kind == FD:          # ghost = 2v - C
                        A[flat, flat] -= coef * wgt * (-1.0)
                        shift[flat] += coef * wgt * 2 * val
                    else:                   # ghost = C - v h_wall
                        xc = m["x"]
                        hw = 2 * xc[g.pad] if off < 0 else 2 * (xc[n - 1 + g.pad] - g.len[ax])
                        A[flat, flat] -= coef * wgt
                        shift[flat] -= coef * wgt * val * hw
    return A, shift


def test_jacobi_cn_equals_dense_implicit_solve():
    n, p = 6, 2
    kappa, dt = 0.05, 2e-3
    tbc = ([FN, FD, FN, FN, FD, FD], [0.3, 0.5, -0.2, 0.1, 1.0, 0.0])
    g, d, s = make(n, periodic=(0, 0, 0), gtype=(1, 2, 2), beta=(1.1, 1.3, 1.5), has_T=1,
                   nu=0.0, kappa=kappa, dt=dt, cn_tol=1e-14, buoy=(0, 0, 0), tbc=tbc, nlscheme=1)
    rng = np.random.default_rng(9)
    T0 = rng.standard_normal(s.shape)
    zero = np.zeros(s.shape)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=T0, pp=zero)
    rt = s.rhs(0)["rt"]
    s.set(t=T0)
    assert s.stage(0) == 0
    T1 = s.get()["t"]
    beta0 = 4.0 / 15.0
    A, shift = dense_cn(g, (n, n, n), beta0 * dt * kappa, tbc)
    sol = np.linalg.solve(A, core(rt, p).ravel() + shift).reshape(n, n, n)
    assert np.abs(core(T1, p) - sol).max() < 1e-11
    assert s.jacobi_iters()[3] >= 1


# ---- RK3 stage: exact discrete heat-equation amplification --------------------

def test_rk3_heat_mode_amplification():
    """V = 0, periodic box, dScheme 1: T = cos(2 pi k.x) is an eigenvector of
    both the explicit and the implicit operator with eigenvalue
    lam = -sum_d 4 sin^2(pi k_d h)/h^2.  Each sub-stage multiplies it by
    (1 + alpha dt kappa lam) / (1 - beta dt kappa lam) with
    alpha = beta = (4/15, 1/15, 1/6)."""
    n, p = 16, 1
    kappa, dt = 0.02, 5e-3
    k = (1, 2, 3)
    g, d, s = make(n, periodic=(1, 1, 1), pad=p, dscheme=1, nlscheme=1, has_T=1, nu=0.0,
                   kappa=kappa, dt=dt, cn_tol=1e-13, buoy=(0, 0, 0))
    x = xi_axis(n, p)
    sh = s.shape
    X, Y, Z = (along(x, a, sh) for a in range(3))
    T0 = np.cos(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z))
    zero = np.zeros(sh)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=T0, pp=zero)
    assert s.step() == 0
    h = 1.0 / n
    lam = -sum(4 * np.sin(np.pi * kd * h) ** 2 / h ** 2 for kd in k)
    G = 1.0
    for a in (4 / 15, 1 / 15, 1 / 6):
        G *= (1 + a * dt * kappa * lam) / (1 - a * dt * kappa * lam)
    T1 = s.get()["t"]
    assert np.abs(core(T1, p) - G * core(T0, p)).max() < 1e-11
    # the same run with a wrong beta (e.g. gamma instead) would give a visibly different factor
    G_wrong = 1.0
    for a, b in zip((4 / 15, 1 / 15, 1 / 6), (8 / 15, 5 / 12, 3 / 4)):
        G_wrong *= (1 + a * dt * kappa * lam) / (1 - b * dt * kappa * lam)
    assert abs(G - G_wrong) > 1e-4


# ---- full step: Taylor-Green decay -----------------------------------------------

def test_taylor_green_2d_decay():
    """u = sin(2 pi x) cos(2 pi z), w = -cos(2 pi x) sin(2 pi z), v = 0 is an
    exact Navier-Stokes solution with KE ~ exp(-2 (2pi)^2 2 nu t)."""
    n, p = 24, 2
    nu, dt, steps = 0.01, 2e-3, 5
    g, d, s = make(n, periodic=(1, 1, 1), has_T=0, nu=nu, kappa=0.0, dt=dt, cn_tol=1e-12, nlscheme=3,
                   mg=O.mg_params(vc_count=30, vc_tol=1e-12, solve_flag=0, pre=3, post=3))
    x = xi_axis(n, p)
    sh = s.shape
    X, Z = along(x, 0, sh), along(x, 2, sh)
    tp = 2 * np.pi
    u = np.sin(tp * X) * np.cos(tp * Z)
    w = -np.cos(tp * X) * np.sin(tp * Z)
    P = 0.25 * (np.cos(2 * tp * X) + np.cos(2 * tp * Z))
    s.set(vx=u, vy=np.zeros(sh), vz=w, p=P, t=np.zeros(sh), pp=np.zeros(sh))
    ke0 = s.diag()["ke"]
    assert abs(ke0 - 0.25) < 1e-12
    for _ in range(steps):
        assert s.step() == 0
    ke = s.diag()["ke"]
    exact = ke0 * np.exp(-2 * tp ** 2 * 2 * nu * steps * dt)
    decay = ke0 - exact
    assert abs(ke - exact) < 0.02 * decay
    assert s.diag()["divmax"] < 0.05


# ---- diagnostics -----------------------------------------------------------------

def test_diagnostics_conduction_and_uniform_flow():
    n, p = 16, 2
    g, d, s = make(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), has_T=1,
                   kappa=1e-3, nu=1e-3)
    z = O.field_metrics(g, 2)["x"]
    sh = s.shape
    T = along(1.0 - z, 2, sh)
    U = np.full(sh, 0.3)
    zero = np.zeros(sh)
    s.set(vx=U, vy=zero, vz=zero, p=zero, t=T, pp=zero)
    dg = s.diag()
    assert dg["nuvol"] == 1.0
    assert abs(dg["nuwall"] - 1.0) < 1e-3
    assert abs(dg["ke"] - 0.5 * 0.09) < 0.5 * 0.09 * 1e-3
    # a constant x-velocity is divergence-free except at the first/last z cells where the
    # no-slip ghosts (2*0 - 0.3) do not enter: d/dx of a constant is exactly zero
    assert dg["divmax"] == 0.0


def test_thermal_energy_conduction_and_offset():
    """tseries.cc:410-446: the thermal energy of the conduction profile
    T = 1 - z is zero (theta = T + z - 1 vanishes to rounding), and a uniform
    offset c gives 1/2 c^2 (the cell volumes dXi / xi_x sum to the box volume
    to the metric's accuracy, as for the kinetic energy above)."""
    n, p = 16, 2
    g, d, s = make(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), has_T=1,
                   kappa=1e-3, nu=1e-3)
    z = O.field_metrics(g, 2)["x"]
    sh = s.shape
    zero = np.zeros(sh)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=along(1.0 - z, 2, sh), pp=zero)
    assert s.thermal_energy() < 1e-30
    c = 0.2
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=along(1.0 - z + c, 2, sh), pp=zero)
    assert abs(s.thermal_energy() - 0.5 * c * c) < 0.5 * c * c * 1e-3


def test_max_iterations_formula():
    g, d, s = make(16)
    assert s.max_iterations == int(np.log(16 ** 3) ** 3)


def test_euler_cn_heat_mode_amplification():
    """eulerCN_d3: explicit half of the diffusion, implicit half (beta 1/2),
    so a Fourier mode is multiplied by (1 + dt kappa lam / 2) / (1 - dt kappa lam / 2)
    per step."""
    n, p = 16, 1
    kappa, dt = 0.02, 5e-3
    k = (2, 1, 3)
    g, d, s = make(n, periodic=(1, 1, 1), pad=p, dscheme=1, nlscheme=1, has_T=1, nu=0.0,
                   kappa=kappa, dt=dt, cn_tol=1e-13, buoy=(0, 0, 0), scheme=1)
    x = xi_axis(n, p)
    sh = s.shape
    X, Y, Z = (along(x, a, sh) for a in range(3))
    T0 = np.cos(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z))
    zero = np.zeros(sh)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=T0, pp=zero)
    assert s.step() == 0
    assert s.step() == 0
    h = 1.0 / n
    lam = -sum(4 * np.sin(np.pi * kd * h) ** 2 / h ** 2 for kd in k)
    G = (1 + 0.5 * dt * kappa * lam) / (1 - 0.5 * dt * kappa * lam)
    T2 = s.get()["t"]
    assert np.abs(core(T2, p) - G * G * core(T0, p)).max() < 1e-11


def test_euler_cn_taylor_green_decay():
    n, p = 24, 2
    nu, dt, steps = 0.01, 2e-3, 5
    g, d, s = make(n, periodic=(1, 1, 1), has_T=0, nu=nu, kappa=0.0, dt=dt, cn_tol=1e-12, nlscheme=3, scheme=1,
                   mg=O.mg_params(vc_count=30, vc_tol=1e-12, solve_flag=0, pre=3, post=3))
    x = xi_axis(n, p)
    sh = s.shape
    X, Z = along(x, 0, sh), along(x, 2, sh)
    tp = 2 * np.pi
    u = np.sin(tp * X) * np.cos(tp * Z)
    w = -np.cos(tp * X) * np.sin(tp * Z)
    P = 0.25 * (np.cos(2 * tp * X) + np.cos(2 * tp * Z))
    s.set(vx=u, vy=np.zeros(sh), vz=w, p=P, t=np.zeros(sh), pp=np.zeros(sh))
    ke0 = s.diag()["ke"]
    for _ in range(steps):
        assert s.step() == 0
    ke = s.diag()["ke"]
    exact = ke0 * np.exp(-2 * tp ** 2 * 2 * nu * steps * dt)
    assert abs(ke - exact) < 0.03 * (ke0 - exact)


def _upwind_case(dscheme, nu, bias, u0, poly):
    n, p = 12, 2
    g, d, s = make(n, periodic=(0, 0, 0), gtype=(2, 0, 1), beta=(1.2, 1.0, 0.8), dscheme=dscheme, nlscheme=2,
                   nu=nu, kappa=nu, central_bias=bias, pe_limit=2.0)
    xi = xi_axis(n, p)
    sh = s.shape
    X = along(xi, 0, sh)
    zero = np.zeros(sh)
    s.set(vx=np.full(sh, u0), vy=zero, vz=zero, p=zero, t=zero, pp=zero)
    F = poly(X)
    H = s.op(7, F)
    m = O.field_metrics(g, 0)["xi_x"]
    exp = -u0 * along(m, 0, sh) * poly.deriv(1)(X)
    return H, exp, p, n


@pytest.mark.parametrize("u0", [0.7, -0.7])
def test_upwind_second_order_biased_exact_on_quadratics(u0):
    """Pure upwinding (Central Bias 0) at high Peclet number: the one-sided
    second-order stencils (3F - 4F- + F--)/2h and (-F++ + 4F+ - 3F)/2h
    are exact for quadratics (interior cells)."""
    H, exp, p, n = _upwind_case(2, 1e-9, 0.0, u0, POLY2)
    sl = (slice(p + 2, -p - 2),) + (slice(p, -p),) * 2
    assert np.abs(H[sl] - exp[sl]).max() < 1e-9


def test_upwind_first_order_exact_on_linears():
    H, exp, p, n = _upwind_case(1, 1e-9, 0.0, 0.9, np.polynomial.Polynomial([0.4, -1.3]))
    sl = (slice(p, -p),) * 3
    assert np.abs(H[sl] - exp[sl]).max() < 1e-9


def test_upwind_low_peclet_is_central():
    """Below the Peclet limit the scheme is the central one: 4th order
    interior / 2nd order edges, the same operator as the central NL scheme."""
    n, p = 12, 2
    rng = np.random.default_rng(3)
    F = rng.standard_normal((n + 2 * p,) * 3)
    V = {k: 0.01 * rng.standard_normal((n + 2 * p,) * 3) for k in ("vx", "vy", "vz")}
    out = []
    for nl in (2, 1):
        g, d, s = make(n, periodic=(1, 0, 0), gtype=(0, 2, 1), beta=(1, 1.1, 0.9), dscheme=2, nlscheme=nl,
                       nu=10.0, kappa=10.0)
        s.set(p=np.zeros_like(F), t=np.zeros_like(F), pp=np.zeros_like(F), **V)
        out.append(core(s.op(7, F), p))
    assert np.abs(out[0] - out[1]).max() < 1e-12 * max(1.0, np.abs(out[1]).max())


def test_upwind_full_bias_is_second_order_central():
    """Central Bias 1 turns the biased stencil into (F+ - F-)/2h."""
    n, p = 12, 2
    H, _, _, _ = _upwind_case(2, 1e-9, 1.0, 0.7, POLY4)
    g = O.grid_desc(n, periodic=(0, 0, 0), gtype=(2, 0, 1), beta=(1.2, 1.0, 0.8), pad=p)
    m = O.field_metrics(g, 0)["xi_x"]
    xi = xi_axis(n, p)
    F = POLY4(xi)
    c2 = -0.7 * m[p:-p] * (F[p + 1:n + p + 1] - F[p - 1:n + p - 1]) * (n / 2.0)
    got = H[p:-p, p + 3, p + 3]
    assert np.abs(got - c2).max() < 1e-9


def test_cfl_time_step():
    n, p = 8, 2
    g, d, s = make(n)
    sh = s.shape
    rng = np.random.default_rng(2)
    V = {k: rng.standard_normal(sh) for k in ("vx", "vy", "vz")}
    s.set(p=np.zeros(sh), t=np.zeros(sh), pp=np.zeros(sh), **V)
    c = (slice(p, -p),) * 3
    mx = (np.abs(V["vx"][c]) * n + np.abs(V["vy"][c]) * n + np.abs(V["vz"][c]) * n).max()
    assert abs(s.cfl(0.1, 1.0) - 0.1 / mx) < 1e-15
    assert s.cfl(0.1, 1e-5) == 1e-5


# ---- stretched spiral vortex LES (lib/les/spiral.cc) ---------------------------

def _tg_les(n=16, seed=5, noise=0.01):
    p = 2
    g, d, s = make(n, periodic=(1, 1, 1), has_T=0, nu=1.0 / 1600, dt=1e-3, les=1)
    x = xi_axis(n, p)
    sh = s.shape
    X, Y, Z = (along(x, a, sh) for a in range(3))
    tp = 2 * np.pi
    rng = np.random.default_rng(seed)
    u = np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z) + noise * rng.uniform(-1, 1, sh)
    v = -np.cos(tp * X) * np.sin(tp * Y) * np.cos(tp * Z) + noise * rng.uniform(-1, 1, sh)
    w = noise * rng.uniform(-1, 1, sh)
    s.set(vx=u, vy=v, vz=w, p=np.zeros(sh), t=np.zeros(sh), pp=np.zeros(sh))
    return g, s, p


def test_les_energy_and_structure_integrals_match_closed_forms():
    """keIntegral ~ (1/2) k^(2/3) Gamma(-1/3, k^2) (stated max error 0.17 %)
    and sfIntegral ~ int_0^1 4 x^(-5/3) (1 - J0(pi d x)) dx (stated 2.71 %),
    checked through single-cell K values: K / (F2/Qd) is keIntegral(kc lv)."""
    import scipy.special as sp
    from scipy.integrate import quad

    def ke_exact(k):
        x = k * k
        g23 = sp.gammaincc(2.0 / 3.0, x) * sp.gamma(2.0 / 3.0)
        gm13 = (g23 - x ** (-1.0 / 3.0) * np.exp(-x)) / (-1.0 / 3.0)
        return 0.5 * k ** (2.0 / 3.0) * gm13

    def sf_exact(d):
        return quad(lambda x: 4 * x ** (-5.0 / 3.0) * (1 - sp.j0(np.pi * d * x)), 0, 1, limit=200)[0]

    g, s, p = _tg_les()
    r, H, dg = s.sgs()
    assert r == 0
    T = s.stresses()
    n = 16
    c = (slice(p, -p),) * 3
    # at every cell T is K times the projector orthogonal to the vortex axis e:
    # trace T = 2K, T e = 0 with e the most extensive strain direction
    K = 0.5 * (T[0] + T[1] + T[2])
    assert (K[c] > 0).all()
    fields = s.get()
    for (i, j, k) in [(3, 5, 7), (0, 0, 0), (15, 8, 2), (9, 14, 11)]:
        dudx = np.array([[s.op(a, fields[comp])[i + p, j + p, k + p] for a in range(3)]
                         for comp in ("vx", "vy", "vz")])
        S = 0.5 * (dudx + dudx.T)
        w, vec = np.linalg.eigh(S)
        e = vec[:, 2]
        Tm = np.array([[T[0], T[3], T[5]], [T[3], T[1], T[4]], [T[5], T[4], T[2]]])[:, :, i + p, j + p, k + p]
        Kc = K[i + p, j + p, k + p]
        assert np.abs(Tm @ e).max() < 1e-9 * Kc
        assert np.allclose(Tm, Kc * (np.eye(3) - np.outer(e, e)), rtol=0, atol=1e-9 * Kc)
        # K = (F2 / Qd) keIntegral(kc lv) with the 26-neighbour structure function
        h = 1.0 / n
        nb = [(a, b, cc) for a in (-1, 0, 1) for b in (-1, 0, 1) for cc in (-1, 0, 1) if (a, b, cc) != (0, 0, 0)]
        F2 = np.mean([sum((fields[f][i + p + a, j + p + b, k + p + cc] - fields[f][i + p, j + p, k + p]) ** 2
                          for f in ("vx", "vy", "vz")) for a, b, cc in nb])
        Qd = np.mean([sf_exact(np.sqrt(max(h * h * (a * a + b * b + cc * cc) - (h * (a * e[0] + b * e[1] + cc * e[2])) ** 2, 0.0)) / h)
                      for a, b, cc in nb])
        axial = e @ S @ e
        lv = np.sqrt(2.0 * (1.0 / 1600) / (3.0 * (abs(axial) + 2e-3)))
        K_exact = F2 / Qd * ke_exact(np.pi / h * lv)
        assert abs(Kc - K_exact) < 0.04 * K_exact


def test_les_divergence_of_stress_and_diagnostics():
    g, s, p = _tg_les(seed=6)
    r, H, dg = s.sgs()
    assert r == 0
    T = s.stresses()
    c = (slice(p, -p),) * 3
    # H -= div T with the field's own first derivative operator
    Txx, Tyy, Tzz, Txy, Tyz, Tzx = T
    exp = [-(s.op(0, Txx) + s.op(1, Txy) + s.op(2, Tzx)),
           -(s.op(0, Txy) + s.op(1, Tyy) + s.op(2, Tyz)),
           -(s.op(0, Tzx) + s.op(1, Tyz) + s.op(2, Tzz))]
    for q in range(3):
        assert np.allclose(H[q][c], exp[q][c], rtol=1e-12, atol=1e-14 * np.abs(exp[q][c]).max())
    # diagnostics: sum |K| dV (dV = h^3 on a uniform unit box) and the
    # dissipation -S:T summed over cells
    K = 0.5 * (Txx + Tyy + Tzz)
    assert abs(dg[0] - np.abs(K[c]).sum() / 16 ** 3) < 1e-12 * dg[0]
    assert dg[1] != 0.0 and dg[2] > 0.0


def test_les_starts_after_five_steps():
    """The model is added once solTime > 5 tStp (lsRK3_d3.cc:238); before
    that a LES run and a plain run are identical."""
    g, s1, p = _tg_les(n=8, seed=7)
    _, d0, s0 = make(8, periodic=(1, 1, 1), has_T=0, nu=1.0 / 1600, dt=1e-3, les=0)
    f = s1.get()
    s0.set(**f)
    for step in range(7):
        assert s1.step() == 0 and s0.step() == 0
        same = all(np.array_equal(s1.get()[k], s0.get()[k]) for k in ("vx", "vy", "vz"))
        # time after step q is (q+1) dt; the model enters the step that starts with time > 5 dt
        assert same == (step <= 5), step


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler-cn"])
def test_channel_poiseuille_closed_form(scheme):
    """Force type 5 (constantPGrad, force.h:148-153) driving the channel of the
    reference's channelTest deck: x and y periodic, no-slip z walls, Re 10 so
    nu = 0.1, mean pressure gradient G = 1, Lz = 2.  On a uniform z grid the
    steady state of the discrete equations, nu D2 u = -G with the no-slip
    ghost u(-1) = -u(0), is the Poiseuille parabola raised by G h^2 / (8 nu):
    the second difference is exact on quadratics, and the ghost rule puts the
    zero at the mean of the two cells beside the wall rather than on the
    wall.  A dropped forcing, a wrong sign, a forcing on another component or
    a missing (alpha + beta) weight lands elsewhere."""
    n, p, L, nu, G = (4, 4, 16), 1, 2.0, 0.1, 1.0
    g = O.grid_desc(n, length=(2.0, 2.0, L), periodic=(1, 1, 0), pad=p)
    d = O.nse_desc(g, dscheme=1, nlscheme=1, has_T=0, nu=nu, kappa=nu, dt=0.5, cn_tol=1e-13, scheme=scheme,
                   vbc=[([FD] * 6, [0.0] * 6)] * 3, pbc=([FN] * 6, [0.0] * 6), mean_pgrad=G,
                   mg=O.mg_params(vc_count=5, vc_tol=1e-12, solve_flag=0, pre=2, post=2))
    s = O.Nse(d)
    zero = np.zeros(s.shape)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=zero, pp=zero)
    for _ in range(300):
        assert s.step() == 0
    f = s.get()
    h = L / n[2]
    z = (np.arange(n[2]) + 0.5) * h
    u = G / (2 * nu) * (z * (L - z) + h * h / 4)
    assert np.abs(core(f["vx"], p) - u[None, None, :]).max() < 1e-9 * u.max()
    assert np.abs(core(f["vy"], p)).max() == 0 and np.abs(core(f["vz"], p)).max() == 0


def plume_desc(n=(8, 8, 8), t_bot=None, radius=0.3, scheme=0):
    """The plume problem's walls (probType 9, hydro.cc:287-440, scalar.cc:129-191):
    no-slip Vx, Vy; Vz zero-gradient on the x walls and the top (outflow),
    zero on the y walls and the bottom; T zero-gradient on the side walls,
    the hot plate (or t_bot) at the bottom, plumeOut at the top."""
    g = O.grid_desc(n, periodic=(0, 0, 0), pad=1)
    vz_t = [FN, FN, FD, FD, FD, FN]
    vbc = [([FD] * 6, [0.0] * 6), ([FD] * 6, [0.0] * 6), (vz_t, [0.0] * 6)]
    tb = t_bot or (O.FBC_HOT_PLATE, 0.0)
    tbc = ([FN] * 4 + [tb[0], O.FBC_PLUME_OUT], [0.0] * 4 + [tb[1], 0.0])
    return g, O.nse_desc(g, dscheme=1, nlscheme=1, has_T=1, nu=1e-2, kappa=1e-2, dt=2e-3, cn_tol=1e-10,
                         scheme=scheme, vbc=vbc, tbc=tbc, plate_radius=radius,
                         mg=O.mg_params(vc_count=20, vc_tol=1e-10, solve_flag=0, pre=3, post=3))


def plume_fields(shape, p, seed=3):
    rng = np.random.default_rng(seed)
    c = (slice(p, -p),) * 3
    f = {k: np.zeros(shape) for k in ("vx", "vy", "vz", "p", "t", "pp")}
    for k in ("vx", "vy", "vz"):
        f[k][c] = 0.3 * rng.standard_normal(core(f[k], p).shape)
    f["t"][c] = rng.uniform(0.0, 1.0, core(f["t"], p).shape)
    return f


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler-cn"])
def test_hot_plate_is_dirichlet_or_neumann_at_extreme_radii(scheme):
    """hotPlate (hotPlate.cc:74-123) puts T = 1 (2 - F(edge)) on the disc and
    a zero gradient off it: a disc that covers the wall gives the run of
    dirichlet(1), an empty disc the run of neumann(0), bit for bit."""
    for radius, plain in ((1e9, (FD, 1.0)), (-1.0, (FN, 0.0))):
        _, d1 = plume_desc(radius=radius, scheme=scheme)
        _, d2 = plume_desc(t_bot=plain, scheme=scheme)
        a, b = O.Nse(d1), O.Nse(d2)
        f = plume_fields(a.shape, 1)
        a.set(**f)
        b.set(**f)
        for _ in range(2):
            assert a.step() == 0 and b.step() == 0
        fa, fb = a.get(), b.get()
        for k in ("vx", "vy", "vz", "p", "t"):
            assert np.array_equal(core(fa[k], 1), core(fb[k], 1)), (radius, k)


def test_hot_plate_disc_and_plume_outflow_ghosts():
    """Ghost layers after the initial imposeBCs and after a step: bottom
    ghosts 2 - T(0) on the disc of radius 0.3 about the wall centre (the
    distance measured to the ghost cell, z = -h/2) and T(0) off it; top
    ghosts T(n-1) where the ghost Vz (= Vz(n-1), zero gradient) is positive,
    0 where it is not."""
    n, p, R = 16, 1, 0.3
    g, d = plume_desc(n=(n, n, n), radius=R)
    s = O.Nse(d)
    f = plume_fields(s.shape, p, seed=9)
    s.set(**f)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    disc = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (0.5 / n) ** 2) <= R
    assert 0 < disc.sum() < n * n
    for step in range(2):
        o = s.get()
        T, Vz = o["t"][p:-p, p:-p], o["vz"][p:-p, p:-p]
        bot = np.where(disc, 2.0 - T[:, :, p], T[:, :, p])
        assert np.array_equal(T[:, :, p - 1], bot), step
        top = np.where(Vz[:, :, -p] > 0, T[:, :, -p - 1], 0.0)
        assert np.array_equal(Vz[:, :, -p], Vz[:, :, -p - 1])
        assert np.array_equal(T[:, :, -p], top), step
        assert 0 < (Vz[:, :, -p] > 0).sum() < n * n
        assert s.step() == 0


def test_open_top_wall_nusselt():
    """tseries for an open top (probType 9, tseries.cc:114-128): the bottom
    wall alone, at its own distance z(0), over one wall's area.  With T = 0
    the flux density is 1 / wallDn everywhere on the bottom and 0 on the top:
    NuWall = 1 / z(0) with the open top, 1 / (2 (Lz - z(n-1))) with both
    walls; a single-sided tanh z grid keeps the two distances apart."""
    n, p = 8, 1
    for top, want in ((O.FBC_PLUME_OUT, "open"), (O.FBC_DIRICHLET, "both")):
        g = O.grid_desc((n, n, n), periodic=(1, 1, 0), gtype=(0, 0, 1), beta=(1.0, 1.0, 1.3), pad=p)
        tbc = ([FN] * 4 + [FD, top], [0.0] * 4 + [1.0, 0.0])
        d = O.nse_desc(g, dscheme=1, nlscheme=1, has_T=1, tbc=tbc,
                       mg=O.mg_params(vc_count=5, vc_tol=1e-10, solve_flag=0, pre=2, post=2))
        s = O.Nse(d)
        zero = np.zeros(s.shape)
        s.set(vx=zero, vy=zero, vz=zero, p=zero, t=zero, pp=zero)
        z = O.field_metrics(g, 2)["x"][p:-p]
        assert abs(z[0] - (1.0 - z[-1])) > 1e-3
        expect = 1.0 / z[0] if want == "open" else 1.0 / (2 * (1.0 - z[-1]))
        assert abs(s.diag()["nuwall"] - expect) < 1e-12 * expect, (want, s.diag()["nuwall"], expect)


# ---- vector operators (vfield::divergence, sfield::gradient, computeDiff, computeNLin)

@pytest.mark.parametrize("dscheme", [1, 2])
def test_vector_operators_closed_forms(dscheme):
    """On a uniform grid with linear fields extended into the pads: the
    divergence of u = (a x, b y, c z) is a + b + c; the gradient of a
    quadratic is exact; computeDiff of a quadratic velocity adds its constant
    Laplacian to the accumulated output; computeNLin of u = (x, 0, 0)
    subtracts u du/dx = x (central) and the skew-symmetric
    1/2 (u du/dx + d(u u)/dx) = 3x/2 (Morinishi), both exact on these
    polynomials.  The operators accumulate (2, 3) or assign (0, 1) as the
    reference's methods do."""
    n, p = 12, 2
    for nl in (1, 3):
        g, d, s = make(n, periodic=(0, 0, 0), dscheme=dscheme, nlscheme=nl)
        sh = s.shape
        X, Y, Z = (along((np.arange(-p, n + p) + 0.5) / n, a, sh) for a in range(3))
        c = (slice(p, -p),) * 3
        div = s.vec_op(0, 2.0 * X, -3.0 * Y, 0.5 * Z)[0]
        assert np.allclose(div[c], -0.5, rtol=0, atol=1e-11)
        F = 1.0 + X ** 2 - 2.0 * Y + 0.5 * Z ** 2
        gx, gy, gz = s.vec_op(1, F)
        assert np.allclose(gx[c], 2 * X[c], atol=1e-11) and np.allclose(gy[c], -2.0, atol=1e-11)
        assert np.allclose(gz[c], Z[c], atol=1e-11)
        acc = [np.full(sh, 0.25), np.full(sh, -1.0), np.zeros(sh)]
        lap = s.vec_op(2, X ** 2, Y ** 2 + Z ** 2, np.zeros(sh), out=acc)
        assert np.allclose(lap[0][c], 0.25 + 2.0, atol=1e-9) and np.allclose(lap[1][c], -1.0 + 4.0, atol=1e-9)
        assert np.allclose(lap[2][c], 0.0, atol=1e-9)
        zero = np.zeros(sh)
        h = s.vec_op(3, X, zero, zero, out=[np.full(sh, 1.0), zero, zero])
        want = 1.0 - (1.5 if nl == 3 else 1.0) * X
        assert np.allclose(h[0][c], want[c], atol=1e-10), (nl, np.abs(h[0][c] - want[c]).max())
        assert np.abs(h[1][c]).max() == 0 and np.abs(h[2][c]).max() == 0


def test_vector_operators_agree_with_scalar_operators():
    """Bit for bit: the gradient's components are calcDerivative1_x/y/z of the
    field, computeDiff of a component is sfield::computeDiff of it, and the
    divergence is ((0 + dVx/dx) + dVy/dy) + dVz/dz of those derivatives."""
    n, p = 10, 2
    g, d, s = make((n, 8, 12), periodic=(1, 0, 1), gtype=(0, 2, 1), beta=(1.0, 1.1, 0.7), dscheme=2, nlscheme=3)
    rng = np.random.default_rng(3)
    f = [rng.standard_normal(s.shape) for _ in range(3)]
    c = (slice(p, -p),) * 3
    grad = s.vec_op(1, f[0])
    for a in range(3):
        assert np.array_equal(grad[a][c], s.op(a, f[0])[c])
    div = s.vec_op(0, *f)[0]
    want = ((0.0 + s.op(0, f[0])) + s.op(1, f[1])) + s.op(2, f[2])
    assert np.array_equal(div[c], want[c])
    lap = s.vec_op(2, *f)
    for q in range(3):
        assert np.array_equal(lap[q][c], s.op(6, f[q])[c])


# ---- the PLANAR build (lsRK3_d2 / eulerCN_d2 with the PLANAR branches) ------

def _tg2(dim, n=16, ny=4, dscheme=1, nlscheme=1, scheme=0, nu=0.05, dt=2e-3, steps=6):
    """2-D Taylor-Green vortex in a periodic (2 pi)^2 box: as a PLANAR run
    (dim 2), or as a 3-D run whose fields are uniform along y (dim 3)."""
    pad = dscheme
    L = 2 * np.pi
    g = O.grid_desc((n, 1 if dim == 2 else ny, n), dim=dim, periodic=(1, 1, 1),
                    length=(L, L if dim == 3 else 1.0, L), pad=pad)
    d = O.nse_desc(g, mg=O.mg_params(vc_count=30, vc_tol=1e-12, solve_flag=0, pre=3, post=3), dscheme=dscheme,
                   nlscheme=nlscheme, has_T=0, nu=nu, kappa=nu, dt=dt, cn_tol=1e-13, scheme=scheme)
    s = O.Nse(d)
    sh = s.shape
    x = (np.arange(-pad, n + pad) + 0.5) * (L / n)
    u = np.broadcast_to((np.sin(x)[:, None] * np.cos(x)[None, :])[:, None, :], sh).copy()
    w = np.broadcast_to((-np.cos(x)[:, None] * np.sin(x)[None, :])[:, None, :], sh).copy()
    zero = np.zeros(sh)
    s.set(vx=u, vy=zero, vz=w, p=zero, t=zero, pp=zero)
    ke = [s.diag()["ke"]]
    for _ in range(steps):
        assert s.step() == 0
        ke.append(s.diag()["ke"])
    return s, np.array(ke)


@pytest.mark.parametrize("dscheme,nlscheme,scheme", [(1, 1, 0), (2, 3, 0), (2, 2, 1), (1, 3, 1)])
def test_planar_taylor_green_decay_and_y_uniform_3d(dscheme, nlscheme, scheme):
    """The PLANAR restatement: the 2-D Taylor-Green vortex, an exact
    Navier-Stokes solution, decays as KE ~ exp(-4 nu t); the planar run agrees
    with a 3-D run of the same flow made uniform along y (whose y terms are
    exact zeros, its pressure iterate y-uniform to rounding); a planar run
    carries one y row with no pads and no Vy solve.  Morinishi with dScheme 2
    is not compared with 3-D: its z edge test reads core.ubound(1)
    (vfield.cc:329, 434), which is 0 in a PLANAR build and ny - 1 in 3-D."""
    nu, dt, steps = 0.05, 2e-3, 6
    s2, ke2 = _tg2(2, dscheme=dscheme, nlscheme=nlscheme, scheme=scheme, nu=nu, dt=dt, steps=steps)
    s3, ke3 = _tg2(3, dscheme=dscheme, nlscheme=nlscheme, scheme=scheme, nu=nu, dt=dt, steps=steps)
    assert s2.shape[1] == 1
    assert s2.jacobi_iters()[1] == 0
    t = dt * np.arange(steps + 1)
    assert np.max(np.abs(ke2 / (ke2[0] * np.exp(-4 * nu * t)) - 1)) < 2e-3
    if dscheme == 2 and nlscheme == 3:
        return
    assert np.max(np.abs(ke2 / ke3 - 1)) < 1e-10
    f2, f3 = s2.get(), s3.get()
    p = dscheme
    for k in ("vx", "vz"):
        a, b = f2[k][p:-p, 0, p:-p], f3[k][p:-p, p + 1, p:-p]
        assert np.abs(a - b).max() < 1e-9 * np.abs(b).max(), k
    assert np.abs(f2["vy"]).max() == 0


def test_planar_jacobi_and_walls():
    """A planar box with walls in z and x periodic: the time step keeps Vy at
    zero, the wall ghosts of T follow their rules on the single y row, and
    conduction (T = 1 - z, u = 0: second differences of a linear profile are
    exact zeros on a uniform grid, its wall ghosts its linear extension) is a
    steady state of the planar scalar step to the Jacobi tolerance."""
    n, pad = 16, 1
    g = O.grid_desc((n, 1, n), dim=2, periodic=(1, 0, 0), pad=pad)
    d = O.nse_desc(g, mg=O.mg_params(vc_count=20, vc_tol=1e-12, solve_flag=0, pre=3, post=3), dscheme=1,
                   nlscheme=1, has_T=1, nu=1e-2, kappa=1e-2, dt=1e-3, cn_tol=1e-13, buoy=(0.0, 0.0, 0.0))
    s = O.Nse(d)
    z = O.field_metrics(g, 2)["x"]
    sh = s.shape
    T = np.broadcast_to(1.0 - z[None, None, :], sh).copy()
    zero = np.zeros(sh)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=T, pp=zero)
    for _ in range(3):
        assert s.step() == 0
    f = s.get()
    c = (slice(pad, -pad), 0, slice(pad, -pad))
    assert np.abs(f["t"][c] - T[c]).max() < 1e-9
    assert np.abs(f["vx"]).max() < 1e-12 and np.abs(f["vz"]).max() < 1e-12 and np.abs(f["vy"]).max() == 0
    # Dirichlet ghosts of T: 2 v - T(edge), v = 1 at the bottom and 0 at the top
    assert np.allclose(f["t"][pad:-pad, 0, 0], 2.0 - f["t"][pad:-pad, 0, 1], rtol=0, atol=1e-15)
    assert np.allclose(f["t"][pad:-pad, 0, -1], -f["t"][pad:-pad, 0, -2], rtol=0, atol=1e-15)


def inflow_desc(n=(16, 8, 8), pad=2, scheme=0, null=True):
    """The inflow problem's walls (probType 3, hydro.cc:349-484): Vx = 1 on
    x-lo, zero gradient of every component on x-hi, no-slip y and z walls;
    P null on x-lo (or zero gradient when `null` is off), zero gradient
    elsewhere."""
    g = O.grid_desc(n, periodic=(0, 0, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.2), pad=pad)
    vbc = [([FD, FN, FD, FD, FD, FD], [1.0, 0.0, 0.0, 0.0, 0.0, 0.0])] + \
          [([FD, FN, FD, FD, FD, FD], [0.0] * 6)] * 2
    pbc = ([O.FBC_NULL if null else FN] + [FN] * 5, [0.0] * 6)
    return g, O.nse_desc(g, dscheme=2, nlscheme=3, has_T=0, nu=2e-2, dt=2e-3, cn_tol=1e-10, scheme=scheme,
                         vbc=vbc, pbc=pbc, mg=O.mg_params(vc_count=6, vc_tol=1e-10, solve_flag=0, pre=3, post=3))


def inflow_fields(shape, p, seed=4):
    rng = np.random.default_rng(seed)
    c = (slice(p, -p),) * 3
    f = {k: np.zeros(shape) for k in ("vx", "vy", "vz", "p", "t", "pp")}
    f["vx"][c] = 1.0 + 0.1 * rng.standard_normal(core(f["vx"], p).shape)
    for k in ("vy", "vz"):
        f[k][c] = 0.1 * rng.standard_normal(core(f[k], p).shape)
    f["p"][:] = rng.uniform(-1.0, 1.0, shape)   # ghosts too: a null wall keeps them
    return f


@pytest.mark.parametrize("scheme", [0, 1], ids=["rk3", "euler-cn"])
def test_null_pressure_wall_keeps_the_whole_array_sum(scheme):
    """nullBC (boundary.h:205-212) writes nothing, so the x-lo ghosts of P
    carry the caller's values plus every correction's ghost: after each
    stage, P(-1) = P(-1) before + Pp(-1), and Pp(-1) = Pp(0) (the solver's
    zero-gradient wall).  A zero-gradient P wall instead gives P(-1) = P(0),
    which differs; the x-hi wall does that in both runs."""
    p = 2
    _, d = inflow_desc(scheme=scheme)
    s = O.Nse(d)
    f = inflow_fields(s.shape, p)
    s.set(**f)
    c = slice(p, -p)
    ghost = lambda a: a[p - 1, c, c]   # noqa: E731
    before = ghost(s.get()["p"]).copy()
    assert np.array_equal(before, ghost(f["p"]))   # the initial imposeBCs leave it
    stages = 1 if scheme == 1 else 3
    for step in range(2):
        for st in range(stages):
            assert s.stage(st) == 0
            o = s.get()
            pp_g = ghost(o["pp"])
            assert np.array_equal(pp_g, o["pp"][p, c, c])
            assert np.array_equal(ghost(o["p"]), before + pp_g), (step, st)
            assert not np.array_equal(ghost(o["p"]), o["p"][p, c, c])
            assert np.array_equal(o["p"][-p, c, c], o["p"][-p - 1, c, c])
            before = ghost(o["p"]).copy()
    # the same run with a zero-gradient x-lo wall differs in the core: the
    # pressure gradient of the first cells reads the ghost
    _, d2 = inflow_desc(scheme=scheme, null=False)
    s2 = O.Nse(d2)
    s2.set(**f)
    for _ in range(2):
        assert s2.step() == 0
    assert not np.array_equal(core(s2.get()["vx"], p), core(s.get()["vx"], p))


@pytest.mark.parametrize("n,reads", [((8, 8, 8), False), ((8, 6, 8), True)], ids=["ny=nz", "ny<nz"])
def test_morinishi_z_edge_test_reads_second_top_layer_only_when_ny_ne_nz(n, reads):
    """Morinishi's z edge test compares k with core.ubound(1) (vfield.cc:329,
    434; sfield.cc:236, 295): with n_y == n_z the last z cell is an edge cell
    and the second ghost layer above the top wall is never read; with
    n_y != n_z that cell takes the 4th-order stencil and reads it.  The
    layer's values change the RHS of the last z cells alone, or nothing."""
    p = 2
    g, d, s = make(n, periodic=(1, 1, 0), pad=p, dscheme=2, nlscheme=3, has_T=1, nu=3e-3, kappa=2e-3)
    rng = np.random.default_rng(8)
    base = {k: rng.uniform(-0.3, 0.3, s.shape) for k in ("vx", "vy", "vz", "t")}
    base.update(p=np.zeros(s.shape), pp=np.zeros(s.shape))
    out = []
    for bump in (0.0, 0.5):
        f = {k: v.copy() for k, v in base.items()}
        for k in ("vx", "vy", "vz", "t"):
            f[k][:, :, -1] += bump
        s2 = O.Nse(d)
        s2.set(**f)
        out.append(s2.rhs(0))
    c = (slice(p, -p),) * 3
    for k in out[0]:
        a, b = out[0][k][c], out[1][k][c]
        diff = np.argwhere(a != b)
        if not reads:
            assert diff.size == 0, k
        elif k in ("hx", "hy", "hz", "ht", "rx", "ry", "rz", "rt"):
            assert diff.size > 0, k
            assert (diff[:, 2] == n[2] - 1).all(), k
