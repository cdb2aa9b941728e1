"""ctypes binding of the FP64 CPU oracle (oracle/_build/libsaras_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product package
(saras-cc_amd/) never imports it.

Parity status of each oracle function is pinned by tests/test_oracle_*.py:
  or_vc_depth, or_level_metrics, or_field_metrics  -> closed forms, tables, FD
  or_mg_* (3-D and 2-D V-cycle)                     -> dense brute force, closed
                                                       forms, the reference's
                                                       own r^2/6 mgTest solution
  or_nse_* (field stencils, Jacobi-CN, RK3 step)    -> polynomial exactness,
                                                       convergence order, dense
                                                       implicit solve, Fourier
                                                       decay, Taylor-Green
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsaras_oracle.so")

GRID_UNIFORM, GRID_TANH_SINGLE, GRID_TANH_DOUBLE = 0, 1, 2
BC_NEUMANN, BC_DIRICHLET, BC_DIRICHLET_TEST = 0, 1, 2


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("len", C.c_double * 3),
                ("periodic", C.c_int * 3), ("gtype", C.c_int * 3),
                ("beta", C.c_double * 3), ("pad", C.c_int)]


class MgParams(C.Structure):
    _fields_ = [("vc_count", C.c_int), ("vc_tol", C.c_double), ("solve_flag", C.c_int),
                ("mg_tol", C.c_double), ("sor", C.c_double), ("pre_smooth", C.c_int),
                ("post_smooth", C.c_int), ("res_type", C.c_int), ("bc", C.c_int),
                ("no_early_exit", C.c_int), ("top_dirichlet", C.c_int)]


FBC_DIRICHLET, FBC_NEUMANN, FBC_HOT_PLATE, FBC_PLUME_OUT, FBC_NULL = 0, 1, 2, 3, 4


class NseDesc(C.Structure):
    _fields_ = [("g", GridDesc), ("mg", MgParams), ("dscheme", C.c_int), ("nlscheme", C.c_int),
                ("has_T", C.c_int), ("nu", C.c_double), ("kappa", C.c_double), ("dt", C.c_double),
                ("cn_tol", C.c_double), ("buoy", C.c_double * 3),
                ("vbc_type", (C.c_int * 6) * 3), ("vbc_val", (C.c_double * 6) * 3),
                ("tbc_type", C.c_int * 6), ("tbc_val", C.c_double * 6),
                ("pbc_type", C.c_int * 6), ("pbc_val", C.c_double * 6), ("scheme", C.c_int),
                ("pe_limit", C.c_double), ("central_bias", C.c_double), ("les", C.c_int),
                ("mean_pgrad", C.c_double), ("plate_radius", C.c_double)]


class MgStats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("residual", C.c_double),
                ("coarse_iters_max", C.c_int), ("coarse_capped", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` first")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.or_vc_depth.argtypes = [C.POINTER(GridDesc)]
        L.or_vc_depth.restype = C.c_int
        L.or_level_metrics.argtypes = [C.POINTER(GridDesc), C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.or_field_metrics.argtypes = [C.POINTER(GridDesc), C.c_int, dp, dp, dp, dp]
        L.or_mg_create.argtypes = [C.POINTER(GridDesc), C.POINTER(MgParams)]
        L.or_mg_create.restype = C.c_void_p
        L.or_mg_destroy.argtypes = [C.c_void_p]
        L.or_mg_levels.argtypes = [C.c_void_p]
        L.or_mg_max_count.argtypes = [C.c_void_p]
        L.or_mg_level_shape.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.or_mg_solve.argtypes = [C.c_void_p, dp, dp, C.c_int, dp, C.POINTER(MgStats)]
        L.or_mg_set_level.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.or_mg_get_level.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.or_mg_smooth.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.or_mg_coarse_solve.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_mg_residual_restrict.argtypes = [C.c_void_p, C.c_int]
        L.or_mg_prolong_correct.argtypes = [C.c_void_p, C.c_int]
        L.or_mg_error.argtypes = [C.c_void_p, C.c_int]
        L.or_mg_error.restype = C.c_double
        L.or_mg_impose_bc.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_mg_vcycles.argtypes = [C.c_void_p, C.c_int]
        L.or_mg_residual.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.or_nse_create.argtypes = [C.POINTER(NseDesc)]
        L.or_nse_create.restype = C.c_void_p
        L.or_nse_destroy.argtypes = [C.c_void_p]
        L.or_nse_max_iterations.argtypes = [C.c_void_p]
        L.or_nse_set.argtypes = [C.c_void_p] + [dp] * 6
        L.or_nse_get.argtypes = [C.c_void_p] + [dp] * 6
        L.or_nse_rhs.argtypes = [C.c_void_p, C.c_int]
        L.or_nse_get_rhs.argtypes = [C.c_void_p] + [dp] * 8
        L.or_nse_stage.argtypes = [C.c_void_p, C.c_int]
        L.or_nse_step.argtypes = [C.c_void_p]
        L.or_nse_jacobi_iters.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.or_nse_mg_stats.argtypes = [C.c_void_p, C.POINTER(MgStats)]
        L.or_nse_op.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.or_nse_vec_op.argtypes = [C.c_void_p, C.c_int] + [dp] * 6
        L.or_nse_diag.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_nse_thermal_energy.argtypes = [C.c_void_p]
        L.or_nse_thermal_energy.restype = C.c_double
        L.or_nse_cfl.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.or_nse_sgs.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_nse_sgs_diag.argtypes = [C.c_void_p, dp]
        L.or_nse_get_sgs.argtypes = [C.c_void_p, dp]
        L.or_nse_time.argtypes = [C.c_void_p]
        L.or_nse_time.restype = C.c_double
        L.or_nse_cfl.restype = C.c_double
        L.or_nse_physics.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def grid_desc(n, dim=3, length=(1.0, 1.0, 1.0), periodic=(0, 0, 0), gtype=(0, 0, 0),
              beta=(1.0, 1.0, 1.0), pad=1):
    g = GridDesc()
    g.dim = dim
    n = tuple(n) if hasattr(n, "__len__") else (n, n, n)
    for a in range(3):
        g.n[a] = n[a]
        g.len[a] = length[a]
        g.periodic[a] = int(periodic[a])
        g.gtype[a] = int(gtype[a])
        g.beta[a] = beta[a]
    g.pad = pad
    return g


def mg_params(vc_count=20, vc_tol=1e-6, solve_flag=1, mg_tol=1e-2, sor=1.0, pre=3, post=6,
              res_type=0, bc=BC_NEUMANN, no_early_exit=0, top_dirichlet=0):
    p = MgParams()
    p.vc_count, p.vc_tol, p.solve_flag, p.mg_tol = vc_count, vc_tol, int(solve_flag), mg_tol
    p.sor, p.pre_smooth, p.post_smooth, p.res_type = sor, pre, post, res_type
    p.bc, p.no_early_exit = bc, int(no_early_exit)
    p.top_dirichlet = int(top_dirichlet)
    return p


def vc_depth(g):
    return lib().or_vc_depth(C.byref(g))


def level_metrics(g, level, axis):
    """Return dict of arrays over [-1, n_l] for one axis."""
    n0 = 1 if (g.dim == 2 and axis == 1) else g.n[axis]
    nl = n0 if (g.dim == 2 and axis == 1) else n0 >> level
    out = {k: np.zeros(nl + 2) for k in ("x", "xi", "xi_x", "xixx", "xix2")}
    r = lib().or_level_metrics(C.byref(g), level, axis, _dp(out["x"]), _dp(out["xi"]),
                               _dp(out["xi_x"]), _dp(out["xixx"]), _dp(out["xix2"]))
    if r < 0:
        raise ValueError("bad level")
    return out


def field_metrics(g, axis):
    p = 1 if (g.dim == 2 and axis == 1) else g.pad
    n = 1 if (g.dim == 2 and axis == 1) else g.n[axis]
    out = {k: np.zeros(n + 2 * p) for k in ("x", "xi_x", "xixx", "xix2")}
    lib().or_field_metrics(C.byref(g), axis, _dp(out["x"]), _dp(out["xi_x"]),
                           _dp(out["xixx"]), _dp(out["xix2"]))
    return out


def field_shape(g, pad):
    if g.dim == 3:
        return (g.n[0] + 2 * pad, g.n[1] + 2 * pad, g.n[2] + 2 * pad)
    return (g.n[0] + 2 * pad, g.n[2] + 2 * pad)


def core(a, pad):
    if a.ndim == 3:
        return a[pad:-pad, pad:-pad, pad:-pad]
    return a[pad:-pad, pad:-pad]


class Multigrid:
    """Oracle multigrid_d3 / multigrid_d2 (poisson.h:125-205)."""

    def __init__(self, g, p):
        self.g, self.p = g, p
        self.h = lib().or_mg_create(C.byref(g), C.byref(p))
        if not self.h:
            raise ValueError("oracle: bad grid")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mg_destroy(self.h)
            self.h = None

    @property
    def levels(self):
        return lib().or_mg_levels(self.h)

    @property
    def max_count(self):
        return lib().or_mg_max_count(self.h)

    def level_shape(self, level):
        n = (C.c_int * 3)()
        lib().or_mg_level_shape(self.h, level, n)
        return tuple(n)

    def padded_shape(self, level):
        n = self.level_shape(level)
        if self.g.dim == 3:
            return (n[0] + 2, n[1] + 2, n[2] + 2)
        return (n[0] + 2, n[2] + 2)

    def solve(self, x_full, b_full, pad, history=False):
        x_full = np.ascontiguousarray(x_full, dtype=np.float64)
        b_full = np.ascontiguousarray(b_full, dtype=np.float64)
        st = MgStats()
        hist = np.zeros(max(self.p.vc_count, 1))
        lib().or_mg_solve(self.h, _dp(x_full), _dp(b_full), pad, _dp(hist), C.byref(st))
        out = dict(cycles=st.cycles, residual=st.residual, coarse_iters_max=st.coarse_iters_max,
                   coarse_capped=st.coarse_capped)
        if history:
            out["history"] = hist[: st.cycles].copy()
        return x_full, out

    # single operations on padded level arrays
    def set_level(self, level, lhs=None, rhs=None):
        la = np.ascontiguousarray(lhs, dtype=np.float64) if lhs is not None else None
        ra = np.ascontiguousarray(rhs, dtype=np.float64) if rhs is not None else None
        if la is not None:
            assert la.shape == self.padded_shape(level)
        if ra is not None:
            assert ra.shape == self.padded_shape(level)
        lib().or_mg_set_level(self.h, level, _dp(la) if la is not None else None,
                              _dp(ra) if ra is not None else None)

    def get_level(self, level):
        shp = self.padded_shape(level)
        lhs, rhs = np.zeros(shp), np.zeros(shp)
        lib().or_mg_get_level(self.h, level, _dp(lhs), _dp(rhs))
        return lhs, rhs

    def smooth(self, level, count, zero_bc):
        lib().or_mg_smooth(self.h, level, count, int(zero_bc))

    def coarse_solve(self, level, zero_bc=1):
        return lib().or_mg_coarse_solve(self.h, level, int(zero_bc))

    def residual_restrict(self, level):
        lib().or_mg_residual_restrict(self.h, level)

    def prolong_correct(self, level):
        lib().or_mg_prolong_correct(self.h, level)

    def error(self, res_type):
        return lib().or_mg_error(self.h, res_type)

    def residual(self, level, zero_bc):
        out = np.zeros(self.padded_shape(level))
        lib().or_mg_residual(self.h, level, int(zero_bc), _dp(out))
        return out

    def apply(self, x_pad):
        """A x at level 0 (homogeneous walls), as a padded array."""
        self.set_level(0, x_pad, np.zeros_like(x_pad))
        return -self.residual(0, 1)

    def vcycles(self, count):
        lib().or_mg_vcycles(self.h, count)

    def impose_bc(self, level, zero_bc):
        lib().or_mg_impose_bc(self.h, level, int(zero_bc))


def set_threads(n):
    lib().or_set_threads(int(n))


def nse_desc(g, dscheme=2, nlscheme=3, has_T=1, nu=1e-2, kappa=1e-2, dt=1e-3, cn_tol=1e-6,
             buoy=(0.0, 0.0, -1.0), vbc=None, tbc=None, pbc=None, mg=None, scheme=0, pe_limit=2.0,
             central_bias=0.8, les=0, mean_pgrad=0.0, plate_radius=0.25):
    """Descriptor of one run.  vbc: [(types[6], vals[6]) per component];
    tbc / pbc: (types[6], vals[6]).  Defaults: the Rayleigh-Benard wiring of
    hydro.cc:287-485 / scalar.cc:129-201 (no-slip, T = 1 bottom, 0 top,
    Neumann P, Neumann T on side walls)."""
    d = NseDesc()
    d.g = g
    d.mg = mg if mg is not None else mg_params()
    d.dscheme, d.nlscheme, d.has_T = dscheme, nlscheme, int(has_T)
    d.nu, d.kappa, d.dt, d.cn_tol = nu, kappa, dt, cn_tol
    d.scheme = scheme
    d.pe_limit, d.central_bias = pe_limit, central_bias
    d.les = int(les)
    d.mean_pgrad = mean_pgrad
    d.plate_radius = plate_radius
    for a in range(3):
        d.buoy[a] = buoy[a]
    if vbc is None:
        vbc = [([FBC_DIRICHLET] * 6, [0.0] * 6)] * 3
    if tbc is None:
        tbc = ([FBC_NEUMANN] * 4 + [FBC_DIRICHLET] * 2, [0.0] * 4 + [1.0, 0.0])
    if pbc is None:
        pbc = ([FBC_NEUMANN] * 6, [0.0] * 6)
    for c in range(3):
        for w in range(6):
            d.vbc_type[c][w], d.vbc_val[c][w] = vbc[c][0][w], vbc[c][1][w]
    for w in range(6):
        d.tbc_type[w], d.tbc_val[w] = tbc[0][w], tbc[1][w]
        d.pbc_type[w], d.pbc_val[w] = pbc[0][w], pbc[1][w]
    return d


class Nse:
    """Oracle lsRK3_d3 time stepper with its fields (V, P, T, Pp)."""

    NAMES = ("vx", "vy", "vz", "p", "t", "pp")

    def __init__(self, d):
        self.d = d
        self.h = lib().or_nse_create(C.byref(d))
        g = d.g
        # planar (dim 2) arrays: one y row, no y pad
        self.shape = (g.n[0] + 2 * g.pad, 1 if g.dim == 2 else g.n[1] + 2 * g.pad, g.n[2] + 2 * g.pad)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_nse_destroy(self.h)
            self.h = None

    @property
    def max_iterations(self):
        return lib().or_nse_max_iterations(self.h)

    def set(self, **f):
        arrs = [np.ascontiguousarray(f[k], dtype=np.float64) if f.get(k) is not None else None
                for k in self.NAMES]
        for a in arrs:
            assert a is None or a.shape == self.shape
        lib().or_nse_set(self.h, *[_dp(a) if a is not None else None for a in arrs])

    def get(self):
        out = [np.zeros(self.shape) for _ in self.NAMES]
        lib().or_nse_get(self.h, *[_dp(a) for a in out])
        return dict(zip(self.NAMES, out))

    def rhs(self, stage):
        lib().or_nse_rhs(self.h, stage)
        out = [np.zeros(self.shape) for _ in range(8)]
        lib().or_nse_get_rhs(self.h, *[_dp(a) for a in out])
        return dict(zip(("rx", "ry", "rz", "rt", "hx", "hy", "hz", "ht"), out))

    def stage(self, stage):
        return lib().or_nse_stage(self.h, stage)

    def step(self):
        return lib().or_nse_step(self.h)

    def jacobi_iters(self):
        out = (C.c_int * 4)()
        lib().or_nse_jacobi_iters(self.h, out)
        return list(out)

    def mg_stats(self):
        st = MgStats()
        lib().or_nse_mg_stats(self.h, C.byref(st))
        return dict(cycles=st.cycles, residual=st.residual, coarse_iters_max=st.coarse_iters_max,
                    coarse_capped=st.coarse_capped)

    def op(self, op, F):
        F = np.ascontiguousarray(F, dtype=np.float64)
        out = np.zeros(self.shape)
        lib().or_nse_op(self.h, op, _dp(F), _dp(out))
        return out

    def vec_op(self, op, fx, fy=None, fz=None, out=None):
        """Vector operators (divergence, gradient, computeDiff, computeNLin):
        `out` (three arrays) are the initial outputs, copied."""
        f = [np.ascontiguousarray(a if a is not None else fx, dtype=np.float64) for a in (fx, fy, fz)]
        o = [np.array(a, dtype=np.float64, copy=True) for a in out] if out is not None else [np.zeros(self.shape) for _ in range(3)]
        lib().or_nse_vec_op(self.h, op, *[_dp(a) for a in f], *[_dp(a) for a in o])
        return o

    def cfl(self, courant, dt_max):
        return lib().or_nse_cfl(self.h, courant, dt_max)

    def sgs(self, H=None):
        """spiral::computeSG on H (zeros by default): returns (status, H, diag[3])."""
        H = [np.zeros(self.shape) for _ in range(3)] if H is None else [np.array(h, dtype=np.float64) for h in H]
        r = lib().or_nse_sgs(self.h, *[_dp(h) for h in H])
        dg = np.zeros(3)
        lib().or_nse_sgs_diag(self.h, _dp(dg))
        return r, H, dg

    def sgs_diag(self):
        """Diagnostics of the last sub-grid evaluation (no new evaluation)."""
        dg = np.zeros(3)
        lib().or_nse_sgs_diag(self.h, _dp(dg))
        return dg

    def stresses(self):
        T = np.zeros((6,) + self.shape)
        lib().or_nse_get_sgs(self.h, _dp(T))
        return T

    @property
    def time(self):
        return lib().or_nse_time(self.h)

    def diag(self):
        v = [C.c_double() for _ in range(4)]
        lib().or_nse_diag(self.h, *[C.byref(x) for x in v])
        return dict(ke=v[0].value, divmax=v[1].value, nuvol=v[2].value, nuwall=v[3].value)

    def thermal_energy(self):
        """tseries.cc:410-446: volume mean of 1/2 (T + z - 1)^2."""
        return lib().or_nse_thermal_energy(self.h)


def nse_physics(prob_type, re=1.0, ra=1.0, pr=1.0, rbc_type=2, g=(0.0, 0.0, -1.0)):
    """nu, kappa and buoyancy vector by the deck (timestep.cc:70-98,
    buoyantForce.cc:48-60); None where the reference aborts."""
    nu, kappa = C.c_double(), C.c_double()
    buoy = np.zeros(3)
    gv = np.asarray(g, dtype=np.float64)
    r = lib().or_nse_physics(prob_type, re, ra, pr, rbc_type, _dp(gv), C.byref(nu), C.byref(kappa), _dp(buoy))
    if r != 0:
        return None
    return dict(nu=nu.value, kappa=kappa.value, buoy=tuple(buoy))
