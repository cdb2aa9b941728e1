"""Python binding of libsaras_hip.so (include/saras_hip.h) for tests and the
benchmark.  The product is the C-ABI library; this module only marshals
arguments.  It raises if the library is missing: there is no CPU fallback.

The class names mirror the reference's drop-in: ``Multigrid`` stands for
multigrid_d3 / multigrid_d2 (lib/poisson/poisson.h:125-205) and ``solve``
for poisson::mgSolve (lib/poisson/poisson.cc:109-189).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsaras_hip.so")

OK, NOT_CONVERGED, JACOBI_DIVERGED, BAD_SHAPE, OOM, HIP_ERROR, BAD_ARG, SGS_FAILED = range(8)
GRID_UNIFORM, GRID_TANH_SINGLE, GRID_TANH_DOUBLE = 0, 1, 2
BC_NEUMANN, BC_DIRICHLET, BC_DIRICHLET_TEST = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
PTR_HOST, PTR_DEVICE = 0, 1


class SarasError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"saras status {status}: {msg}")
        self.status = status


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("len", C.c_double * 3),
                ("periodic", C.c_int * 3), ("gtype", C.c_int * 3),
                ("beta", C.c_double * 3), ("pad", C.c_int)]


class MgParams(C.Structure):
    _fields_ = [("vc_count", C.c_int), ("vc_tol", C.c_double), ("solve_flag", C.c_int),
                ("mg_tol", C.c_double), ("sor", C.c_double), ("pre_smooth", C.c_int),
                ("post_smooth", C.c_int), ("res_type", C.c_int), ("bc", C.c_int),
                ("precision", C.c_int), ("inner_cycles", C.c_int), ("max_refine", C.c_int),
                ("top_dirichlet", C.c_int)]


class MgStats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("refinements", C.c_int), ("residual", C.c_double),
                ("converged", C.c_int), ("coarse_iters_max", C.c_int),
                ("coarse_capped", C.c_int), ("device_ms", C.c_double)]


FBC_DIRICHLET, FBC_NEUMANN, FBC_HOT_PLATE, FBC_PLUME_OUT, FBC_NULL = 0, 1, 2, 3, 4


class NseParams(C.Structure):
    _fields_ = [("dscheme", C.c_int), ("nlscheme", C.c_int), ("has_T", C.c_int),
                ("nu", C.c_double), ("kappa", C.c_double), ("dt", C.c_double),
                ("cn_tol", C.c_double), ("buoy", C.c_double * 3),
                ("vbc_type", (C.c_int * 6) * 3), ("vbc_val", (C.c_double * 6) * 3),
                ("tbc_type", C.c_int * 6), ("tbc_val", C.c_double * 6),
                ("pbc_type", C.c_int * 6), ("pbc_val", C.c_double * 6), ("scheme", C.c_int),
                ("pe_limit", C.c_double), ("central_bias", C.c_double), ("les", C.c_int),
                ("mean_pgrad", C.c_double), ("plate_radius", C.c_double), ("les_sf32", C.c_int)]


class NseStats(C.Structure):
    _fields_ = [("jacobi_iters", C.c_int * 4), ("mg", MgStats), ("device_ms", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("mem_budget", C.c_size_t), ("slots", C.c_int), ("use_graphs", C.c_int),
                ("keep_resident", C.c_int), ("largest_first", C.c_int)]


class Request(C.Structure):
    _fields_ = [("grid", GridDesc), ("mg", MgParams), ("b", C.c_void_p), ("x", C.c_void_p),
                ("tag", C.c_longlong)]


class NseRequest(C.Structure):
    _fields_ = [("grid", GridDesc), ("mg", MgParams), ("nse", NseParams), ("steps", C.c_int),
                ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p), ("p", C.c_void_p), ("t", C.c_void_p),
                ("tag", C.c_longlong)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_longlong), ("status", C.c_int), ("stats", MgStats), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("message", C.c_char * 120)]


# every symbol include/saras_hip.h declares
EXPORTS = [
    "saras_mg_create", "saras_mg_solve", "saras_mg_destroy", "saras_mg_levels",
    "saras_mg_level_shape", "saras_vc_depth", "saras_level_metrics", "saras_mg_set_level",
    "saras_mg_get_level", "saras_mg_smooth", "saras_mg_coarse_solve",
    "saras_mg_residual_restrict", "saras_mg_prolong_correct", "saras_mg_error",
    "saras_device_alloc", "saras_device_free", "saras_memcpy_h2d", "saras_memcpy_d2h",
    "saras_device_sync", "saras_device_zero", "saras_last_error", "saras_mg_bench_cycles", "saras_mg_bench_sweeps",
    "saras_bench_stream_copy", "saras_nse_create", "saras_nse_physics", "saras_nse_wiring", "saras_nse_destroy", "saras_nse_set_fields",
    "saras_nse_get_fields", "saras_nse_step", "saras_nse_stage", "saras_nse_rhs", "saras_nse_get_rhs",
    "saras_nse_solve_field", "saras_field_op", "saras_field_vec_op", "saras_nse_diag", "saras_nse_bench_steps", "saras_nse_bench_sgs",
    "saras_nse_cfl", "saras_nse_sgs_diag", "saras_nse_tsdata", "saras_nse_probe", "saras_nse_probe_noalloc", "saras_mg_set_graphs", "saras_mg_footprint", "saras_nse_footprint", "saras_nse_set_time", "saras_nse_set_graphs",
    "saras_server_create", "saras_server_submit", "saras_server_submit_nse",
    "saras_server_poll", "saras_server_drain", "saras_server_clock", "saras_server_mem_used",
    "saras_server_destroy",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.saras_last_error.restype = C.c_char_p
        L.saras_vc_depth.argtypes = [C.POINTER(GridDesc)]
        L.saras_level_metrics.argtypes = [C.POINTER(GridDesc), C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.saras_mg_create.argtypes = [C.POINTER(GridDesc), C.POINTER(MgParams), C.POINTER(vp)]
        L.saras_mg_solve.argtypes = [vp, vp, vp, C.c_int, vp, C.POINTER(MgStats)]
        L.saras_mg_destroy.argtypes = [vp]
        L.saras_mg_levels.argtypes = [vp]
        L.saras_mg_level_shape.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        L.saras_mg_set_level.argtypes = [vp, C.c_int, dp, dp]
        L.saras_mg_get_level.argtypes = [vp, C.c_int, dp, dp]
        L.saras_mg_smooth.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.saras_mg_coarse_solve.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.saras_mg_residual_restrict.argtypes = [vp, C.c_int]
        L.saras_mg_prolong_correct.argtypes = [vp, C.c_int]
        L.saras_mg_error.argtypes = [vp, C.c_int, dp]
        L.saras_mg_bench_cycles.argtypes = [vp, C.c_int, dp]
        L.saras_mg_bench_sweeps.argtypes = [vp, C.c_int, dp]
        L.saras_bench_stream_copy.argtypes = [C.c_size_t, C.c_int, dp]
        L.saras_device_alloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.saras_device_free.argtypes = [vp]
        L.saras_memcpy_h2d.argtypes = [vp, vp, C.c_size_t]
        L.saras_memcpy_d2h.argtypes = [vp, vp, C.c_size_t]
        L.saras_device_zero.argtypes = [vp, C.c_size_t]
        L.saras_nse_footprint.argtypes = [C.POINTER(GridDesc), C.POINTER(MgParams), C.POINTER(NseParams)]
        L.saras_nse_footprint.restype = C.c_size_t
        L.saras_nse_set_time.argtypes = [vp, C.c_double]
        L.saras_nse_set_graphs.argtypes = [vp, C.c_int]
        L.saras_server_submit_nse.argtypes = [vp, C.POINTER(NseRequest)]
        L.saras_nse_physics.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp]
        L.saras_nse_wiring.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(NseParams)]
        L.saras_nse_create.argtypes = [C.POINTER(GridDesc), C.POINTER(MgParams), C.POINTER(NseParams),
                                       C.POINTER(vp)]
        L.saras_nse_destroy.argtypes = [vp]
        L.saras_nse_set_fields.argtypes = [vp] + [vp] * 6 + [C.c_int]
        L.saras_nse_get_fields.argtypes = [vp] + [vp] * 6 + [C.c_int]
        L.saras_nse_step.argtypes = [vp, C.POINTER(NseStats)]
        L.saras_nse_stage.argtypes = [vp, C.c_int, C.POINTER(NseStats)]
        L.saras_nse_rhs.argtypes = [vp, C.c_int]
        L.saras_nse_get_rhs.argtypes = [vp] + [vp] * 8 + [C.c_int]
        L.saras_nse_solve_field.argtypes = [vp, C.c_int, C.c_double, C.POINTER(C.c_int)]
        L.saras_field_op.argtypes = [vp, C.c_int, vp, vp, C.c_int]
        L.saras_field_vec_op.argtypes = [vp, C.c_int] + [vp] * 6 + [C.c_int]
        L.saras_nse_diag.argtypes = [vp, dp]
        L.saras_nse_tsdata.argtypes = [vp, dp]
        L.saras_nse_probe.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp]
        L.saras_nse_probe_noalloc.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp]
        L.saras_nse_bench_steps.argtypes = [vp, C.c_int, dp]
        L.saras_nse_bench_sgs.argtypes = [vp, C.c_int, dp]
        L.saras_nse_cfl.argtypes = [vp, C.c_double, C.c_double, dp]
        L.saras_nse_sgs_diag.argtypes = [vp, dp, dp]
        L.saras_mg_set_graphs.argtypes = [vp, C.c_int]
        L.saras_mg_footprint.argtypes = [C.POINTER(GridDesc), C.POINTER(MgParams)]
        L.saras_mg_footprint.restype = C.c_size_t
        L.saras_server_create.argtypes = [C.POINTER(ServerParams), C.POINTER(vp)]
        L.saras_server_submit.argtypes = [vp, C.POINTER(Request)]
        L.saras_server_poll.argtypes = [vp, C.POINTER(Completion), C.c_int, C.c_int]
        L.saras_server_drain.argtypes = [vp]
        L.saras_server_clock.argtypes = [vp]
        L.saras_server_clock.restype = C.c_double
        L.saras_server_mem_used.argtypes = [vp]
        L.saras_server_mem_used.restype = C.c_size_t
        L.saras_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def check(status, allow=(OK,)):
    if status not in allow:
        raise SarasError(status, lib().saras_last_error().decode())
    return status


def device_sync():
    check(lib().saras_device_sync())


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def nse_physics(prob_type, re=1.0, ra=1.0, pr=1.0, rbc_type=2, g=(0.0, 0.0, -1.0)):
    """nu, kappa and the buoyancy vector gTerms from the deck (timestep.cc:70-98,
    buoyantForce.cc:48-60), for nse_params(nu=, kappa=, buoy=)."""
    nu, kappa = C.c_double(), C.c_double()
    buoy = np.zeros(3)
    gv = np.ascontiguousarray(g, dtype=np.float64)
    check(lib().saras_nse_physics(prob_type, re, ra, pr, rbc_type, _dp(gv), C.byref(nu), C.byref(kappa), _dp(buoy)))
    return dict(nu=nu.value, kappa=kappa.value, buoy=tuple(buoy))


def nse_wiring(prob_type, planar=False, non_hg=False):
    """The drivers' wall kinds and values for a "Problem Type" (hydro.cc:349-484,
    scalar.cc:163-201), as the vbc / tbc / pbc arguments of nse_params."""
    d = NseParams()
    check(lib().saras_nse_wiring(prob_type, int(planar), int(non_hg), C.byref(d)))
    vbc = [([d.vbc_type[c][w] for w in range(6)], [d.vbc_val[c][w] for w in range(6)]) for c in range(3)]
    tbc = ([d.tbc_type[w] for w in range(6)], [d.tbc_val[w] for w in range(6)])
    pbc = ([d.pbc_type[w] for w in range(6)], [d.pbc_val[w] for w in range(6)])
    return dict(vbc=vbc, tbc=tbc, pbc=pbc)


def nse_footprint(g, mg, prm):
    return lib().saras_nse_footprint(C.byref(g), C.byref(mg), C.byref(prm))


def grid_desc(n, dim=3, length=(1.0, 1.0, 1.0), periodic=(0, 0, 0), gtype=(0, 0, 0),
              beta=(1.0, 1.0, 1.0), pad=1):
    g = GridDesc()
    g.dim = dim
    n = tuple(n) if hasattr(n, "__len__") else (n, n, n)
    for a in range(3):
        g.n[a], g.len[a], g.periodic[a] = n[a], length[a], int(periodic[a])
        g.gtype[a], g.beta[a] = int(gtype[a]), beta[a]
    g.pad = pad
    return g


def mg_params(vc_count=20, vc_tol=1e-6, solve_flag=1, mg_tol=1e-2, sor=1.0, pre=3, post=6,
              res_type=0, bc=BC_NEUMANN, precision=PREC_FP64, inner_cycles=3, max_refine=20, top_dirichlet=0):
    p = MgParams()
    p.vc_count, p.vc_tol, p.solve_flag, p.mg_tol = vc_count, vc_tol, int(solve_flag), mg_tol
    p.sor, p.pre_smooth, p.post_smooth, p.res_type = sor, pre, post, res_type
    p.bc, p.precision, p.inner_cycles, p.max_refine = bc, precision, inner_cycles, max_refine
    p.top_dirichlet = int(top_dirichlet)
    return p


def vc_depth(g):
    return lib().saras_vc_depth(C.byref(g))


def level_metrics(g, level, axis):
    n0 = 1 if (g.dim == 2 and axis == 1) else g.n[axis]
    nl = n0 if (g.dim == 2 and axis == 1) else n0 >> level
    out = {k: np.zeros(nl + 2) for k in ("x", "xi", "xi_x", "xixx", "xix2")}
    check(lib().saras_level_metrics(C.byref(g), level, axis, _dp(out["x"]), _dp(out["xi"]),
                                    _dp(out["xi_x"]), _dp(out["xixx"]), _dp(out["xix2"])))
    return out


class DeviceArray:
    """A device buffer owned through the library's own allocator."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        check(lib().saras_device_alloc(C.byref(self.ptr), nbytes))
        self.nbytes = nbytes

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        check(lib().saras_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_host(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        check(lib().saras_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        check(lib().saras_device_zero(self.ptr, self.nbytes))

    def free(self):
        if self.ptr:
            lib().saras_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Multigrid:
    """multigrid_d3 / multigrid_d2 backed by the device hierarchy."""

    def __init__(self, g, p):
        self.g, self.p = g, p
        self.h = C.c_void_p()
        check(lib().saras_mg_create(C.byref(g), C.byref(p), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().saras_mg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def levels(self):
        return lib().saras_mg_levels(self.h)

    def level_shape(self, level):
        n = (C.c_int * 3)()
        check(lib().saras_mg_level_shape(self.h, level, n))
        return tuple(n)

    def padded_shape(self, level):
        n = self.level_shape(level)
        return (n[0] + 2, n[1] + 2, n[2] + 2) if self.g.dim == 3 else (n[0] + 2, n[2] + 2)

    def solve(self, x_full, b_full, allow_not_converged=True):
        """poisson::mgSolve on host arrays; returns (x, stats dict, status)."""
        x = np.array(x_full, dtype=np.float64, order="C", copy=True)
        b = np.ascontiguousarray(b_full, dtype=np.float64)
        st = MgStats()
        allow = (OK, NOT_CONVERGED) if allow_not_converged else (OK,)
        s = check(lib().saras_mg_solve(self.h, x.ctypes.data, b.ctypes.data, PTR_HOST, None,
                                       C.byref(st)), allow)
        return x, stats_dict(st), s

    def solve_device(self, x_dev, b_dev, stream=None):
        st = MgStats()
        s = check(lib().saras_mg_solve(self.h, x_dev.ptr, b_dev.ptr, PTR_DEVICE, stream, C.byref(st)),
                  (OK, NOT_CONVERGED))
        return stats_dict(st), s

    def set_level(self, level, lhs=None, rhs=None):
        la = np.ascontiguousarray(lhs, dtype=np.float64) if lhs is not None else None
        ra = np.ascontiguousarray(rhs, dtype=np.float64) if rhs is not None else None
        for a in (la, ra):
            if a is not None:
                assert a.shape == self.padded_shape(level)
        check(lib().saras_mg_set_level(self.h, level, _dp(la) if la is not None else None,
                                       _dp(ra) if ra is not None else None))

    def get_level(self, level):
        shp = self.padded_shape(level)
        lhs, rhs = np.zeros(shp), np.zeros(shp)
        check(lib().saras_mg_get_level(self.h, level, _dp(lhs), _dp(rhs)))
        return lhs, rhs

    def smooth(self, level, count, zero_bc):
        check(lib().saras_mg_smooth(self.h, level, count, int(zero_bc)))

    def coarse_solve(self, level, zero_bc=1):
        it = C.c_int()
        check(lib().saras_mg_coarse_solve(self.h, level, int(zero_bc), C.byref(it)))
        return it.value

    def residual_restrict(self, level):
        check(lib().saras_mg_residual_restrict(self.h, level))

    def prolong_correct(self, level):
        check(lib().saras_mg_prolong_correct(self.h, level))

    def bench_cycles(self, count):
        ms = C.c_double()
        check(lib().saras_mg_bench_cycles(self.h, count, C.byref(ms)))
        return ms.value

    def bench_sweeps(self, reps):
        ms = C.c_double()
        check(lib().saras_mg_bench_sweeps(self.h, reps, C.byref(ms)))
        return ms.value

    def error(self, res_type):
        v = C.c_double()
        check(lib().saras_mg_error(self.h, res_type, C.byref(v)))
        return v.value

    def set_graphs(self, enable):
        check(lib().saras_mg_set_graphs(self.h, int(enable)))


def stats_dict(st):
    return dict(cycles=st.cycles, refinements=st.refinements, residual=st.residual,
                converged=bool(st.converged), coarse_iters_max=st.coarse_iters_max,
                coarse_capped=st.coarse_capped, device_ms=st.device_ms)


def bench_stream_copy(n_doubles, reps):
    """ms for `reps` copies of n doubles (reads + writes 16 n bytes each)."""
    ms = C.c_double()
    check(lib().saras_bench_stream_copy(n_doubles, reps, C.byref(ms)))
    return ms.value


def nse_params(dscheme=2, nlscheme=3, has_T=1, nu=1e-2, kappa=1e-2, dt=1e-3, cn_tol=1e-6,
               buoy=(0.0, 0.0, -1.0), vbc=None, tbc=None, pbc=None, scheme=0, pe_limit=2.0, central_bias=0.8,
               les=0, mean_pgrad=0.0, plate_radius=0.25, les_sf32=0):
    """Time-step parameters.  vbc: [(types[6], vals[6]) per component];
    tbc / pbc: (types[6], vals[6]).  Defaults: the Rayleigh-Benard wiring of
    hydro.cc:287-485 / scalar.cc:129-201 (no-slip walls, T = 1 at the bottom
    and 0 at the top, Neumann T on side walls, Neumann P).  mean_pgrad: the
    constant x pressure gradient of force type 5 (constantPGrad, force.h:148-153;
    hydrodynamics runs, has_T = 0).  les_sf32: the sub-grid model's
    structure-function integral in single precision."""
    d = NseParams()
    d.dscheme, d.nlscheme, d.has_T = dscheme, nlscheme, int(has_T)
    d.nu, d.kappa, d.dt, d.cn_tol = nu, kappa, dt, cn_tol
    d.scheme = scheme
    d.pe_limit, d.central_bias = pe_limit, central_bias
    d.les = int(les)
    d.les_sf32 = int(les_sf32)
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


def nse_stats_dict(st):
    return dict(jacobi_iters=list(st.jacobi_iters), mg=stats_dict(st.mg), device_ms=st.device_ms)


class Nse:
    """lsRK3_d3 time stepper with its fields (V, P, T, Pp) on the device
    (lsRK3_d3.cc:58-70, 190-305)."""

    NAMES = ("vx", "vy", "vz", "p", "t", "pp")

    def __init__(self, g, mg, p):
        self.g, self.mgp, self.p = g, mg, p
        self.h = C.c_void_p()
        check(lib().saras_nse_create(C.byref(g), C.byref(mg), C.byref(p), C.byref(self.h)))
        # planar (dim 2) arrays have no y extent beyond the single row
        ypad = 0 if g.dim == 2 else g.pad
        self.shape = (g.n[0] + 2 * g.pad, g.n[1] + 2 * ypad, g.n[2] + 2 * g.pad)

    def close(self):
        if self.h:
            lib().saras_nse_destroy(self.h)
            self.h = C.c_void_p()

    def set_time(self, t):
        check(lib().saras_nse_set_time(self.h, float(t)))

    def set_graphs(self, enable):
        check(lib().saras_nse_set_graphs(self.h, int(enable)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, **f):
        arrs = [np.ascontiguousarray(f[k], dtype=np.float64) if f.get(k) is not None else None
                for k in self.NAMES]
        for a in arrs:
            assert a is None or a.shape == self.shape
        check(lib().saras_nse_set_fields(self.h, *[a.ctypes.data if a is not None else None for a in arrs],
                                         PTR_HOST))

    def get(self, base=None):
        out = [np.zeros(self.shape) if base is None else np.array(base, dtype=np.float64) for _ in self.NAMES]
        check(lib().saras_nse_get_fields(self.h, *[a.ctypes.data for a in out], PTR_HOST))
        return dict(zip(self.NAMES, out))

    def rhs(self, stage):
        check(lib().saras_nse_rhs(self.h, stage))
        out = [np.zeros(self.shape) for _ in range(8)]
        check(lib().saras_nse_get_rhs(self.h, *[a.ctypes.data for a in out], PTR_HOST))
        return dict(zip(("rx", "ry", "rz", "rt", "hx", "hy", "hz", "ht"), out))

    def stage(self, stage):
        st = NseStats()
        s = lib().saras_nse_stage(self.h, stage, C.byref(st))
        return s, nse_stats_dict(st)

    def step(self):
        st = NseStats()
        s = lib().saras_nse_step(self.h, C.byref(st))
        return s, nse_stats_dict(st)

    def solve_field(self, field, beta):
        it = C.c_int()
        s = lib().saras_nse_solve_field(self.h, field, beta, C.byref(it))
        return s, it.value

    def op(self, op, F, out=None):
        F = np.ascontiguousarray(F, dtype=np.float64)
        out = np.zeros(self.shape) if out is None else np.array(out, dtype=np.float64)
        check(lib().saras_field_op(self.h, op, F.ctypes.data, out.ctypes.data, PTR_HOST))
        return out

    def vec_op(self, op, fx, fy=None, fz=None, out=None):
        """saras_field_vec_op: 0 divergence, 1 gradient of fx, 2 computeDiff,
        3 computeNLin; `out` (three arrays) are the initial outputs, copied."""
        f = [np.ascontiguousarray(a if a is not None else fx, dtype=np.float64) for a in (fx, fy, fz)]
        o = [np.array(a, dtype=np.float64, copy=True) for a in out] if out is not None else \
            [np.zeros(self.shape) for _ in range(3)]
        check(lib().saras_field_vec_op(self.h, op, *[a.ctypes.data for a in f], *[a.ctypes.data for a in o], PTR_HOST))
        return o

    def cfl(self, courant, dt_max):
        dt = C.c_double()
        check(lib().saras_nse_cfl(self.h, courant, dt_max, C.byref(dt)))
        return dt.value

    def diag(self):
        v = np.zeros(4)
        check(lib().saras_nse_diag(self.h, _dp(v)))
        return dict(ke=v[0], divmax=v[1], nuvol=v[2], nuwall=v[3])

    def tsdata(self):
        """The quantities of a TimeSeries.dat line (tseries.cc:209-490)."""
        v = np.zeros(5)
        check(lib().saras_nse_tsdata(self.h, _dp(v)))
        return dict(ke=v[0], te=v[1], divmax=v[2], nuvol=v[3], nuwall=v[4])

    def probe(self, points):
        """Vx, Vy, Vz, P, T at core cells `points` ((n, 3) integer indices)."""
        ijk = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 3))
        out = np.zeros((ijk.shape[0], 5))
        check(lib().saras_nse_probe(self.h, ijk.shape[0], ijk.ctypes.data_as(C.POINTER(C.c_int)), _dp(out)))
        return out

    def probe_noalloc(self, points):
        """As probe(), staged in the library's fixed device array (no device
        allocation; at most PROBE_NOALLOC_MAX points)."""
        ijk = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 3))
        out = np.zeros((ijk.shape[0], 5))
        check(lib().saras_nse_probe_noalloc(self.h, ijk.shape[0], ijk.ctypes.data_as(C.POINTER(C.c_int)), _dp(out)))
        return out

    def sgs(self):
        """Sub-grid diagnostics of the last LES evaluation and the run time."""
        v, t = np.zeros(3), C.c_double()
        check(lib().saras_nse_sgs_diag(self.h, _dp(v), C.byref(t)))
        return dict(sgke=v[0], disp=v[1], nusg=v[2]), t.value

    def bench_steps(self, count):
        ms = C.c_double()
        s = lib().saras_nse_bench_steps(self.h, count, C.byref(ms))
        return s, ms.value

    def bench_sgs(self, count):
        ms = C.c_double()
        check(lib().saras_nse_bench_sgs(self.h, count, C.byref(ms)))
        return ms.value


def footprint(g, p):
    return lib().saras_mg_footprint(C.byref(g), C.byref(p))


class Server:
    """The persistent batched request loop (saras_server_*).  Arrays handed
    to submit() are kept referenced until their completion is polled."""

    def __init__(self, mem_budget, slots, use_graphs=True, keep_resident=True, largest_first=False):
        sp = ServerParams()
        sp.mem_budget, sp.slots = int(mem_budget), int(slots)
        sp.use_graphs, sp.keep_resident = int(use_graphs), int(keep_resident)
        sp.largest_first = int(largest_first)
        self.h = C.c_void_p()
        check(lib().saras_server_create(C.byref(sp), C.byref(self.h)))
        self.live = {}

    def submit(self, g, p, b, x, tag):
        """b, x: C-contiguous float64 arrays in the caller layout; x is in/out."""
        assert b.dtype == np.float64 and x.dtype == np.float64 and b.flags.c_contiguous and x.flags.c_contiguous
        r = Request()
        r.grid, r.mg, r.b, r.x, r.tag = g, p, b.ctypes.data, x.ctypes.data, int(tag)
        self.live[int(tag)] = (b, x)
        check(lib().saras_server_submit(self.h, C.byref(r)))

    def submit_nse(self, g, mg, prm, steps, fields, tag):
        """A short Navier-Stokes run (saras_server_submit_nse): `fields` maps
        vx, vy, vz, p (and t with a scalar) to C-contiguous float64 arrays in
        the caller layout, updated in place to the state after `steps` steps."""
        names = ("vx", "vy", "vz", "p") + (("t",) if prm.has_T else ())
        for k in names:
            a = fields[k]
            assert a.dtype == np.float64 and a.flags.c_contiguous, k
        r = NseRequest()
        r.grid, r.mg, r.nse, r.steps, r.tag = g, mg, prm, int(steps), int(tag)
        for k in names:
            setattr(r, k, fields[k].ctypes.data)
        self.live[int(tag)] = tuple(fields[k] for k in names)
        check(lib().saras_server_submit_nse(self.h, C.byref(r)))

    def poll(self, max_n=256, wait_ms=0):
        buf = (Completion * max_n)()
        n = lib().saras_server_poll(self.h, buf, max_n, int(wait_ms))
        out = []
        for q in range(n):
            c = buf[q]
            self.live.pop(int(c.tag), None)
            out.append(dict(tag=c.tag, status=c.status, stats=stats_dict(c.stats), t_submit=c.t_submit,
                            t_admit=c.t_admit, t_done=c.t_done, message=c.message.decode(errors="replace")))
        return out

    def drain(self):
        check(lib().saras_server_drain(self.h))

    def clock(self):
        return lib().saras_server_clock(self.h)

    def mem_used(self):
        return lib().saras_server_mem_used(self.h)

    def close(self):
        if self.h:
            lib().saras_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
