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

OK, NOT_CONVERGED, JACOBI_DIVERGED, BAD_SHAPE, OOM, HIP_ERROR, BAD_ARG = range(7)
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
                ("precision", C.c_int), ("inner_cycles", C.c_int), ("max_refine", C.c_int)]


class MgStats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("refinements", C.c_int), ("residual", C.c_double),
                ("converged", C.c_int), ("coarse_iters_max", C.c_int),
                ("coarse_capped", C.c_int), ("device_ms", C.c_double)]


# every symbol include/saras_hip.h declares
EXPORTS = [
    "saras_mg_create", "saras_mg_solve", "saras_mg_destroy", "saras_mg_levels",
    "saras_mg_level_shape", "saras_vc_depth", "saras_level_metrics", "saras_mg_set_level",
    "saras_mg_get_level", "saras_mg_smooth", "saras_mg_coarse_solve",
    "saras_mg_residual_restrict", "saras_mg_prolong_correct", "saras_mg_error",
    "saras_device_alloc", "saras_device_free", "saras_memcpy_h2d", "saras_memcpy_d2h",
    "saras_device_sync", "saras_last_error", "saras_mg_bench_cycles", "saras_mg_bench_sweeps",
    "saras_bench_stream_copy",
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
        _lib = L
    return _lib


def check(status, allow=(OK,)):
    if status not in allow:
        raise SarasError(status, lib().saras_last_error().decode())
    return status


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


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
              res_type=0, bc=BC_NEUMANN, precision=PREC_FP64, inner_cycles=3, max_refine=20):
    p = MgParams()
    p.vc_count, p.vc_tol, p.solve_flag, p.mg_tol = vc_count, vc_tol, int(solve_flag), mg_tol
    p.sor, p.pre_smooth, p.post_smooth, p.res_type = sor, pre, post, res_type
    p.bc, p.precision, p.inner_cycles, p.max_refine = bc, precision, inner_cycles, max_refine
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


def stats_dict(st):
    return dict(cycles=st.cycles, refinements=st.refinements, residual=st.residual,
                converged=bool(st.converged), coarse_iters_max=st.coarse_iters_max,
                coarse_capped=st.coarse_capped, device_ms=st.device_ms)


def bench_stream_copy(n_doubles, reps):
    """ms for `reps` copies of n doubles (reads + writes 16 n bytes each)."""
    ms = C.c_double()
    check(lib().saras_bench_stream_copy(n_doubles, reps, C.byref(ms)))
    return ms.value
