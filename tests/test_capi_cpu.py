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
