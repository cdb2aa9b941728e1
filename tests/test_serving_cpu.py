"""Host-side pieces of the serving loop: the request mix, its right-hand
sides, the memory footprint model and argument checks (no device calls)."""
import numpy as np

import saras_amd as S
from saras_amd import serving as V


def test_request_mix_matches_the_workload():
    reqs = V.draw_requests(4000, seed=4)
    d3 = [r for r in reqs if r["dim"] == 3]
    d2 = [r for r in reqs if r["dim"] == 2]
    assert 0.72 < len(d3) / len(reqs) < 0.78
    assert all(r["n"] % 16 == 0 and 32 <= r["n"] <= 256 for r in d3)
    assert all(r["n"] % 16 == 0 and 64 <= r["n"] <= 2048 for r in d2)
    # log-uniform: about half the 3-D draws below the geometric midpoint 90.5
    assert 0.4 < np.mean([r["n"] <= 90 for r in d3]) < 0.6
    kinds = [r["kind"] for r in reqs]
    for k in V.KINDS:
        assert 0.3 < kinds.count(k) / len(kinds) < 0.37
    assert V.draw_requests(50, seed=4) == V.draw_requests(50, seed=4)


def test_mix_extents_coarsen_to_small_levels():
    """Every drawn extent is a multigrid-friendly one: a multiple of 16 whose
    odd part of n/16 is at most 7, so the reference's depth rule
    (grid.cc:678-742) leaves a coarsest level of at most 7 cells per axis
    (14 with walls on every face); the extents are the nearest such values."""
    assert V.mg_extents(32, 256) == [32, 48, 64, 80, 96, 112, 128, 160, 192, 224, 256]
    assert V.nearest_extent(208, 32, 256) == 192 and V.nearest_extent(209, 32, 256) == 224
    assert V.nearest_extent(2000, 64, 2048) == 2048 and V.nearest_extent(33, 32, 256) == 32
    for r in V.draw_requests(600, seed=4):
        g = V.request_grid(r)
        depth = S.vc_depth(g)
        assert depth >= 2, r
        coarse = r["n"] >> depth
        assert coarse << depth == r["n"]
        assert coarse <= (7 if r["kind"] == "periodic" else 14), (r, coarse)


def test_request_rhs_compatibility():
    for kind in V.KINDS:
        for dim in (2, 3):
            req = dict(tag=0, dim=dim, n=32, kind=kind, rhs_seed=7)
            b = V.request_rhs(req)
            core = b[(slice(1, -1),) * dim]
            assert core.shape == (32,) * dim
            assert np.abs(b).max() > 0.01
            if kind != "dirichlet":
                assert abs(core.mean()) < 1e-15
            # ghosts are zero (the solver reads the core only)
            assert b[0].max() == 0 and b[-1].max() == 0


def test_footprint_model_grows_with_the_grid_and_precision():
    f = lambda n, prec: S.footprint(S.grid_desc(n, periodic=(1, 1, 1), pad=1), S.mg_params(precision=prec))
    assert 0 < f(32, S.PREC_FP64) < f(64, S.PREC_FP64) < f(128, S.PREC_FP64)
    assert f(128, S.PREC_MIXED) > f(128, S.PREC_FP64)
    # 256^3 FP64: level arrays x, b, xalt (8/7 of the finest) plus host staging, under a gigabyte
    assert 0.5e9 < f(256, S.PREC_FP64) < 1.0e9
    assert S.footprint(S.grid_desc((3, 3, 3), pad=1), S.mg_params()) >= 0


def test_server_rejects_bad_arguments():
    L = S.lib()
    sp = S.ServerParams()
    sp.mem_budget, sp.slots = 0, 4
    h = S.C.c_void_p()
    assert L.saras_server_create(S.C.byref(sp), S.C.byref(h)) == S.BAD_ARG
    sp.mem_budget, sp.slots = 1 << 30, 0
    assert L.saras_server_create(S.C.byref(sp), S.C.byref(h)) == S.BAD_ARG
    assert L.saras_server_submit(None, None) == S.BAD_ARG
    assert L.saras_server_poll(None, None, 0, 0) == -1
