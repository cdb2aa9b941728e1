"""GPU parity of the FP64 multigrid path against the FP64 CPU oracle.

Bar: bit-exact.  Both sides evaluate the reference's expressions in the same
order with separately rounded operations (FMA contraction off), so every
smoother sweep, restriction, prolongation, coarse solve, L-inf norm and whole
solve must agree to the last bit.  Sum-based norms (L1, L2) reduce in a
different order and are held to 1e-13 relative.
"""
import numpy as np
import pytest

import oracle as O
import saras_amd as S

pytestmark = pytest.mark.gpu

BCMAP = {O.BC_NEUMANN: S.BC_NEUMANN, O.BC_DIRICHLET: S.BC_DIRICHLET,
         O.BC_DIRICHLET_TEST: S.BC_DIRICHLET_TEST}

CASES = [
    # n, dim, periodic, gtype, bc
    ((8, 8, 8), 3, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET),
    ((16, 16, 16), 3, (0, 0, 0), (2, 2, 2), O.BC_DIRICHLET),
    ((16, 16, 32), 3, (1, 1, 0), (0, 0, 2), O.BC_NEUMANN),
    ((16, 8, 16), 3, (1, 1, 1), (0, 0, 0), O.BC_NEUMANN),
    ((12, 12, 12), 3, (1, 0, 1), (0, 1, 0), O.BC_NEUMANN),
    ((16, 16, 16), 3, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET_TEST),
    ((24, 24, 24), 3, (1, 1, 1), (0, 0, 0), O.BC_NEUMANN),
    ((3, 3, 3), 3, (1, 1, 1), (0, 0, 0), O.BC_NEUMANN),
    ((32, 1, 16), 2, (0, 0, 0), (2, 0, 0), O.BC_DIRICHLET),
    ((32, 1, 32), 2, (1, 0, 0), (0, 0, 2), O.BC_NEUMANN),
    ((32, 1, 32), 2, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET_TEST),
    ((5, 1, 7), 2, (1, 0, 1), (0, 0, 0), O.BC_NEUMANN),
]
IDS = [f"{c[1]}d-{'x'.join(map(str, c[0]))}-p{''.join(map(str, c[2]))}-g{''.join(map(str, c[3]))}-bc{c[4]}"
       for c in CASES]


def pair(n, dim, per, gt, bc, beta=(0.7, 0.6, 0.8), **kw):
    go = O.grid_desc(n, dim=dim, periodic=per, gtype=gt, beta=beta)
    gs = S.grid_desc(n, dim=dim, periodic=per, gtype=gt, beta=beta)
    po = O.mg_params(bc=bc, **kw)
    ps = S.mg_params(bc=BCMAP[bc], **kw)
    return O.Multigrid(go, po), S.Multigrid(gs, ps)


def inner(a):
    return a[(slice(1, -1),) * a.ndim]


def rand_pad(shape, rng):
    a = np.zeros(shape)
    a[(slice(1, -1),) * len(shape)] = rng.standard_normal(tuple(s - 2 for s in shape))
    return a


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("zero_bc", [0, 1])
@pytest.mark.parametrize("sor", [1.15, 1.0])
def test_smooth_bitexact(case, zero_bc, sor):
    mo, ms = pair(*case, sor=sor)
    rng = np.random.default_rng(11)
    shp = mo.padded_shape(0)
    x, b = rand_pad(shp, rng), rand_pad(shp, rng)
    mo.set_level(0, x, b)
    ms.set_level(0, x, b)
    mo.smooth(0, 3, zero_bc)
    ms.smooth(0, 3, zero_bc)
    np.testing.assert_array_equal(inner(ms.get_level(0)[0]), inner(mo.get_level(0)[0]))


@pytest.mark.parametrize("case", [c for c in CASES if c[0] != (3, 3, 3) and c[0] != (5, 1, 7)],
                         ids=[i for c, i in zip(CASES, IDS) if c[0] != (3, 3, 3) and c[0] != (5, 1, 7)])
def test_restrict_prolong_bitexact(case):
    mo, ms = pair(*case)
    rng = np.random.default_rng(12)
    levels = mo.levels
    assert ms.levels == levels
    for lev in range(levels - 1):
        shp = mo.padded_shape(lev)
        x, b = rand_pad(shp, rng), rand_pad(shp, rng)
        mo.set_level(lev, x, b)
        ms.set_level(lev, x, b)
        mo.impose_bc(lev, 0 if lev == 0 else 1)
        mo.residual_restrict(lev)
        ms.residual_restrict(lev)
        lo, ro = mo.get_level(lev + 1)
        ls, rs = ms.get_level(lev + 1)
        np.testing.assert_array_equal(inner(rs), inner(ro))
        np.testing.assert_array_equal(inner(ls), inner(lo))
        e = rand_pad(mo.padded_shape(lev + 1), rng)
        mo.set_level(lev + 1, e, None)
        ms.set_level(lev + 1, e, None)
        mo.prolong_correct(lev)
        ms.prolong_correct(lev)
        np.testing.assert_array_equal(inner(ms.get_level(lev)[0]), inner(mo.get_level(lev)[0]))


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_coarse_solve_bitexact(case):
    mo, ms = pair(*case, mg_tol=1e-9)
    rng = np.random.default_rng(13)
    D = mo.levels - 1
    shp = mo.padded_shape(D)
    x, b = rand_pad(shp, rng), rand_pad(shp, rng)
    mo.set_level(D, x, b)
    ms.set_level(D, x, b)
    zb = 1 if D > 0 else 0
    io = mo.coarse_solve(D, zb)
    is_ = ms.coarse_solve(D, zb)
    assert is_ == io
    np.testing.assert_array_equal(inner(ms.get_level(D)[0]), inner(mo.get_level(D)[0]))


@pytest.mark.parametrize("case", CASES[:6] + CASES[8:11], ids=IDS[:6] + IDS[8:11])
@pytest.mark.parametrize("res_type", [0, 1, 2])
def test_error_norms(case, res_type):
    mo, ms = pair(*case)
    rng = np.random.default_rng(14)
    shp = mo.padded_shape(0)
    x, b = rand_pad(shp, rng), rand_pad(shp, rng)
    mo.set_level(0, x, b)
    ms.set_level(0, x, b)
    mo.impose_bc(0, 0)
    eo, es = mo.error(res_type), ms.error(res_type)
    if res_type == 0:
        assert es == eo
    else:
        assert es == pytest.approx(eo, rel=1e-13)


def rhs_for(mo, rng, compatible):
    g = mo.g
    shp = O.field_shape(g, 1)
    b = np.zeros(shp)
    c = rng.standard_normal(tuple(s - 2 for s in shp))
    if compatible:
        c -= c.mean()
    b[(slice(1, -1),) * len(shp)] = c
    return b


SOLVE_CASES = [c for c in CASES if c[0] not in ((3, 3, 3), (5, 1, 7))]
SOLVE_IDS = [i for c, i in zip(CASES, IDS) if c[0] not in ((3, 3, 3), (5, 1, 7))]


@pytest.mark.parametrize("tail", [True, False], ids=["tail-kernel", "per-level"])
@pytest.mark.parametrize("case", SOLVE_CASES, ids=SOLVE_IDS)
def test_full_solve_bitexact(case, tail, monkeypatch):
    """Whole solves agree bit for bit, with the small levels run by the
    single-workgroup tail kernel and with every level run level by level."""
    if not tail:
        monkeypatch.setenv("SARAS_NO_TAIL", "1")
    kw = dict(vc_count=12, vc_tol=1e-9, mg_tol=1e-6, pre=2, post=2)
    mo, ms = pair(*case, **kw)
    rng = np.random.default_rng(15)
    b = rhs_for(mo, rng, case[4] == O.BC_NEUMANN)
    x0 = rhs_for(mo, rng, False) * 0.1
    xo, so = mo.solve(x0.copy(), b, 1, history=True)
    xs, ss, status = ms.solve(x0.copy(), b)
    assert ss["cycles"] == so["cycles"]
    assert ss["residual"] == so["residual"]
    assert ss["coarse_iters_max"] == so["coarse_iters_max"]
    np.testing.assert_array_equal(xs, xo)   # core and the whole ghost shell


def test_config1_dirichlet_sine_64():
    """64^3 unit cube, homogeneous Dirichlet, b = -3 pi^2 sin sin sin + U(-1e-3,
    1e-3) (seed 1), V(2,2) RB-GS, coarsest 4^3 solved to 1e-6, L-inf relative
    residual 1e-8: identical V-cycle count, residual and solution bits."""
    n = 64
    mo, ms = pair((n, n, n), 3, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET,
                  vc_count=100, vc_tol=1e-8, mg_tol=1e-6, pre=2, post=2)
    assert mo.levels == 5 and mo.level_shape(4) == (4, 4, 4)
    xc = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(xc, xc, xc, indexing="ij")
    rng = np.random.default_rng(1)
    core = -3 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    core += rng.uniform(-1e-3, 1e-3, core.shape)
    b = np.pad(core, 1)
    xo, so = mo.solve(np.zeros_like(b), b, 1)
    xs, ss, status = ms.solve(np.zeros_like(b), b)
    assert status == S.OK and ss["converged"]
    assert ss["cycles"] == so["cycles"] and ss["residual"] == so["residual"]
    np.testing.assert_array_equal(xs, xo)
    u = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    assert np.abs(xs[1:-1, 1:-1, 1:-1] - u).max() < 2e-3


def test_device_pointer_path_and_pad2():
    n = 32
    go = O.grid_desc((n, n, n), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 0.9), pad=2)
    gs = S.grid_desc((n, n, n), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 0.9), pad=2)
    kw = dict(vc_count=8, vc_tol=1e-12, mg_tol=1e-3, pre=3, post=6)
    mo = O.Multigrid(go, O.mg_params(bc=O.BC_NEUMANN, **kw))
    ms = S.Multigrid(gs, S.mg_params(bc=S.BC_NEUMANN, **kw))
    rng = np.random.default_rng(16)
    shp = (n + 4,) * 3
    b = rng.standard_normal(shp)
    x0 = rng.standard_normal(shp)   # outer pad layer must survive untouched
    xo, so = mo.solve(x0.copy(), b, 2)
    xd, bd = S.DeviceArray.from_host(x0), S.DeviceArray.from_host(b)
    ss, status = ms.solve_device(xd, bd)
    xs = xd.to_host(shp)
    assert status == S.NOT_CONVERGED and ss["cycles"] == 8
    np.testing.assert_array_equal(xs, xo)
    np.testing.assert_array_equal(xs[0], x0[0])


@pytest.mark.parametrize("fused", [True, False], ids=["fused-sweep", "two-pass"])
@pytest.mark.parametrize("case", [
    ((32, 16, 64), 3, (1, 1, 0), (0, 0, 2), O.BC_NEUMANN),
    ((32, 32, 32), 3, (0, 0, 0), (2, 1, 2), O.BC_DIRICHLET),
    ((32, 32, 32), 3, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET_TEST),
    ((16, 16, 32), 3, (1, 1, 1), (0, 0, 0), O.BC_NEUMANN),
], ids=["ppn-z2", "nnn-stretched-dir", "nnn-dirtest", "ppp"])
def test_smoother_paths_bitexact(case, fused, monkeypatch):
    """The fused out-of-place sweep and the two in-place colour passes both
    reproduce the reference smoother bit for bit, for several sweeps."""
    if not fused:
        monkeypatch.setenv("SARAS_NO_FUSED_SWEEP", "1")
    mo, ms = pair(*case, sor=1.2)
    rng = np.random.default_rng(21)
    shp = mo.padded_shape(0)
    x, b = rand_pad(shp, rng), rand_pad(shp, rng)
    for lev in range(mo.levels - 1):
        shp = mo.padded_shape(lev)
        x, b = rand_pad(shp, rng), rand_pad(shp, rng)
        mo.set_level(lev, x, b)
        ms.set_level(lev, x, b)
        zb = 0 if lev == 0 else 1
        mo.smooth(lev, 3, zb)
        ms.smooth(lev, 3, zb)
        np.testing.assert_array_equal(inner(ms.get_level(lev)[0]), inner(mo.get_level(lev)[0]),
                                      err_msg=f"level {lev}")


def oracle_residual(g_kw, bc, x_full, b_full, res_type=0):
    """Residual of a returned solution evaluated independently on the CPU."""
    go = O.grid_desc(**g_kw)
    mo = O.Multigrid(go, O.mg_params(bc=bc))
    mo.set_level(0, x_full, b_full)
    mo.impose_bc(0, 0)
    return mo.error(res_type)


MIXED_CASES = [
    (dict(n=(64, 64, 64), periodic=(1, 1, 1)), O.BC_NEUMANN, "periodic"),
    (dict(n=(64, 64, 64), periodic=(0, 0, 0)), O.BC_DIRICHLET, "dirichlet"),
    (dict(n=(32, 32, 32), periodic=(0, 0, 0)), O.BC_DIRICHLET_TEST, "dirichlet-r2"),
    # stretched singular problems: the reference's coarsest-level solve makes
    # the V-cycle diverge (recorded in test_oracle_mg), so smooth it instead
    (dict(n=(64, 32, 32), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 0.3)), O.BC_NEUMANN, "ppn-stretched"),
    (dict(n=(128, 1, 64), dim=2, periodic=(0, 0, 0)), O.BC_DIRICHLET, "2d-dirichlet"),
]


@pytest.mark.parametrize("g_kw,bc,name", MIXED_CASES, ids=[c[2] for c in MIXED_CASES])
def test_mixed_precision_reaches_fp64_tolerance(g_kw, bc, name):
    """FP32 V-cycles inside FP64 refinement reach a relative L-inf residual of
    1e-10 (checked on the CPU by the oracle's computeError), and the solution
    matches the FP64 oracle's own converged solve."""
    tol = 1e-10
    kw = dict(vc_count=200, vc_tol=tol, mg_tol=1e-6, pre=2, post=2, solve_flag=0 if "gtype" in g_kw else 1)
    gs = S.grid_desc(**g_kw)
    ms = S.Multigrid(gs, S.mg_params(bc=BCMAP[bc], precision=S.PREC_MIXED, inner_cycles=3, max_refine=30, **kw))
    go = O.grid_desc(**g_kw)
    mo = O.Multigrid(go, O.mg_params(bc=bc, **kw))
    rng = np.random.default_rng(31)
    if "gtype" in g_kw and bc == O.BC_NEUMANN:
        # stretched singular operator: its left null vector is not constant, so
        # a zero-mean RHS is not compatible; take b in the range, b = A x_true
        b = mo.apply(rhs_for(mo, rng, False))
    else:
        b = rhs_for(mo, rng, bc == O.BC_NEUMANN)
    xs, ss, status = ms.solve(np.zeros_like(b), b)
    assert status == S.OK and ss["converged"], ss
    assert ss["refinements"] >= 1
    r_cpu = oracle_residual(g_kw, bc, xs, b)
    assert r_cpu < 1.5 * tol, (r_cpu, ss)
    assert abs(r_cpu - ss["residual"]) <= 1e-3 * ss["residual"] + 1e-16
    xo, so = mo.solve(np.zeros_like(b), b, 1)
    c_s, c_o = inner(xs), inner(xo)
    if bc == O.BC_NEUMANN:
        c_s, c_o = c_s - c_s.mean(), c_o - c_o.mean()
    rel = np.abs(c_s - c_o).max() / np.abs(c_o).max()
    assert rel < 1e-7, rel


def test_mixed_inner_cycles_follow_convergence():
    """inner_cycles caps the FP32 cycles of a refinement; the first runs 3 to
    measure the convergence factor, later ones the cycles it says reach the
    tolerance (or the FP32 accuracy floor).  The FP64 oracle needs 9 cycles
    here; a fixed 8 per refinement would run 16."""
    tol = 1e-10
    g_kw = dict(n=(64, 64, 64), periodic=(1, 1, 1))
    kw = dict(vc_count=200, vc_tol=tol, mg_tol=1e-6, pre=2, post=2, solve_flag=1)
    ms = S.Multigrid(S.grid_desc(**g_kw),
                     S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_MIXED, inner_cycles=8, max_refine=30, **kw))
    mo = O.Multigrid(O.grid_desc(**g_kw), O.mg_params(bc=O.BC_NEUMANN, **kw))
    b = rhs_for(mo, np.random.default_rng(31), True)
    xs, ss, status = ms.solve(np.zeros_like(b), b)
    assert status == S.OK and ss["converged"], ss
    xo, so = mo.solve(np.zeros_like(b), b, 1)
    assert so["cycles"] > 8
    assert ss["cycles"] <= so["cycles"] + 2, (ss, so)
    assert oracle_residual(g_kw, O.BC_NEUMANN, xs, b) < 1.5 * tol


@pytest.mark.parametrize("case", [((56, 56, 56), 3, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET),
                                  ((48, 48, 48), 3, (0, 0, 0), (2, 2, 2), O.BC_DIRICHLET),
                                  ((224, 1, 224), 2, (0, 0, 0), (0, 0, 0), O.BC_DIRICHLET)],
                         ids=["3d-56-coarse14", "3d-48-stretched-coarse12", "2d-224-coarse14"])
def test_full_solve_large_coarsest_level(case):
    """Grids whose coarsest level alone exceeds the single-workgroup tail's
    LDS budget (few factors of two): the coarsest solve runs in the
    multi-level path with global scratch, bit-exact with the oracle."""
    kw = dict(vc_count=4, vc_tol=1e-12, mg_tol=1e-6, pre=2, post=2)
    mo, ms = pair(*case, **kw)
    rng = np.random.default_rng(16)
    b = rhs_for(mo, rng, False)
    xo, so = mo.solve(np.zeros_like(b), b, 1, history=True)
    xs, ss, status = ms.solve(np.zeros_like(b), b)
    assert ss["cycles"] == so["cycles"]
    assert ss["residual"] == so["residual"]
    assert ss["coarse_iters_max"] == so["coarse_iters_max"]
    np.testing.assert_array_equal(xs, xo)


@pytest.mark.parametrize("case", [((128, 256, 256), 3, (1, 1, 0), (0, 0, 2), O.BC_NEUMANN),
                                  ((256, 128, 256), 3, (0, 0, 0), (2, 0, 1), O.BC_DIRICHLET)],
                         ids=["ppn-z-stretched", "nnn-dirichlet"])
def test_fused_march_sweeps_bitexact_at_size(case):
    """Levels large enough for the fused out-of-place march sweep (>= 8M
    cells): three sweeps of the finest level agree bit for bit with the
    reference smoother."""
    mo, ms = pair(*case, sor=1.1, beta=(1.0, 1.2, 1.4))
    rng = np.random.default_rng(31)
    shp = mo.padded_shape(0)
    x, b = rand_pad(shp, rng), rand_pad(shp, rng)
    mo.set_level(0, x, b)
    ms.set_level(0, x, b)
    mo.smooth(0, 3, 0)
    ms.smooth(0, 3, 0)
    np.testing.assert_array_equal(inner(ms.get_level(0)[0]), inner(mo.get_level(0)[0]))


SPLIT_CASES = [
    dict(n=(128, 256, 256), dim=3, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.2), bc=S.BC_NEUMANN),
    dict(n=(96, 80, 72), dim=3, periodic=(0, 0, 0), gtype=(2, 0, 1), beta=(1.2, 1, 0.9), bc=S.BC_DIRICHLET),
    dict(n=(64, 64, 64), dim=3, periodic=(1, 1, 1), gtype=(0, 0, 0), beta=(1, 1, 1), bc=S.BC_NEUMANN),
    dict(n=(256, 1, 512), dim=2, periodic=(1, 0, 0), gtype=(0, 0, 2), beta=(1, 1, 1.3), bc=S.BC_NEUMANN),
]


@pytest.mark.parametrize("variant", ["z-march", "z-two-pass", "z-tj8"])
@pytest.mark.parametrize("case", SPLIT_CASES, ids=["ppn-256", "nnn-mixed-layouts", "ppp-64", "2d-pn"])
def test_fp32_split_rows_match_z_contiguous(case, variant, monkeypatch):
    """FP32 cycles of the mixed solver on colour-split rows (the default for
    every level whose half-rows are whole float4 vectors) against the
    z-contiguous layout with the fused march sweep, its 64 x 8 variant or the
    two in-place colour passes: the same FP32 arithmetic per cell, so whole
    mixed solves are bit-identical (cases include hierarchies whose coarse
    levels fall back to z-contiguous rows)."""
    if case["dim"] == 2 and variant == "z-tj8":
        pytest.skip("the march sweep is 3-D only")
    g = S.grid_desc(case["n"], dim=case["dim"], periodic=case["periodic"], gtype=case["gtype"], beta=case["beta"],
                    pad=1)
    p = S.mg_params(vc_count=6, vc_tol=1e-14, solve_flag=0, mg_tol=1e-2, pre=2, post=2, bc=case["bc"],
                    precision=S.PREC_MIXED, inner_cycles=3, max_refine=2)
    rng = np.random.default_rng(32)
    n = case["n"] if case["dim"] == 3 else (case["n"][0], case["n"][2])
    b = np.zeros(tuple(m + 2 for m in n))
    b[(slice(1, -1),) * len(n)] = rng.standard_normal(n)
    split = S.Multigrid(g, p)
    xs, ss, _ = split.solve(np.zeros_like(b), b)
    split.close()
    monkeypatch.setenv("SARAS_NO_SPLIT", "1")
    if variant == "z-two-pass":
        monkeypatch.setenv("SARAS_NO_FUSED_SWEEP", "1")
    elif variant == "z-tj8":
        monkeypatch.setenv("SARAS_MARCH_TJ8", "1")
    other = S.Multigrid(g, p)
    xo, so, _ = other.solve(np.zeros_like(b), b)
    other.close()
    assert ss["cycles"] == so["cycles"]
    assert ss["residual"] == so["residual"]
    np.testing.assert_array_equal(xs, xo)



@pytest.mark.parametrize("switch", [({}, {"SARAS_RR_LANE": "1"}), ({}, {"SARAS_NO_RED_FOLD": "1"}),
                                    ({}, {"SARAS_TAIL_LDS_KB": "8"}),
                                    ({"SARAS_NO_SPLIT": "1"}, {"SARAS_MARCH_WGS": "128"})],
                         ids=["rr-lane", "no-red-fold", "tail-8kb", "z-contiguous-march-128"])
@pytest.mark.parametrize("prec", [S.PREC_FP64, S.PREC_MIXED], ids=["fp64", "mixed"])
def test_kernel_variant_switches_bitexact(switch, prec, monkeypatch):
    """The handle-creation switches that select another kernel arrangement of
    the same arithmetic (restriction with all eight children in one lane; the
    coarse level's first red pass left to the sweep instead of the
    restriction; a smaller tail LDS budget, so fewer levels run in the tail;
    another march grid) give whole solves bit-identical to the default."""
    g = S.grid_desc((64, 64, 64), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 0.8), pad=1)
    p = S.mg_params(vc_count=6, vc_tol=1e-14, solve_flag=0, mg_tol=1e-2, pre=2, post=2, bc=S.BC_NEUMANN,
                    precision=prec, inner_cycles=3, max_refine=2)
    rng = np.random.default_rng(33)
    b = np.zeros((66, 66, 66))
    c = rng.standard_normal((64, 64, 64))
    b[1:-1, 1:-1, 1:-1] = c - c.mean()
    base_env, other_env = switch
    for k, v in base_env.items():
        monkeypatch.setenv(k, v)
    base = S.Multigrid(g, p)
    xs, ss, _ = base.solve(np.zeros_like(b), b)
    base.close()
    for k, v in other_env.items():
        monkeypatch.setenv(k, v)
    other = S.Multigrid(g, p)
    xo, so, _ = other.solve(np.zeros_like(b), b)
    other.close()
    assert ss["cycles"] == so["cycles"] and ss["residual"] == so["residual"], (ss, so)
    np.testing.assert_array_equal(xs, xo)


@pytest.mark.parametrize("tail", [True, False], ids=["tail-kernel", "per-level"])
@pytest.mark.parametrize("gt", [(0, 0, 0), (0, 0, 2)], ids=["uniform", "tanh-z"])
def test_2d_top_dirichlet_solve_bitexact(gt, tail, monkeypatch):
    """multigrid_d2's probType 9 walls (poisson_d2.cc:641-647): Neumann with a
    homogeneous Dirichlet top; whole solves agree bit for bit, ghosts included."""
    if not tail:
        monkeypatch.setenv("SARAS_NO_TAIL", "1")
    kw = dict(vc_count=12, vc_tol=1e-9, mg_tol=1e-6, pre=2, post=2, top_dirichlet=1)
    mo, ms = pair((64, 1, 32), 2, (1, 0, 0), gt, O.BC_NEUMANN, **kw)
    rng = np.random.default_rng(31)
    b = rhs_for(mo, rng, False)
    x0 = rhs_for(mo, rng, False) * 0.1
    xo, so = mo.solve(x0.copy(), b, 1, history=True)
    xs, ss, status = ms.solve(x0.copy(), b)
    assert ss["cycles"] == so["cycles"] and ss["residual"] == so["residual"]
    np.testing.assert_array_equal(xs, xo)
    assert np.array_equal(xs[1:-1, -1], -xs[1:-1, -2])


@pytest.mark.parametrize("gt", [(0, 0, 0), (0, 0, 2)], ids=["uniform", "tanh-z"])
def test_2d_top_dirichlet_mixed_precision(gt):
    """The probType 9 walls of multigrid_d2 (Neumann, homogeneous Dirichlet top;
    poisson_d2.cc:641-647) on the mixed-precision path: FP32 split-row cycles
    inside FP64 refinement reach 1e-10, confirmed by the oracle's computeError
    with the same walls, and the solution is the FP64 oracle's to 1e-8.  The
    stretched grid smooths its coarsest level: solving it makes the reference
    algorithm diverge there."""
    tol = 1e-10
    kw = dict(vc_count=200, vc_tol=tol, mg_tol=1e-6, pre=2, post=2, top_dirichlet=1, beta=(0.7, 0.6, 0.8),
              solve_flag=0 if gt[2] else 1)
    mo, _ = pair((64, 1, 32), 2, (1, 0, 0), gt, O.BC_NEUMANN, **{k: v for k, v in kw.items() if k != "beta"})
    gs = S.grid_desc((64, 1, 32), dim=2, periodic=(1, 0, 0), gtype=gt, beta=kw["beta"])
    ms = S.Multigrid(gs, S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_MIXED, inner_cycles=3, max_refine=40,
                                     **{k: v for k, v in kw.items() if k != "beta"}))
    b = rhs_for(mo, np.random.default_rng(32), False)
    xs, ss, status = ms.solve(np.zeros_like(b), b)
    assert status == S.OK and ss["converged"] and ss["refinements"] >= 1, ss
    mo.set_level(0, xs, b)
    mo.impose_bc(0, 0)
    r_cpu = mo.error(0)
    assert r_cpu < 1.5 * tol and abs(r_cpu - ss["residual"]) <= 1e-3 * ss["residual"], (r_cpu, ss)
    assert np.array_equal(xs[1:-1, -1], -xs[1:-1, -2])          # the Dirichlet top ghost
    mo2, _ = pair((64, 1, 32), 2, (1, 0, 0), gt, O.BC_NEUMANN, **{k: v for k, v in kw.items() if k != "beta"})
    xo, so = mo2.solve(np.zeros_like(b), b, 1)
    assert so["residual"] < tol
    assert np.abs(inner(xs) - inner(xo)).max() <= 1e-8 * np.abs(inner(xo)).max()


@pytest.mark.parametrize("case", [
    ((32, 16, 64), 3, (1, 1, 0), (2, 1, 2), O.BC_NEUMANN),
    ((16, 32, 32), 3, (0, 0, 0), (2, 2, 2), O.BC_DIRICHLET),
    ((64, 1, 64), 2, (0, 0, 0), (2, 0, 2), O.BC_NEUMANN),
], ids=["ppn-stretched", "nnn-dir-stretched", "2d-nn-stretched"])
def test_unit_sor_signed_zeros_bitwise(case):
    """With SOR parameter 1 the colour passes skip the cells' own values except
    where the quotient is a zero, whose sign the reference's (1 - w) x + q
    takes from x.  Iterates and right-hand sides of mixed +0, -0 and a few
    non-zero values on stretched grids (first-derivative coefficients of both
    signs): the result is identical to the oracle bit for bit, zero signs
    included."""
    mo, ms = pair(*case, sor=1.0)
    rng = np.random.default_rng(41)
    neg = 0
    for lev in range(min(2, mo.levels - 1)):
        shp = mo.padded_shape(lev)
        core = tuple(t - 2 for t in shp)
        x, b = np.zeros(shp), np.zeros(shp)
        x[(slice(1, -1),) * len(shp)] = np.where(rng.random(core) < 0.01, rng.standard_normal(core),
                                                  rng.choice([0.0, -0.0], core))
        b[(slice(1, -1),) * len(shp)] = np.where(rng.random(core) < 0.005, rng.standard_normal(core),
                                                  rng.choice([0.0, -0.0], core))
        for sweeps in (1, 1):
            mo.set_level(lev, x, b)
            ms.set_level(lev, x, b)
            zb = 0 if lev == 0 else 1
            mo.smooth(lev, sweeps, zb)
            ms.smooth(lev, sweeps, zb)
            a, e = inner(ms.get_level(lev)[0]), inner(mo.get_level(lev)[0])
            np.testing.assert_array_equal(a.view(np.int64), e.view(np.int64), err_msg=f"level {lev}")
            neg += int((np.signbit(e) & (e == 0)).sum())
            x = mo.get_level(lev)[0]
    assert neg > 0   # negative zeros occur, so their sign rule was exercised



def test_large_level_tiled_rows_bitexact():
    """On levels whose two i-planes of iterate and right-hand side exceed half
    of an XCD's L2 (here FP64 rows of 512 cells, 512 rows per plane) the colour
    passes visit rows in j-tiles; every cell is still updated once per pass,
    bit for bit as the reference's loops do, for both SOR parameters."""
    for sor in (1.0, 1.3):
        mo, ms = pair((16, 512, 512), 3, (1, 0, 1), (0, 2, 0), O.BC_NEUMANN, sor=sor)
        rng = np.random.default_rng(43)
        shp = mo.padded_shape(0)
        x, b = rand_pad(shp, rng), rand_pad(shp, rng)
        mo.set_level(0, x, b)
        ms.set_level(0, x, b)
        mo.smooth(0, 2, 0)
        ms.smooth(0, 2, 0)
        np.testing.assert_array_equal(inner(ms.get_level(0)[0]), inner(mo.get_level(0)[0]), err_msg=f"sor {sor}")
