"""Pins of the oracle's grid restatement (lib/grid/grid.cc) against tables,
closed forms and finite differences of the mapping."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n,dim,per,depth", [
    ((64, 64, 64), 3, (0, 0, 0), 4),      # 2^7 > 64 -> 6; coarsest 1 -> 5; all walls -> 4
    ((256, 256, 256), 3, (1, 1, 0), 7),   # -> 8; coarsest 1 -> 7
    ((512, 512, 512), 3, (1, 1, 1), 8),
    ((384, 384, 384), 3, (1, 1, 1), 7),   # 384 = 3*128; coarsest 3
    ((96, 96, 96), 3, (0, 0, 0), 4),      # mgTest deck: 96 = 3*32 -> 5; all walls -> 4
    ((2048, 1, 2048), 2, (1, 0, 1), 10),
    ((64, 1, 128), 2, (0, 0, 0), 4),      # min axis decides: 2^7 !| 64 -> 6; 64/64=1 -> 5; walls -> 4
    ((48, 48, 48), 3, (1, 1, 1), 4),      # 48 = 3*16 -> coarsest 3
    ((32, 32, 32), 3, (0, 0, 0), 3),
])
def test_vc_depth_table(n, dim, per, depth):
    g = O.grid_desc(n, dim=dim, periodic=per)
    assert O.vc_depth(g) == depth


def test_uniform_metrics_all_levels():
    g = O.grid_desc((32, 16, 8), length=(2.0, 1.0, 0.5), periodic=(1, 1, 1))
    d = O.vc_depth(g)
    assert d == 2
    for lev in range(d + 1):
        for ax in range(3):
            m = O.level_metrics(g, lev, ax)
            n = g.n[ax] >> lev
            L = g.len[ax]
            i = np.arange(-1, n + 1)
            xi = (i + 0.5) / n
            # cell centres, ghosts: level 0 continues the lattice; coarse ghost x(-1) = -x(0), x(n) = L + x(0)
            np.testing.assert_allclose(m["xi"][1:-1], xi[1:-1], rtol=0, atol=1e-15)
            np.testing.assert_allclose(m["x"][1:-1], L * xi[1:-1], rtol=1e-15)
            np.testing.assert_allclose(m["x"][0], -L * xi[1], rtol=1e-15)
            np.testing.assert_allclose(m["x"][-1], L + L * xi[1], rtol=1e-15)
            assert np.all(m["xi_x"] == 1.0 / L)
            assert np.all(m["xixx"] == 0.0)
            assert np.all(m["xix2"] == (1.0 / L) ** 2)


def _inverse_double(x, L, beta):
    # xi(x) for x = L (1 - tanh(beta (1 - 2 xi)) / tanh beta) / 2
    return 0.5 * (1.0 - np.arctanh((1.0 - 2.0 * x / L) * np.tanh(beta)) / beta)


def _inverse_single(x, L, beta):
    return 1.0 - np.arctanh((1.0 - x / L) * np.tanh(beta)) / beta


@pytest.mark.parametrize("gtype,inv", [(O.GRID_TANH_DOUBLE, _inverse_double),
                                        (O.GRID_TANH_SINGLE, _inverse_single)])
@pytest.mark.parametrize("beta,L", [(1.5, 1.0), (1.0, 2.0)])
def test_tanh_metrics_match_inverse_mapping(gtype, inv, beta, L):
    g = O.grid_desc((16, 16, 64), length=(1.0, 1.0, L), periodic=(1, 1, 0),
                    gtype=(0, 0, gtype), beta=(1, 1, beta), pad=2)
    d = O.vc_depth(g)
    for lev in range(d + 1):
        m = O.level_metrics(g, lev, 2)
        n = 64 >> lev
        x = m["x"][1:-1]
        # the mapping evaluated at the level's xi reproduces the coordinate
        np.testing.assert_allclose(inv(x, L, beta), m["xi"][1:-1], rtol=0, atol=1e-12)
        # xi_x and xi_xx by central differences of the inverse mapping
        eps = 1e-5 * L
        d1 = (inv(x + eps, L, beta) - inv(x - eps, L, beta)) / (2 * eps)
        d2 = (inv(x + eps, L, beta) - 2 * inv(x, L, beta) + inv(x - eps, L, beta)) / eps ** 2
        np.testing.assert_allclose(m["xi_x"][1:-1], d1, rtol=1e-6)
        np.testing.assert_allclose(m["xixx"][1:-1], d2, rtol=2e-4, atol=1e-3)
        np.testing.assert_allclose(m["xix2"][1:-1], m["xi_x"][1:-1] ** 2, rtol=1e-15)
        # ghost rules: level 0 wraps periodically, coarse levels mirror about 0
        if lev == 0:
            assert m["x"][0] == pytest.approx(m["x"][n] - L, abs=1e-15)
        else:
            assert m["x"][0] == -m["x"][1]
        assert m["x"][-1] == pytest.approx(L + m["x"][1], abs=1e-15)
        assert m["xi_x"][0] == m["xi_x"][n] and m["xi_x"][-1] == m["xi_x"][1]


def test_coarse_xi_is_mean_of_children():
    g = O.grid_desc((32, 32, 32), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5))
    for lev in range(1, O.vc_depth(g) + 1):
        f = O.level_metrics(g, lev - 1, 2)["xi"][1:-1]
        c = O.level_metrics(g, lev, 2)["xi"][1:-1]
        np.testing.assert_array_equal(c, (f[0::2] + f[1::2]) / 2.0)


def test_field_metrics_pad2_wraps():
    g = O.grid_desc((16, 16, 32), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 1.5), pad=2)
    fm = O.field_metrics(g, 2)
    lm = O.level_metrics(g, 0, 2)
    # the level-0 multigrid arrays are the [-1, n] slice of the pad-2 field arrays
    np.testing.assert_array_equal(fm["x"][1:-1], lm["x"])
    np.testing.assert_array_equal(fm["xix2"][1:-1], lm["xix2"])
    assert fm["x"][0] == pytest.approx(fm["x"][32] - 1.0, abs=1e-15)
    assert fm["x"][-1] == pytest.approx(1.0 + fm["x"][3], abs=1e-15)
