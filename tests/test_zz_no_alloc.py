"""The allocation-free probe entry point (saras_nse_probe_noalloc) against
saras_nse_probe: the same values, bit for bit, on two calls in a row (the
second with fewer points, so stale staging would show), on two handles of
different sizes, and a clear error past the point limit."""
import numpy as np
import pytest

import saras_amd as S

pytestmark = pytest.mark.gpu

LIMIT = 8192   # SARAS_PROBE_NOALLOC_MAX


def _run(n, seed):
    pad = 2
    g = S.grid_desc(n, periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1.0, 1.0, 1.5), pad=pad)
    mg = S.mg_params(bc=S.BC_NEUMANN, precision=S.PREC_FP64, vc_count=8, vc_tol=1e-6, solve_flag=0, mg_tol=1e-2,
                     pre=3, post=6)
    nu = (1.0 / 1e7) ** 0.5
    s = S.Nse(g, mg, S.nse_params(nu=nu, kappa=nu, dt=1e-3))
    rng = np.random.default_rng(seed)
    shape = tuple(q + 2 * pad for q in n)
    f = {k: rng.standard_normal(shape) for k in ("vx", "vy", "vz", "p", "t")}
    f["pp"] = np.zeros(shape)
    s.set(**f)
    r, _ = s.step()
    assert r == 0
    return s, rng


def _points(rng, n, count):
    return np.stack([rng.integers(0, n[a], count) for a in range(3)], axis=1).astype(np.int32)


@pytest.mark.parametrize("n", [(16, 16, 16), (32, 16, 16)])
def test_probe_noalloc_matches_probe(n):
    s, rng = _run(n, 11)
    first = _points(rng, n, 3000)
    second = _points(rng, n, 37)
    for pts in (first, second):
        ref = s.probe(pts)
        got = s.probe_noalloc(pts)
        assert got.shape == ref.shape == (len(pts), 5)
        assert np.array_equal(got, ref)
    # the largest accepted batch, then one point past the limit
    pts = _points(rng, n, LIMIT)
    assert np.array_equal(s.probe_noalloc(pts), s.probe(pts))
    with pytest.raises(S.SarasError):
        s.probe_noalloc(_points(rng, n, LIMIT + 1))
    with pytest.raises(S.SarasError):
        s.probe_noalloc([(n[0], 0, 0)])
    # after the failed calls the entry point still agrees
    pts = _points(rng, n, 5)
    assert np.array_equal(s.probe_noalloc(pts), s.probe(pts))
