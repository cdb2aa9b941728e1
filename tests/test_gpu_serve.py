"""GPU tests of the serving path: hipGraph-replayed V-cycles against eager
launches, and the batched request loop (each result the same alone and in
the mix, memory-based admission).  Bar: bit-identical, since every kernel is
deterministic and a request's arithmetic does not depend on what else runs."""
import numpy as np
import pytest

import oracle as O
import saras_amd as S
from saras_amd import serving as V

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [S.PREC_FP64, S.PREC_MIXED])
@pytest.mark.parametrize("sweeps", [(2, 2), (3, 6)])
@pytest.mark.parametrize("case", [dict(n=(64, 64, 64), periodic=(1, 1, 1)),
                                  dict(n=(32, 32, 64), periodic=(1, 1, 0), gtype=(0, 0, 2), beta=(1, 1, 0.5)),
                                  dict(n=(128, 1, 96), dim=2, periodic=(0, 0, 0))],
                         ids=["ppp64", "ppn-stretched", "2d"])
def test_graph_replay_matches_eager(prec, sweeps, case):
    """Three solves in a row on a graph-replaying handle equal the same three
    solves on an eager handle, with other eager work interleaved between the
    replays (V(3,6) swaps buffers an odd number of times per cycle, so both
    cached graph states are exercised)."""
    g = S.grid_desc(case["n"], dim=case.get("dim", 3), periodic=case["periodic"], gtype=case.get("gtype", (0, 0, 0)),
                    beta=case.get("beta", (1, 1, 1)), pad=1)
    p = S.mg_params(vc_count=6, vc_tol=1e-14, solve_flag=1, mg_tol=1e-3, pre=sweeps[0], post=sweeps[1],
                    bc=S.BC_NEUMANN, precision=prec, inner_cycles=2, max_refine=3)
    eager, graph, other = S.Multigrid(g, p), S.Multigrid(g, p), S.Multigrid(g, p)
    graph.set_graphs(1)
    rng = np.random.default_rng(5)
    shape = tuple(s + 2 for s in (case["n"] if case.get("dim", 3) == 3 else (case["n"][0], case["n"][2])))
    for rep in range(3):
        b = np.zeros(shape)
        b[(slice(1, -1),) * len(shape)] = rng.standard_normal(tuple(s - 2 for s in shape))
        b[(slice(1, -1),) * len(shape)] -= b[(slice(1, -1),) * len(shape)].mean()
        xe, se, _ = eager.solve(np.zeros(shape), b)
        other.solve(np.zeros(shape), b)           # unrelated eager work between replays
        xg, sg, _ = graph.solve(np.zeros(shape), b)
        assert se["cycles"] == sg["cycles"], rep
        assert np.array_equal(xe, xg), (rep, float(np.abs(xe - xg).max()))


def small_mix(count=24, seed=9):
    reqs = V.draw_requests(count, seed=seed, n3=(32, 64), n2=(64, 256))
    return reqs


@pytest.mark.parametrize("graphs,largest_first", [(0, 0), (1, 0), (1, 1)])
def test_alone_equals_in_the_mix(graphs, largest_first):
    reqs = small_mix()
    rhs = V.RhsCache()
    mix = S.Server(mem_budget=16 << 30, slots=8, use_graphs=graphs, largest_first=largest_first)
    recs, sols = V.run_stream(mix, reqs, rhs, S.PREC_MIXED, collect={r["tag"] for r in reqs})
    mix.close()
    assert len(recs) == len(reqs)
    for r in reqs:
        rec = recs[r["tag"]]
        assert rec["status"] == S.OK and rec["stats"]["converged"], (r, rec)
        assert rec["stats"]["residual"] < 1e-8
    # each request alone, on an idle server of one slot, then on a plain handle
    for r in reqs[:12]:
        alone = S.Server(mem_budget=16 << 30, slots=1, use_graphs=graphs)
        rec1, sol1 = V.run_stream(alone, [r], rhs, S.PREC_MIXED, collect={r["tag"]})
        alone.close()
        assert np.array_equal(sol1[r["tag"]], sols[r["tag"]]), r
        assert rec1[r["tag"]]["stats"]["cycles"] == recs[r["tag"]]["stats"]["cycles"]
        h = S.Multigrid(V.request_grid(r), V.request_params(r, S.PREC_MIXED))
        x, st, _ = h.solve(np.zeros_like(rhs.get(r)), rhs.get(r))
        h.close()
        assert np.array_equal(x, sols[r["tag"]]), r


def test_memory_admission_bounds_residency():
    """A budget of two large hierarchies: everything still completes, requests
    that do not fit wait or evict idle hierarchies, and resident memory never
    exceeds the budget once the queue is empty."""
    reqs = small_mix(count=30, seed=3)
    rhs = V.RhsCache()
    big = max(S.footprint(V.request_grid(r), V.request_params(r, S.PREC_MIXED)) for r in reqs)
    budget = 2 * big
    srv = S.Server(mem_budget=budget, slots=8, use_graphs=1)
    recs, _ = V.run_stream(srv, reqs, rhs, S.PREC_MIXED)
    assert srv.mem_used() <= budget
    srv.close()
    assert len(recs) == len(reqs)
    assert all(r["status"] in (S.OK, S.NOT_CONVERGED) for r in recs.values())


def test_request_larger_than_budget_runs_alone():
    r = dict(tag=0, dim=3, n=64, kind="neumann", rhs_seed=1)
    rhs = V.RhsCache()
    srv = S.Server(mem_budget=1 << 20, slots=4, use_graphs=0)
    recs, sols = V.run_stream(srv, [r], rhs, S.PREC_FP64, collect={0})
    srv.close()
    assert recs[0]["status"] in (S.OK, S.NOT_CONVERGED)
    h = S.Multigrid(V.request_grid(r), V.request_params(r, S.PREC_FP64))
    x, _, _ = h.solve(np.zeros_like(rhs.get(r)), rhs.get(r))
    assert np.array_equal(x, sols[0])


def full_size_sample(seed=4, count=240):
    """From the serving stream itself (seed 4, 240 requests, 32^3-256^3 and
    64^2-2048^2): the smallest and the largest request of every (dimension,
    boundary kind), so both ends of the size range and every wall rule occur."""
    reqs = V.draw_requests(count, seed=seed)
    pick = {}
    for r in reqs:
        for end, better in (("lo", lambda a, b: a < b), ("hi", lambda a, b: a > b)):
            key = (r["dim"], r["kind"], end)
            if key not in pick or better(r["n"], pick[key]["n"]):
                pick[key] = r
    out = sorted({r["tag"]: r for r in pick.values()}.values(), key=lambda r: r["tag"])
    return out


def oracle_error(req, x_full, b_full):
    """computeError (poisson_d3.cc:563-663, poisson_d2.cc) of a returned
    solution, evaluated by the FP64 CPU oracle on its own level arrays."""
    g = V.request_grid(req)
    go = O.grid_desc(tuple(g.n), dim=g.dim, periodic=tuple(g.periodic), pad=1)
    bc = O.BC_DIRICHLET if req["kind"] == "dirichlet" else O.BC_NEUMANN
    mo = O.Multigrid(go, O.mg_params(bc=bc))
    mo.set_level(0, x_full, b_full)
    mo.impose_bc(0, 0)
    return mo.error(0)


def test_full_size_mix_alone_equals_in_the_mix():
    """The serving workload at its real sizes (up to 256^3 and 2048^2): every
    request converges to 1e-8 in the batched mix, its solution is bit-identical
    to a plain one-request solve, and the oracle's computeError of the
    returned solution confirms the residual the server reported."""
    reqs = full_size_sample()
    assert max(r["n"] for r in reqs if r["dim"] == 3) == 256
    assert max(r["n"] for r in reqs if r["dim"] == 2) >= 1792
    rhs = V.RhsCache()
    mix = S.Server(mem_budget=64 << 30, slots=8, use_graphs=1, largest_first=1)
    recs, sols = V.run_stream(mix, reqs, rhs, S.PREC_MIXED, collect={r["tag"] for r in reqs})
    mix.close()
    for r in reqs:
        print(f"request {r['tag']}: {r['dim']}-D n={r['n']} {r['kind']}", flush=True)
        rec = recs[r["tag"]]
        assert rec["status"] == S.OK and rec["stats"]["converged"], (r, rec)
        h = S.Multigrid(V.request_grid(r), V.request_params(r, S.PREC_MIXED))
        x, st, status = h.solve(np.zeros_like(rhs.get(r)), rhs.get(r))
        h.close()
        assert status == S.OK and st["cycles"] == rec["stats"]["cycles"], (r, st, rec)
        assert np.array_equal(x, sols[r["tag"]]), r
        e = oracle_error(r, sols[r["tag"]], rhs.get(r))
        assert e < 1e-8, (r, e)
        assert abs(e - rec["stats"]["residual"]) <= 1e-12 * e, (r, e, rec["stats"]["residual"])


def solo_nse(req):
    """A short run on its own plain handle: set, steps, get."""
    g, mg, prm, f = V.nse_request_setup(req)
    h = S.Nse(g, mg, prm)
    h.set(**f, pp=np.zeros_like(f["vx"]))
    for _ in range(req["steps"]):
        r, _ = h.step()
        assert r == S.OK
    out = h.get()
    h.close()
    return out


def test_nse_requests_alone_equal_in_the_mix():
    """Short Navier-Stokes runs (Rayleigh-Benard and Taylor-Green boxes of
    16^3-64^3, 2-5 steps) interleaved with Poisson solves in the batched
    loop: every run's fields are bit-identical to the same run alone on a
    plain handle; a repeat of the first runs reuses their resident handles
    (run time and pressure correction reset) with the same results."""
    reqs = small_mix(count=12, seed=5)
    nreqs = V.draw_nse_requests(10, seed=6)
    assert {r["kind"] for r in nreqs} == {"rbc", "tg"}
    rhs = V.RhsCache()
    srv = S.Server(mem_budget=16 << 30, slots=8, use_graphs=1, largest_first=1)
    recs, sols, outs = V.run_mixed_stream(srv, reqs, nreqs, rhs, S.PREC_MIXED)
    # the same runs again: each reuses the idle handle its first run left
    again = [dict(r, tag=r["tag"] + 500) for r in nreqs[:4]]
    recs2, _, outs2 = V.run_mixed_stream(srv, [], again, rhs, S.PREC_MIXED)
    srv.close()
    for r in reqs:
        assert recs[r["tag"]]["status"] == S.OK, (r, recs[r["tag"]])
    core = (slice(2, -2),) * 3
    for r in nreqs:
        print(f"nse request {r['tag']}: {r['kind']} n={r['n']} steps={r['steps']}", flush=True)
        assert recs[r["tag"]]["status"] == S.OK, (r, recs[r["tag"]])
        ref = solo_nse(r)
        names = ("vx", "vy", "vz", "p") + (("t",) if r["kind"] == "rbc" else ())
        for k in names:
            assert np.array_equal(outs[r["tag"]][k][core], ref[k][core]), (r, k)
    for r, r0 in zip(again, nreqs[:4]):
        assert recs2[r["tag"]]["status"] == S.OK
        for k in ("vx", "vy", "vz", "p"):
            assert np.array_equal(outs2[r["tag"]][k][core], outs[r0["tag"]][k][core]), (r, k)
