"""Serving loop on the GPU: a Poisson-arrival stream of mixed requests (P2 Poisson, Q2
reaction-diffusion, Taylor-Hood Navier-Stokes) served with HBM-budget admission and
continuous batching on two streams. Every result must equal the same request solved
alone (bitwise: all reductions are per sub-problem) and agree with the oracle to 1e-8."""
import numpy as np
import pytest

from oracle import newton as on
from tests.oracle_bridge import oracle_problem

pytestmark = pytest.mark.gpu


def _oracle(req, mp):
    if req.kind == "poisson":
        op = oracle_problem("poisson", N=req.N, seed=req.seed)
        tol = 1e-8
    elif req.kind == "rd":
        op = oracle_problem("rd", N=req.N, f=mp.equations.source_field)
        tol = 1e-8
    else:
        op = oracle_problem("ns", N=req.N, Re=100.0)
        tol = 1e-10
    return on.newton_solve(op, np.zeros(op.ndof), tol=tol)


def test_served_results_equal_alone_and_oracle(gpu):
    from pyoomph_amd.serving import GpuExecutor, ServingLoop, poisson_stream, summarize
    reqs = poisson_stream(rate=400.0, n=14, seed=0, dof_range=(800, 9000), ns_max_dofs=2500)
    assert {r.kind for r in reqs} == {"poisson", "rd", "ns"}
    ex = GpuExecutor()
    loop = ServingLoop(ex, hbm_budget=8e9, workers=2, max_batch_dofs=40_000)
    res = loop.run(reqs)
    assert any(b[1][1:] for b in loop.batches), "no request was batched with another"
    for req, r in zip(sorted(reqs, key=lambda q: q.rid), res):
        print("request", req.rid, req.kind, req.N, "batch", r.batch, r.batch_size, flush=True)
        assert r.stats["converged"], (req, r.stats)
        x_alone, st_alone = ex.run_alone(req)
        assert np.array_equal(r.dofs, x_alone), req
        assert r.stats["newton_iters"] == st_alone["newton_iters"]
    # the oracle on one request of each kind
    done = set()
    for req, r in zip(sorted(reqs, key=lambda q: q.rid), res):
        if req.kind in done:
            continue
        done.add(req.kind)
        xo = _oracle(req, req.mesh_problem())
        assert np.max(np.abs(r.dofs - xo)) <= 1e-8 * max(np.max(np.abs(xo)), 1e-300), req
    s = summarize(res)
    assert s["requests"] == len(reqs) and loop.peak_reserved <= 8e9
