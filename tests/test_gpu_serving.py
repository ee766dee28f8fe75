"""Serving loop on the GPU: a Poisson-arrival stream of mixed requests (P2 Poisson, Q2
reaction-diffusion, Taylor-Hood Navier-Stokes at several mesh sizes) served by slot pools
with Newton-step continuous batching on two streams. Requests join pools that are already
stepping; every result equals the same request solved alone in a one-slot pool bit for bit
(every reduction, test and preconditioner is per sub-problem) and agrees with the oracle's
converged solution to 1e-8 relative."""
import numpy as np
import pytest

from oracle import newton as on
from tests.oracle_bridge import oracle_problem

pytestmark = pytest.mark.gpu

SMALL = {"poisson": (8, 12), "rd": (3, 4), "ns": (6, 8)}


def _oracle(req):
    from pyoomph_amd.serving import ns_reynolds
    if req.kind == "poisson":
        op = oracle_problem("poisson", N=req.N, seed=req.seed)
    elif req.kind == "rd":
        op = oracle_problem("rd", N=req.N, f=req.payload[1])
    else:
        op = oracle_problem("ns", N=req.N, Re=ns_reynolds(req.seed))
    return on.converged_solve(op, np.zeros(op.ndof), tol=1e-12)


@pytest.mark.parametrize("cg_precond", ["amg", "jacobi"])
def test_served_results_equal_alone_and_oracle(gpu, cg_precond):
    from pyoomph_amd.serving import GpuPools, ServingLoop, attach_payloads, poisson_stream, summarize
    reqs = attach_payloads(poisson_stream(rate=300.0, n=18, seed=0, sizes=SMALL))
    assert {r.kind for r in reqs} == {"poisson", "rd", "ns"}
    ex = GpuPools(cg_precond=cg_precond)
    loop = ServingLoop(ex, hbm_budget=8e9, workers=2, slots=lambda k, N: 3)
    loop.prepare(sorted({r.signature for r in reqs}))
    res = loop.run(reqs)
    # some request joined a pool while another request was in it
    assert any(a.pool == b.pool and a.admitted < b.admitted < a.done for a in res for b in res)
    for req, r in zip(sorted(reqs, key=lambda q: q.rid), res):
        print("request", req.rid, req.kind, req.N, "slot", r.slot, r.stats, flush=True)
        assert r.stats["converged"], (req, r.stats)
        x_alone, st_alone = ex.solve_alone(req)
        assert np.array_equal(r.dofs, x_alone), req
        assert r.stats["newton_iters"] == st_alone["newton_iters"]
        xo = _oracle(req)
        assert np.max(np.abs(r.dofs - xo)) <= 1e-8 * np.max(np.abs(xo)), req
    s = summarize(res)
    assert s["requests"] == len(reqs) and s["converged"] == len(reqs)
    assert loop.reserved <= 8e9
    loop.close()


@pytest.mark.parametrize("kind,N,precond", [("poisson", 8, "amg"), ("rd", 3, "jacobi"), ("ns", 6, None)])
def test_a_failed_request_leaves_its_pool_usable(gpu, kind, N, precond):
    """A request whose data is NaN fails (its Newton residual is not finite) while the
    request beside it converges; the slot it leaves holds NaN element rows and factors,
    and the next request in that slot converges with the bits it has alone."""
    from pyoomph_amd.serving import GpuPools, Request, SlotPool
    ex = GpuPools(cg_precond=precond)
    pool = SlotPool(ex, (kind, N), 2, warm=False)
    good = [Request(i, kind, N, 11 + i, 0.0) for i in range(2)]
    bad = Request(9, kind, N, 99, 0.0)
    params, aux = bad.data(pool.mesh)
    if kind == "rd":
        aux = aux.copy()
        aux[len(aux) // 2] = np.nan
    else:
        params = params.copy()
        params[-1] = np.nan
    bad.payload = (params, aux)

    def run(loads):
        for s, r in loads:
            pool.load(s, r)
        out = {}
        for _ in range(40):
            for s in pool.step():
                out[s] = pool.take(s)
            if len(out) == len(loads):
                return out
        raise AssertionError("pool did not finish")

    out = run([(0, bad), (1, good[0])])
    assert not out[0][1]["converged"]
    assert out[1][1]["converged"]
    assert np.array_equal(out[1][0], ex.solve_alone(good[0])[0])
    out = run([(0, good[1])])
    assert out[0][1]["converged"]
    assert np.array_equal(out[0][0], ex.solve_alone(good[1])[0])
