"""The allocation-free assembly entry point (pyhip_assemble_noalloc) against
pyhip_assemble, through the C ABI: the same residual and CSR values bit for bit, on two
calls in a row (the second on a smaller problem, so the shared static scratch holds the
first call's data), in FP64 and FP32, residual-only and with the Jacobian; a problem whose
element buffers exceed the static scratch is rejected with a clear error."""
import ctypes

import numpy as np
import pytest

from pyoomph_amd import _lib
from pyoomph_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _assemble(P, fn, flag, precision):
    import torch
    L = _lib.load()
    vdt = torch.float64 if precision == _lib.F64 else torch.float32
    with torch.cuda.stream(P._stream):
        res = torch.full((max(P.ndof, 1),), float("nan"), dtype=torch.float64, device=P.device)
        vals = torch.full((max(P.nnz, 1),), float("nan"), dtype=vdt, device=P.device)
    P._stream.synchronize()
    rc = getattr(L, fn)(P._handle, _lib.ptr(P.dofs), flag, -1, precision, _lib.ptr(res),
                        _lib.ptr(vals) if flag else None)
    P._stream.synchronize()
    _lib.check(rc)
    return res.cpu().numpy(), vals.cpu().numpy()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_assemble_on_two_calls(gpu, precision):
    from pyoomph_amd.problem import Problem
    big = Problem(wl.bratu(40, "C2", 2.0))            # first call: the larger problem
    small = Problem(wl.lid_driven_cavity(6))          # second call: smaller, other form
    for P, seed in ((big, 1), (small, 2)):
        x = np.random.default_rng(seed).uniform(-0.3, 0.3, P.ndof)
        P.set_current_dofs(x)
        for flag in (1, 0):
            r_ref, v_ref = _assemble(P, "pyhip_assemble", flag, precision)
            r_new, v_new = _assemble(P, "pyhip_assemble_noalloc", flag, precision)
            assert np.array_equal(r_new, r_ref), (P.ndof, flag)
            if flag:
                assert np.array_equal(v_new, v_ref), (P.ndof, flag)


def test_noalloc_rejects_a_problem_beyond_the_scratch(gpu):
    from pyoomph_amd.problem import Problem
    # 23^3 Q2 bricks: 12167 x (27 + 27^2) doubles = 73.6 MB > 64 MiB
    P = Problem(wl.reaction_diffusion_cube(23))
    L = _lib.load()
    with pytest.raises(RuntimeError, match="static scratch"):
        _assemble(P, "pyhip_assemble_noalloc", 1, _lib.F64)
    # the residual alone fits (12167 x 27 doubles)
    r_ref, _ = _assemble(P, "pyhip_assemble", 0, _lib.F64)
    r_new, _ = _assemble(P, "pyhip_assemble_noalloc", 0, _lib.F64)
    assert np.array_equal(r_new, r_ref)
    assert L.pyhip_last_error() is not None
