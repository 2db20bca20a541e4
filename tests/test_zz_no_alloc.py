"""The allocation-free FP64 CG entry point (amx_cg_solve_noalloc_dev).

It has no counterpart in the reference: amx_cg_solve_dev with its workspace
in a fixed device array, for callers that cannot allocate.  Bar: through the
C ABI, two calls in a row -- the second on a smaller operator, so the array
holds the first call's contents -- return the same solution bit for bit and
the same iteration count as amx_cg_solve_dev on the same inputs (same
arithmetic, same reduction grid), and free device memory does not drop
during the second call, made on an operator that has never solved (the
first call on a device maps the code object that holds the array); an
operator above the stated maximum is rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from amat_amd import AmatError, Maps, aMatFree, problems
from amat_amd import _capi
from amat_amd.amat import NOALLOC_MAX_DOFS

pytestmark = pytest.mark.gpu


def solve(torch, op, b, fn):
    db = torch.from_numpy(b).cuda()
    dx = torch.full_like(db, 3.0)       # the solver starts from x = 0 whatever x holds
    st = _capi.SolveStats()
    s = torch.cuda.current_stream()
    _capi.call(fn, op.handle, C.c_void_p(db.data_ptr()), C.c_void_p(dx.data_ptr()), 1e-10, 20000, C.byref(st),
               C.c_void_p(s.cuda_stream))
    torch.cuda.synchronize()
    return dx.cpu().numpy(), st


def rhs(p, seed):
    b = p.random_u(seed)
    b[p.constrained.astype(np.int64)] = 0.0
    return b


def test_noalloc_cg_matches_cg_on_two_calls(torch_gpu):
    torch = torch_gpu
    big = problems.elasticity_hex20(5)      # 7,986 DoFs
    small = problems.poisson_hex8(9)        # 1,000 DoFs, after the larger call
    ops = [p.build_operator(aMatFree, Maps) for p in (big, small)]
    for k, (p, op, seed) in enumerate(zip((big, small), ops, (21, 22))):
        b = rhs(p, seed)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        x1, s1 = solve(torch, op, b, "amx_cg_solve_noalloc_dev")
        free1 = torch.cuda.mem_get_info()[0]
        if k == 1:
            assert free0 - free1 < (1 << 20), f"device memory fell by {free0 - free1} bytes during the call"
        x0, s0 = solve(torch, op, b, "amx_cg_solve_dev")
        assert s1.converged and s1.rel_residual <= 1e-10
        assert s1.iterations == s0.iterations and s1.outer_iterations == s0.outer_iterations
        np.testing.assert_array_equal(x1, x0)
        assert s1.rel_residual == s0.rel_residual


def test_noalloc_cg_rejects_large_operators(torch_gpu):
    torch = torch_gpu
    p = problems.poisson_hex8(102)          # 103^3 = 1,092,727 DoFs > NOALLOC_MAX_DOFS
    assert p.n_dofs > NOALLOC_MAX_DOFS
    op = p.build_operator(aMatFree, Maps)
    b = torch.zeros(p.n_dofs, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    with pytest.raises(AmatError) as e:
        op.cg_noalloc(b, x)
    assert e.value.name == "INVALID_ARGUMENT" and "at most" in str(e.value)
