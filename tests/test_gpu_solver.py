"""Device solvers: FP64 Jacobi-PCG against a direct solve and the oracle's
PCG; FP64 iterative refinement with FP32 / BF16 inner solves reaching the
same 1e-10 relative residual, with their distance to the FP64 solution."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as orc
from amat_amd import BCOP, PREC, Maps, aMatFree, mesh, problems
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def direct(p, bcop, pen=None):
    K = p.element_matrices()
    E, ne = p.gmap.shape
    g = p.gmap.astype(np.int64)
    A = sp.csr_matrix((K.reshape(-1), (np.repeat(g, ne, 1).reshape(-1), np.tile(g, (1, ne)).reshape(-1))),
                      shape=(p.n_dofs, p.n_dofs)).tolil()
    c = p.constrained.astype(np.int64)
    if bcop == BCOP.IDENTITY:
        A[c, :] = 0
        A[:, c] = 0
        A[c, c] = 1.0
    else:
        A[c, c] = pen
    return A.tocsc()


@pytest.mark.parametrize("which", ["hex8", "hex20_penalty", "tet10"])
def test_fp64_cg(torch_gpu, which):
    torch = torch_gpu
    if which == "hex8":
        p, bcop, pen = problems.poisson_hex8(10), BCOP.IDENTITY, None
    elif which == "hex20_penalty":
        p, bcop, pen = problems.elasticity_hex20(3), BCOP.PENALTY_DIAG, 1e6
    else:
        p, bcop, pen = problems.poisson_tet10(4), BCOP.IDENTITY, None
    op = p.build_operator(aMatFree, Maps, bcop=bcop, penalty=pen)
    b = p.rhs if p.rhs is not None else mesh.uniform(5, 5, p.n_dofs, -1, 1)
    if bcop == BCOP.IDENTITY:
        b = b.copy()
        b[p.constrained.astype(np.int64)] = 0.0
    db = dev(torch, b)
    dx = torch.zeros_like(db)
    st = op.cg(db, dx, rtol=1e-10, max_it=5000, stream=torch.cuda.current_stream())
    x = dx.cpu().numpy()
    assert st.converged and st.rel_residual <= 1e-10, (st.iterations, st.rel_residual)
    xd = spla.spsolve(direct(p, bcop, pen), b)
    assert rel_l2(x, xd) < 1e-7
    maps = orc.set_bdr_map(0, 1, [p.n_dofs], p.gmap, p.constrained, p.prescribed)
    xo, it, rr = orc.pcg(0 if bcop == BCOP.IDENTITY else 1, p.gmap, p.element_matrices(), maps["bdr_map"],
                         p.constrained, pen or 0.0, b, rtol=1e-10)
    assert abs(it - st.iterations) <= max(2, 0.05 * it)   # same method; dot-product order differs
    assert rel_l2(x, xo) < 1e-7


@pytest.mark.parametrize("prec", [PREC.FP32, PREC.BF16])
@pytest.mark.parametrize("which", ["hex8", "tet10"])
def test_ir_cg(torch_gpu, prec, which):
    torch = torch_gpu
    p = problems.poisson_hex8(12) if which == "hex8" else problems.poisson_tet10(5)
    op = p.build_operator(aMatFree, Maps, precision=PREC.FP64 | prec)
    b = (p.rhs if p.rhs is not None else mesh.uniform(6, 6, p.n_dofs, -1, 1)).copy()
    b[p.constrained.astype(np.int64)] = 0.0
    db = dev(torch, b)
    x64 = torch.zeros_like(db)
    s64 = op.cg(db, x64, rtol=1e-10, stream=torch.cuda.current_stream())
    xir = torch.zeros_like(db)
    sir = op.ir_cg(prec, db, xir, rtol=1e-10, inner_rtol=1e-2, stream=torch.cuda.current_stream())
    err = rel_l2(xir.cpu().numpy(), x64.cpu().numpy())
    print(f"IR-CG {prec.name} {which}: outer={sir.outer_iterations} inner={sir.iterations} "
          f"(FP64 CG {s64.iterations}) relres={sir.rel_residual:.2e} err_vs_fp64={err:.2e}")
    assert sir.converged and sir.rel_residual <= 1e-10
    assert err < 1e-8


@pytest.mark.parametrize("prec", [PREC.FP32, PREC.BF16])
@pytest.mark.parametrize("which", ["hex8", "tet10", "hex20_penalty"])
def test_reliable_update_cg(torch_gpu, prec, which):
    """Mixed-precision CG with reliable updates: the low-precision recurrence is
    never restarted, so it takes about the FP64 iteration count; convergence
    is decided on the FP64 true residual."""
    torch = torch_gpu
    from amat_amd.amat import IR_RELIABLE
    if which == "hex8":
        p, bcop, pen = problems.poisson_hex8(12), BCOP.IDENTITY, None
    elif which == "tet10":
        p, bcop, pen = problems.poisson_tet10(5), BCOP.IDENTITY, None
    else:
        p, bcop, pen = problems.elasticity_hex20(3), BCOP.PENALTY_DIAG, 1e6
    op = p.build_operator(aMatFree, Maps, precision=PREC.FP64 | prec, bcop=bcop, penalty=pen)
    b = (p.rhs if p.rhs is not None else mesh.uniform(6, 6, p.n_dofs, -1, 1)).copy()
    if bcop == BCOP.IDENTITY:
        b[p.constrained.astype(np.int64)] = 0.0
    db = dev(torch, b)
    x64 = torch.zeros_like(db)
    s64 = op.cg(db, x64, rtol=1e-10, stream=torch.cuda.current_stream())
    xr = torch.zeros_like(db)
    sr = op.ir_cg(int(prec) | IR_RELIABLE, db, xr, rtol=1e-10, inner_rtol=0.1, max_outer=200, max_inner=20000,
                  stream=torch.cuda.current_stream())
    err = rel_l2(xr.cpu().numpy(), x64.cpu().numpy())
    print(f"RU-CG {prec.name} {which}: replacements={sr.outer_iterations} it={sr.iterations} "
          f"(FP64 CG {s64.iterations}) relres={sr.rel_residual:.2e} err_vs_fp64={err:.2e}")
    assert sr.converged and sr.rel_residual <= 1e-10
    assert err < 1e-7
    # FP32: about the FP64 count; BF16's operator error (2^-9 per entry) slows
    # the recurrence on the stiffer elasticity problem (2x measured)
    assert sr.iterations <= (1.35 if prec == PREC.FP32 else 2.5) * s64.iterations + 10
    assert sr.precision == int(prec) | IR_RELIABLE
    # the FP64 solution against scipy on the assembled system
    xd = spla.spsolve(direct(p, bcop, pen), b)
    assert rel_l2(xr.cpu().numpy(), xd) < 1e-7


def test_fp64_cg_residual_replacement(torch_gpu):
    """FP64 CG replaces the stopped recurrence's residual by the true b - A x
    and iterates on while that misses rtol: at a tolerance below what the
    recurrence can certify, replacements run (up to the limit), the reported
    residual is the true one, and the served job takes the same path
    bit for bit."""
    torch = torch_gpu
    from amat_amd import serve
    from amat_amd.serve import Server
    p = serve.build_problem("hex20", 3)
    op = serve.build_operator(p)
    b = serve.rhs_for(p, 11)
    db = torch.from_numpy(b).cuda()
    x = torch.zeros_like(db)
    st = op.cg(db, x, rtol=1e-16, max_it=20000)
    assert st.outer_iterations >= 1          # the first true residual missed rtol
    assert not st.converged and 1e-16 < st.rel_residual < 1e-12
    xs = x.cpu().numpy()
    srv = Server(max_jobs=2)
    j = srv.submit(op, PREC.FP64, b, rtol=1e-16)
    out = {}
    while not out:
        srv.step(8)
        for k in srv.poll():
            out[k] = srv.result(k, p.n_dofs)
    x2, st2 = out[j]
    np.testing.assert_array_equal(x2, xs)
    assert (st2.iterations, st2.outer_iterations, st2.converged) == (st.iterations, st.outer_iterations, st.converged)
    # at an attainable tolerance the first replacement confirms convergence
    x.zero_()
    st3 = op.cg(db, x, rtol=1e-10, max_it=20000)
    assert st3.converged and st3.outer_iterations == 0 and st3.rel_residual <= 1e-10


def _tree(a):
    """The block tree of the device reduction: a[i] += a[i + s] for s = 128 .. 1."""
    a = a.copy()
    s = a.shape[-1] // 2
    while s > 0:
        a = a[..., :s] + a[..., s:2 * s]
        s //= 2
    return a[..., 0]


def host_dot_fixed_order(x, n_blocks):
    """sum(x) in the summation order of the solvers' reduction over n_blocks
    blocks of 256 threads: per-thread strided sums, block trees, group trees
    (32 blocks, zero-padded to 256), the tree over the group sums."""
    S = n_blocks * 256
    rows = max(1, -(-len(x) // S))
    xs = np.zeros(rows * S)
    xs[:len(x)] = x
    acc = np.zeros(S)
    for k in range(rows):
        acc = acc + xs[k * S:(k + 1) * S]
    b = _tree(acc.reshape(n_blocks, 256))
    n_groups = -(-n_blocks // 32)
    gs = np.zeros(n_groups)
    for g in range(n_groups):
        pad = np.zeros(256)
        part = b[32 * g:32 * (g + 1)]
        pad[:len(part)] = part
        gs[g] = _tree(pad)
    pad = np.zeros(256)
    pad[:n_groups] = gs
    return _tree(pad)


def test_reduction_fixed_order_sweep(torch_gpu):
    """The fence-free cross-block reduction of the CG kernels, checked bit for
    bit against a host sum in the same fixed order, for sizes around block and
    group boundaries and grids of 1 to 1024 blocks, each reduction repeated 16
    times in a row on one partials buffer (every repeat equal): a deposit seen
    late by the finishing block, or a counter left non-zero, changes a sum.
    y = 1, so x.y is sum(x) and y.y = n exactly whatever the contraction of
    the products.  Stops at the first mismatch."""
    import ctypes as C

    from amat_amd._capi import lib

    torch = torch_gpu
    L = lib()
    rng = np.random.default_rng(11)
    repeat = 16
    for n in [1, 255, 256, 257, 8191, 32768, 65537, 262143, 1 << 20, (1 << 22) + 17]:
        x = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))  # wide exponents: order-sensitive sums
        xd = dev(torch, x)
        yd = torch.ones(n, dtype=torch.float64, device="cuda")
        for nb in [1, 2, 31, 32, 33, 64, 100, 257, 1000, 1024]:
            out = torch.full((2 * repeat,), np.nan, dtype=torch.float64, device="cuda")
            rc = L.amx_dev_dot2(C.c_void_p(xd.data_ptr()), C.c_void_p(yd.data_ptr()), n, nb, repeat,
                                C.c_void_p(out.data_ptr()), 0, None)
            assert rc == 0
            got = out.cpu().numpy().reshape(repeat, 2)
            want = host_dot_fixed_order(x, nb)
            bad = [k for k in range(repeat) if got[k, 0] != want or got[k, 1] != float(n)]
            assert not bad, (n, nb, bad[:4], got[bad[0]].tolist(), want)
            # and the fixed-order sum is a sum: close to the compensated one
            assert abs(want - math.fsum(x)) <= 1e-10 * np.abs(x).sum(), (n, nb)


def test_reduction_rejects_bad_arguments(torch_gpu):
    import ctypes as C

    from amat_amd._capi import lib

    torch = torch_gpu
    L = lib()
    xd = torch.ones(10, dtype=torch.float64, device="cuda")
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    p = C.c_void_p(xd.data_ptr())
    o = C.c_void_p(out.data_ptr())
    assert L.amx_dev_dot2(p, p, 10, 0, 1, o, 0, None) != 0
    assert L.amx_dev_dot2(p, p, 10, 1025, 1, o, 0, None) != 0
    assert L.amx_dev_dot2(p, p, 10, 4, 0, o, 0, None) != 0
    assert L.amx_dev_dot2(p, p, 10, 4, 1, None, 0, None) != 0
    assert L.amx_dev_dot2(p, p, 10, 4, 1, o, 0, None) == 0
    assert out.cpu().tolist() == [10.0, 10.0]
