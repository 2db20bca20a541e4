"""GPU parity of the near-symmetric records of 60-DoF FP64 element matrices.

A record keeps the upper triangle, one byte per lower entry coding its low
32-bit word against the transposed entry, and full values at the group's
exception positions; the element kernel rebuilds every entry and keeps the
whole layout's products and summation order.  Bar: bit-identical to the whole
layout and to the oracle's sequential matvec / MatMult_mf on the same inputs
(deterministic scatter), 1e-13 relative L2 with the atomic scatter.

The matrices are the elasticity element matrices of the examples (69
exception positions: entries that are zero in exact arithmetic and come out
as +-1e-17 of either sign) and synthetic ones that exercise every code: low
words off by -127..127 (both ends included), sign flips, unrelated values,
entries whose low word would carry into the high word, no exception at all,
and one exception more than a record holds (the group then keeps whole
matrices).
"""
import dataclasses

import numpy as np
import pytest

import oracle as orc
from amat_amd import BCOP, PREC, SCATTER, Maps, MultiRHS, aMatFree, fem, mesh, problems
from conftest import rel_l2

pytestmark = pytest.mark.gpu

N = 60
MAX_EXCEPTIONS = 346   # exception values a record holds for n = 60


def run(torch, fn, u, dtype=np.float64):
    du = torch.from_numpy(np.ascontiguousarray(u.astype(dtype))).cuda()
    dv = torch.empty_like(du)
    fn(dv, du, torch.cuda.current_stream())
    return dv.cpu().numpy()


def near_symmetric(seed, E, n_exc, carry=True):
    """E symmetric matrices whose lower entries are then moved off their
    transposed entry: the low word by a code in [-127, 127] (the ends at fixed
    positions), and at n_exc positions (the same in every matrix) to values
    the byte cannot code."""
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(E, N, N))
    A = A + A.transpose(0, 2, 1)                      # bitwise symmetric
    r, c = np.tril_indices(N, -1)
    bits = A.view(np.uint64)
    up = bits[:, c, r]                                # transposed (upper) entries
    lo = (up & np.uint64(0xFFFFFFFF)).astype(np.int64)
    d = rng.integers(-127, 128, size=up.shape)
    d[:, 0], d[:, 1] = 127, -127
    ok = (lo + d >= 0) & (lo + d <= 0xFFFFFFFF)       # no carry: codable
    d = np.where(ok, d, 0)
    low = (up & np.uint64(0xFFFFFFFF00000000)) | (lo + d).astype(np.uint64)
    pick = rng.choice(np.arange(2, len(r)), size=n_exc, replace=False)
    for k, j in enumerate(pick):
        kind = k % 4
        if kind == 0:      # sign flip
            low[:, j] = (-A[:, c[j], r[j]]).view(np.uint64)
        elif kind == 1:    # tiny, either sign
            low[:, j] = (rng.normal(size=E) * 1e-17).view(np.uint64)
        elif kind == 2:    # unrelated value
            low[:, j] = rng.normal(size=E).view(np.uint64)
        else:              # low word off by more than a byte codes
            low[:, j] = up[:, j] + np.uint64(300)
    B = A.copy()
    B.view(np.uint64)[:, r, c] = low
    if carry and n_exc > 0:
        # one position where the low word would carry: the high words differ
        j = pick[0]
        for e in range(E):
            x = np.array([B[e, c[j], r[j]]]).view(np.uint64)
            x[0] = (x[0] & np.uint64(0xFFFFFFFF00000000)) | np.uint64(0xFFFFFFF0)
            B[e, c[j], r[j]] = x.view(np.float64)[0]
            B.view(np.uint64)[e, r[j], c[j]] = x[0] + np.uint64(32)
    return B


def with_matrices(p, K):
    return dataclasses.replace(p, K=K, k_ref=None, scale=None)


def ops_pair(p, **kw):
    rec = p.build_operator(aMatFree, Maps, sym_storage=1, **kw)
    whole = p.build_operator(aMatFree, Maps, sym_storage=0, **kw)
    return rec, whole


def check_bitexact(torch, p, rec, whole):
    K = p.element_matrices()
    u = p.random_u()
    v = run(torch, rec.matvec_dev, u)
    np.testing.assert_array_equal(v, run(torch, whole.matvec_dev, u))
    np.testing.assert_array_equal(v, orc.matvec(p.gmap, K, u))
    vm = run(torch, rec.matmult_dev, u)
    np.testing.assert_array_equal(vm, run(torch, whole.matmult_dev, u))
    np.testing.assert_array_equal(vm, orc.matmult_mf(p.gmap, K, u, p.constrained))
    np.testing.assert_array_equal(rec.get_diagonal(), whole.get_diagonal())
    assert rec.info().trace == whole.info().trace


def test_elasticity_matrices_records(torch_gpu):
    """The examples' hex20 elasticity matrices (K_e = E_e K_ref): records
    with 69 exception positions, about 0.61 of the whole bytes."""
    p = problems.elasticity_hex20(5, 4, 3)
    rec, whole = ops_pair(p)
    ir, iw = rec.info(), whole.info()
    assert ir.near_symmetric == 1 and iw.near_symmetric == 0
    assert ir.k_bytes_fp64 < 0.62 * iw.k_bytes_fp64
    check_bitexact(torch_gpu, p, rec, whole)
    pen, penw = ops_pair(p, bcop=BCOP.PENALTY_DIAG, penalty=1e8)
    u = p.random_u()
    np.testing.assert_array_equal(run(torch_gpu, pen.matmult_dev, u), run(torch_gpu, penw.matmult_dev, u))


def test_distorted_elements_records(torch_gpu):
    """Element matrices from ke_hex20_iso on jittered nodes (every element
    different); bit-identical whichever layout the group gets."""
    p = problems.elasticity_hex20(3)
    rng = np.random.default_rng(8)
    xe = p.mesh.coords[p.mesh.elem_nodes] + rng.uniform(-0.02, 0.02, size=(p.n_elem, 20, 3)) / 3
    K = np.stack([fem.ke_hex20_iso(x, 1.0, 0.3, 3) for x in xe])
    q = with_matrices(p, K)
    rec, whole = ops_pair(q)
    print("near-symmetric groups:", rec.info().near_symmetric)
    check_bitexact(torch_gpu, q, rec, whole)


@pytest.mark.parametrize("n_exc", [0, 57, MAX_EXCEPTIONS])
def test_synthetic_codes(torch_gpu, n_exc):
    p = problems.elasticity_hex20(4, 3, 2)
    q = with_matrices(p, near_symmetric(n_exc + 1, p.n_elem, n_exc))
    rec, whole = ops_pair(q)
    assert rec.info().near_symmetric == 1
    check_bitexact(torch_gpu, q, rec, whole)
    # identity BC operator (masked gather of u) and the atomic scatter
    u = q.random_u()
    ato = q.build_operator(aMatFree, Maps, sym_storage=1, scatter=SCATTER.ATOMIC)
    assert ato.info().near_symmetric == 1
    assert rel_l2(run(torch_gpu, ato.matmult_dev, u), run(torch_gpu, whole.matmult_dev, u)) <= 1e-13


def test_one_exception_too_many_keeps_whole_matrices(torch_gpu):
    p = problems.elasticity_hex20(3, 2, 2)
    q = with_matrices(p, near_symmetric(5, p.n_elem, MAX_EXCEPTIONS + 1, carry=False))
    rec, whole = ops_pair(q)
    assert rec.info().near_symmetric == 0
    assert rec.info().k_bytes_fp64 == whole.info().k_bytes_fp64
    check_bitexact(torch_gpu, q, rec, whole)


def test_multi_rhs_layout_from_records(torch_gpu):
    """The multi-RHS storage built from records equals the one built from
    whole matrices: the deterministic products are bit-identical."""
    p = problems.elasticity_hex20(3, 3, 2)
    q = with_matrices(p, near_symmetric(9, p.n_elem, 120))
    rec, whole = ops_pair(q)
    assert rec.info().near_symmetric == 1
    nrhs = 8
    U = mesh.uniform(3, 1, q.n_dofs * nrhs, -1.0, 1.0).astype(np.float32).reshape(q.n_dofs, nrhs)
    outs = []
    for op in (rec, whole):
        mr = MultiRHS([op], PREC.FP32)
        mr.set_scatter(SCATTER.DETERMINISTIC)
        dU = torch_gpu.from_numpy(U).cuda()
        dV = torch_gpu.full_like(dU, float("nan"))
        mr.apply_dev(nrhs, [dU], [dV], torch_gpu.cuda.current_stream())
        outs.append(dV.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("prec", [PREC.FP32, PREC.BF16])
@pytest.mark.parametrize("case", ["elasticity", "synthetic", "exact", "too_many"])
def test_low_precision_records(torch_gpu, prec, case):
    """FP32 / BF16 storage of 60-DoF groups: rounded, the matrices are bitwise
    symmetric except at the group's exception positions, so the records hold
    the upper triangle and those values; products bit-identical to the whole
    layout with FP32 vectors, FP64 vectors (the inner solves of
    AMX_IR_F64_VECTORS, through the solver's result) and the atomic scatter
    within tolerance."""
    p = problems.elasticity_hex20(4, 3, 2)
    if case == "synthetic":
        p = with_matrices(p, near_symmetric(21, p.n_elem, 57))
    elif case == "exact":      # exactly symmetric: no exception at all
        K = p.element_matrices()
        p = with_matrices(p, K + K.transpose(0, 2, 1))
    elif case == "too_many":   # more exceptions than a record holds: whole matrices
        rng = np.random.default_rng(5)
        p = with_matrices(p, rng.normal(size=(p.n_elem, N, N)))
    kw = dict(precision=PREC.FP64 | PREC.FP32 | PREC.BF16)
    rec, whole = ops_pair(p, **kw)
    ir, iw = rec.info(), whole.info()
    kb = "k_bytes_fp32" if prec == PREC.FP32 else "k_bytes_bf16"
    if case == "too_many":
        assert getattr(ir, kb) == getattr(iw, kb)
    else:
        assert getattr(ir, kb) < 0.56 * getattr(iw, kb)
    u = p.random_u()
    for fn in ("matmult_f32_dev",):
        a = run(torch_gpu, lambda v, uu, s: rec.matmult_f32_dev(prec, v, uu, s), u, np.float32)
        b = run(torch_gpu, lambda v, uu, s: whole.matmult_f32_dev(prec, v, uu, s), u, np.float32)
        np.testing.assert_array_equal(a, b)
    ato = p.build_operator(aMatFree, Maps, sym_storage=1, scatter=SCATTER.ATOMIC, **kw)
    a = run(torch_gpu, lambda v, uu, s: ato.matmult_f32_dev(prec, v, uu, s), u, np.float32)
    b = run(torch_gpu, lambda v, uu, s: whole.matmult_f32_dev(prec, v, uu, s), u, np.float32)
    assert rel_l2(a, b) <= 1e-6
    # FP64 inner vectors with FP32 / BF16 storage (SPD cases): the solves take the same path bit for bit
    from amat_amd.amat import IR_F64_VECTORS
    if case in ("synthetic", "too_many"):
        return
    b64 = p.random_u(seed=9)
    b64[p.constrained.astype(np.int64)] = 0.0
    xs = []
    for op in (rec, whole):
        db = torch_gpu.from_numpy(b64).cuda()
        dx = torch_gpu.zeros_like(db)
        st = op.ir_cg(int(prec) | IR_F64_VECTORS, db, dx, rtol=1e-8, inner_rtol=1e-2, max_outer=30, max_inner=3000,
                      stream=torch_gpu.cuda.current_stream())
        xs.append((dx.cpu().numpy(), st.iterations))
    np.testing.assert_array_equal(xs[0][0], xs[1][0])
    assert xs[0][1] == xs[1][1]
