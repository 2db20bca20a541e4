"""Pins the oracle's matvec path (amx_oracle.c) with checks that a plausible
mistake would fail: the closed-form 27-point stencil of the hex8 Laplacian,
scipy-assembled sparse matrices, identity-row BC algebra, the penalty
diagonal, rigid-body null spaces, and textbook CG."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as orc
from amat_amd import mesh, problems
from conftest import rel_l2


def assembled(gmap, K, n):
    E, ne = gmap.shape
    rows = np.repeat(gmap, ne, axis=1).reshape(-1)
    cols = np.tile(gmap, (1, ne)).reshape(-1)
    return sp.csr_matrix((np.asarray(K).reshape(-1), (rows, cols)), shape=(n, n))


def stencil27(u3, h):
    """Assembled Q1 Laplacian on a uniform grid, interior rows: centre 8h/3,
    6 face neighbours 0, 12 edge neighbours -h/6, 8 corner neighbours -h/12."""
    out = np.zeros_like(u3)
    c = u3[1:-1, 1:-1, 1:-1]
    acc = (8 * h / 3) * c
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                d = abs(di) + abs(dj) + abs(dk)
                if d == 0:
                    continue
                w = 0.0 if d == 1 else (-h / 6 if d == 2 else -h / 12)
                acc = acc + w * u3[1 + dk:u3.shape[0] - 1 + dk, 1 + dj:u3.shape[1] - 1 + dj, 1 + di:u3.shape[2] - 1 + di]
    out[1:-1, 1:-1, 1:-1] = acc
    return out


def test_matvec_hex8_closed_form_stencil():
    n = 12
    p = problems.poisson_hex8(n)
    u = p.random_u(5)
    v = orc.matvec(p.gmap, p.element_matrices(), u)
    u3 = u.reshape(n + 1, n + 1, n + 1)
    ref = stencil27(u3, 1.0 / n)
    np.testing.assert_allclose(v.reshape(u3.shape)[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-14)


def test_matmult_mf_hex8_identity_rows():
    """MatMult_mf: constrained rows return u_c; free rows see u_c = 0."""
    n = 10
    p = problems.poisson_hex8(n)
    u = p.random_u(6)
    vv = orc.matmult_mf(p.gmap, p.element_matrices(), u, p.constrained)
    u3 = u.reshape(n + 1, n + 1, n + 1).copy()
    bnd = np.zeros(u3.shape, bool)
    bnd[[0, -1], :, :] = bnd[:, [0, -1], :] = bnd[:, :, [0, -1]] = True
    u0 = np.where(bnd, 0.0, u3)
    ref = np.where(bnd, u3, stencil27(u0, 1.0 / n))
    np.testing.assert_allclose(vv.reshape(u3.shape), ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize("which", ["hex20", "tet10"])
def test_matvec_matches_scipy_assembly(which):
    p = problems.elasticity_hex20(3) if which == "hex20" else problems.poisson_tet10(3, with_rhs=False)
    K = p.element_matrices()
    u = p.random_u(7)
    A = assembled(p.gmap.astype(np.int64), K, p.n_dofs)
    v1 = orc.matvec(p.gmap, K, u, nthreads=1)
    assert rel_l2(v1, A @ u) < 1e-14
    # column-major VECTORIZED_OPENMP variant and multithreaded atomics agree
    v2 = orc.matvec(p.gmap, np.ascontiguousarray(K.transpose(0, 2, 1)), u, layout=1, nthreads=1)
    v3 = orc.matvec(p.gmap, K, u, nthreads=4)
    assert rel_l2(v2, v1) < 1e-15 and rel_l2(v3, v1) < 1e-15


def test_hex20_translation_in_null_space():
    p = problems.elasticity_hex20(3)
    t = np.zeros((p.n_dofs // 3, 3))
    t[:, 1] = 1.0
    v = orc.matvec(p.gmap, p.element_matrices(), t.reshape(-1))
    assert np.abs(v).max() < 1e-12 * np.abs(p.k_ref).max() * 2


def test_penalty_equals_ebe_plus_diagonal_correction():
    """aMatBased penalty (assembled, diagonal replaced) = EbE K u with
    v_c += (p - K_cc) u_c: two independent paths."""
    p = problems.elasticity_hex20(3)
    K = p.element_matrices()
    maps = orc.set_bdr_map(0, 1, [p.n_dofs], p.gmap, p.constrained, p.prescribed)
    u = p.random_u(8)
    pen = 1e8
    v = orc.matmult_penalty(p.gmap, K, maps["bdr_map"], pen, u)
    d = orc.get_diagonal(p.gmap, K, None, 0, p.n_dofs)
    ref = orc.matvec(p.gmap, K, u)
    c = p.constrained.astype(np.int64)
    ref[c] += (pen - d[c]) * u[c]
    assert rel_l2(v, ref) < 1e-15
    free = np.setdiff1d(np.arange(p.n_dofs), c)
    assert rel_l2(v[free], ref[free]) < 1e-14
    # a planted mistake (correction with the wrong sign) is caught
    bad = ref.copy()
    bad[c] -= 2 * (pen - d[c]) * u[c]
    assert rel_l2(v, bad) > 1e-3


def test_diagonal_and_trace():
    p = problems.poisson_tet10(3, with_rhs=False)
    K = p.element_matrices()
    A = assembled(p.gmap.astype(np.int64), K, p.n_dofs)
    d = orc.get_diagonal(p.gmap, K, None, 0, p.n_dofs)
    np.testing.assert_allclose(d, A.diagonal(), rtol=1e-15)
    maps = orc.set_bdr_map(0, 1, [p.n_dofs], p.gmap, p.constrained, p.prescribed)
    d1 = orc.get_diagonal(p.gmap, K, maps["bdr_map"], 0, p.n_dofs)
    d2 = orc.get_diagonal(p.gmap, K, maps["bdr_map"], 1, p.n_dofs)
    c = p.constrained.astype(np.int64)
    tr = orc.trace(K)
    assert abs(tr - np.trace(K, axis1=1, axis2=2).sum()) <= 1e-12 * tr
    assert (d1[c] == 1.0).all() and np.allclose(d2[c], 100.0 * tr)


def test_transposed_gather_equals_sequential_matvec_bitwise():
    p = problems.poisson_tet10(2, with_rhs=False)
    K = p.element_matrices()
    u = p.random_u(9)
    off, pos = orc.build_ew_maps(p.gmap, p.n_dofs)
    ve = orc.element_products(p.gmap, K, u)
    v = orc.gather_ew(off, pos, ve, p.n_dofs)
    np.testing.assert_array_equal(v, orc.matvec(p.gmap, K, u, nthreads=1))


def test_blocks_match_full_element_matrix():
    rng = np.random.default_rng(4)
    E, bd, nb, N = 5, 2, 3, 14
    n = bd * nb
    lmap = rng.integers(0, N, size=(E, n))
    present = rng.random((E, bd * bd)) < 0.7
    full = np.zeros((E, n, n))
    flat = []
    for e in range(E):
        for bi in range(bd):
            for bj in range(bd):
                if present[e, bi * bd + bj]:
                    B = rng.normal(size=(nb, nb))
                    flat.append(B.reshape(-1))
                    full[e, bi * nb:(bi + 1) * nb, bj * nb:(bj + 1) * nb] = B
    u = rng.normal(size=N)
    vb = orc.matvec_blocks(np.full(E, n), np.full(E, bd), present.reshape(-1), lmap.reshape(-1),
                           np.concatenate(flat), u, N)
    assert rel_l2(vb, orc.matvec(lmap, full, u)) < 1e-14


@pytest.mark.parametrize("op", [0, 1])
def test_pcg_against_direct_solve(op):
    p = problems.poisson_hex8(6) if op == 0 else problems.elasticity_hex20(2)
    K = p.element_matrices()
    maps = orc.set_bdr_map(0, 1, [p.n_dofs], p.gmap, p.constrained, p.prescribed)
    A = assembled(p.gmap.astype(np.int64), K, p.n_dofs).tolil()
    c = p.constrained.astype(np.int64)
    pen = 1e6
    if op == 0:   # identity rows and columns
        A[c, :] = 0
        A[:, c] = 0
        A[c, c] = 1.0
    else:
        A[c, c] = pen
    b = mesh.uniform(3, 3, p.n_dofs, -1, 1)
    x, it, rr = orc.pcg(op, p.gmap, K, maps["bdr_map"], p.constrained, pen, b, rtol=1e-12)
    xd = spla.spsolve(A.tocsc(), b)
    assert rr <= 1e-12 and it > 5
    assert rel_l2(x, xd) < 1e-8
