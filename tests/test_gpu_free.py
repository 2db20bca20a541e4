"""Method 2 (FREE): element contributions computed in-kernel from coordinates.

Pins: (1) one matvec over disjoint copies of an element recovers K_e column by
column, which must equal the reference's own ke_hex8 / ke_hex20_iso / ke_tet10
outputs (fixtures in tests/golden/fem_ref.npz, written by the reference
sources); (2) on distorted meshes the FREE operator equals the oracle's
stored-matrix matvec with K_e from the element routines, which are pinned to
the same fixtures.  FREE forms K_e u_e in a different operation order than
K_e followed by a GEMV, so the bar is 1e-13 relative L2 (1e-12 relative to
max |K_e| per entry), not bit-exactness; the deterministic scatter is still
bit-identical from run to run.
"""
import dataclasses

import numpy as np
import pytest

import oracle as orc
from amat_amd import BCOP, ELEM, SCATTER, AmatError, Maps, MultiRHS, PREC, aMatFree, fem, mesh, problems
from amat_amd.serve import Server
from conftest import GOLDEN, rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(torch, fn, u):
    du = dev(torch, u)
    dv = torch.empty_like(du)
    fn(dv, du, torch.cuda.current_stream())
    return dv.cpu().numpy()


def one_rank_maps(gmap, n_dofs, constrained=()):
    m = Maps(0, 1)
    E = gmap.shape[0]
    m.set_map(E, gmap.astype(np.uint32), np.full(E, gmap.shape[1], np.uint32), n_dofs,
              np.arange(n_dofs, dtype=np.uint64), 0, n_dofs - 1, n_dofs)
    c = np.asarray(constrained, dtype=np.uint64)
    m.set_bdr_map(c, np.zeros(len(c)))
    return m


def fixture_cases():
    z = np.load(GOLDEN / "fem_ref.npz")
    ref = {k: z[k] for k in z.files}
    cases = []
    for key in sorted(k for k in ref if k.endswith("_ke") and not k.startswith("hex8eig")):
        parts = key.split("_")
        xe = ref[key[:-3] + "_xe"]
        if parts[0] == "hex20":
            cases.append((key, ELEM.HEX20_ELASTIC, xe, float(parts[2][1:]), float(parts[3][2:]), int(parts[4][3:]),
                          ref[key]))
        elif parts[0] == "hex8":
            cases.append((key, ELEM.HEX8_LAPLACE, xe, 1.0, 0.3, int(parts[2][3:]), ref[key]))
        else:
            cases.append((key, ELEM.TET10_LAPLACE, xe, 1.0, 0.3, int(parts[2][3:]), ref[key]))
    return cases


def test_free_element_matches_reference_fixtures(torch_gpu):
    """n disjoint copies of the fixture element, u = e_j on copy j: the FREE
    matvec returns K_e's columns; the diagonal and the trace come from the
    DIAG kernel."""
    torch = torch_gpu
    for key, kind, xe, E, nu, ngt, K in fixture_cases():
        n = K.shape[0]
        gmap = np.arange(n * n, dtype=np.uint64).reshape(n, n)
        op = aMatFree(one_rank_maps(gmap, n * n))
        op.set_element_function(kind, np.repeat(xe[None], n, axis=0), None, E, nu, ngt)
        op.finalize()
        u = np.zeros(n * n)
        u[np.arange(n) * n + np.arange(n)] = 1.0
        v = run(torch, op.matvec_dev, u).reshape(n, n)   # row j = column j of K
        scale = np.abs(K).max()
        assert np.abs(v.T - K).max() <= 1e-12 * scale, key
        d = op.get_diagonal()
        assert np.abs(d - np.tile(np.diag(K), n)).max() <= 1e-12 * scale, key
        assert abs(op.info().trace - n * np.trace(K)) <= 1e-12 * n * abs(np.trace(K)), key


def distorted(p: problems.Problem, amp: float, seed: int = 11) -> problems.Problem:
    """The problem on a mesh whose interior nodes moved by U[-amp h, amp h]; K_e
    from the element routines per element (times the problem's factors)."""
    m = p.mesh
    c = m.coords.copy()
    lo, hi = c.min(axis=0), c.max(axis=0)
    inner = np.all((c > lo + 1e-12) & (c < hi - 1e-12), axis=1)
    h = (hi - lo).min() / max(m.dims)
    c[inner] += mesh.uniform(seed, 9, 3 * int(inner.sum()), -amp * h, amp * h).reshape(-1, 3)
    m2 = dataclasses.replace(m, coords=c)
    xe = c[m.elem_nodes]
    f = p.free
    scale = p.scale if p.scale is not None else np.ones(p.n_elem)
    if f["kind"] == ELEM.HEX8_LAPLACE:
        K = np.stack([fem.ke_hex8(x, f["ngt"]) for x in xe])
    elif f["kind"] == ELEM.HEX20_ELASTIC:
        K = np.stack([fem.ke_hex20_iso(x, f["E"], f["nu"], f["ngt"]) for x in xe])
    else:
        K = fem.ke_tet10_batch(xe, f["ngt"])
    return dataclasses.replace(p, mesh=m2, K=K * scale[:, None, None], k_ref=None)


CASES = {
    "hex8": lambda: distorted(problems.poisson_hex8(7, 6, 5), 0.2),
    "hex20_ngt2": lambda: distorted(problems.elasticity_hex20(4, 3, 3, ngt=2), 0.1),
    "hex20_ngt3": lambda: distorted(problems.elasticity_hex20(3, 3, 2, ngt=3), 0.1),
    "tet10": lambda: problems.poisson_tet10(4, with_rhs=False),
}


@pytest.mark.parametrize("which", list(CASES))
def test_free_matches_oracle_on_distorted_mesh(torch_gpu, which):
    torch = torch_gpu
    p = CASES[which]()
    K = p.element_matrices()
    u = p.random_u()
    op = p.build_free_operator(aMatFree, Maps, bcop=BCOP.IDENTITY)
    v = run(torch, op.matvec_dev, u)
    assert rel_l2(v, orc.matvec(p.gmap, K, u)) <= 1e-13
    np.testing.assert_array_equal(run(torch, op.matvec_dev, u), v)   # deterministic scatter
    vm = run(torch, op.matmult_dev, u)
    assert rel_l2(vm, orc.matmult_mf(p.gmap, K, u, p.constrained)) <= 1e-13
    # host-array forms and the stored-matrix operator on the same K_e
    vh = np.zeros_like(u)
    op.matvec(vh, u, True)
    np.testing.assert_array_equal(vh, v)
    st = p.build_operator(aMatFree, Maps, bcop=BCOP.IDENTITY)
    assert rel_l2(op.get_diagonal(), st.get_diagonal()) <= 1e-13
    assert abs(op.info().trace - st.info().trace) <= 1e-13 * abs(st.info().trace)
    # atomic scatter
    opa = p.build_free_operator(aMatFree, Maps, bcop=BCOP.IDENTITY, scatter=SCATTER.ATOMIC)
    assert rel_l2(run(torch, opa.matmult_dev, u), orc.matmult_mf(p.gmap, K, u, p.constrained)) <= 1e-13
    inf = op.info()
    assert inf.matfree_type == 2 and inf.k_bytes_fp64 == 0 and inf.flops > 0


def test_free_penalty_and_cg(torch_gpu):
    torch = torch_gpu
    p = distorted(problems.elasticity_hex20(3, 3, 3, ngt=2), 0.1)
    K = p.element_matrices()
    u = p.random_u()
    op = p.build_free_operator(aMatFree, Maps, bc_meth=1, bcop=BCOP.PENALTY_DIAG)
    st = p.build_operator(aMatFree, Maps, bc_meth=1, bcop=BCOP.PENALTY_DIAG)
    assert abs(op.info().penalty - st.info().penalty) <= 1e-13 * st.info().penalty
    maps = orc.set_bdr_map(0, 1, [p.n_dofs], p.gmap, p.constrained, p.prescribed)
    want = orc.matmult_penalty(p.gmap, K, maps["bdr_map"], st.info().penalty, u)
    assert rel_l2(run(torch, op.matmult_dev, u), want) <= 1e-13
    b = mesh.uniform(5, 5, p.n_dofs, -1, 1)
    xs = []
    its = []
    for o in (op, st):
        db = dev(torch, b)
        dx = torch.zeros_like(db)
        s = o.cg(db, dx, rtol=1e-10, max_it=5000, stream=torch.cuda.current_stream())
        assert s.converged, (s.iterations, s.rel_residual)
        xs.append(dx.cpu().numpy())
        its.append(s.iterations)
    assert rel_l2(xs[0], xs[1]) <= 1e-8
    assert abs(its[0] - its[1]) <= max(2, 0.02 * its[1])


def test_free_two_rank_partition(torch_gpu):
    """The independent/dependent launch split and the halo path with FREE
    kernels: both ranks of a z-slab partition on the one GPU equal the
    single-rank MatMult_mf."""
    torch = torch_gpu
    gp = distorted(problems.poisson_hex8(6, 5, 8), 0.2)
    m = gp.mesh
    ranks = mesh.zslab_partition(m, 2)
    maps, ops = [], []
    for r in ranks:
        mp = Maps(r.rank, r.size)
        mp.set_map(len(r.elements), r.element_to_rank_map, None, r.n_all_dofs, r.rank_to_global_map,
                   r.owned_begin, r.owned_end, r.n_global, r.rank_dof_counts)
        maps.append(mp)
    from test_maps import alltoallv
    for mp, (c, g) in zip(maps, alltoallv([mp.get_ghost_requests() for mp in maps])):
        mp.set_peer_requests(c, g)
    for r, mp in zip(ranks, maps):
        cd, cv = mesh.constraints_for(m, r.elements, mesh.boundary_nodes(m))
        mp.set_bdr_map(cd, cv)
        op = aMatFree(mp)
        op.set_element_function(ELEM.HEX8_LAPLACE, m.coords[m.elem_nodes[r.elements]], None, ngt=2)
        op.finalize()
        ops.append(op)
    u = gp.random_u()
    s = torch.cuda.current_stream()
    du, dv = [], []
    for r in ranks:
        owned = np.zeros(r.n_all_dofs)
        owned[r.n_pre:r.n_pre + r.n_owned] = u[r.owned_begin:r.owned_end + 1]
        du.append(dev(torch, owned))
        dv.append(torch.zeros(r.n_all_dofs, dtype=torch.float64, device="cuda"))
    for op, a, b in zip(ops, du, dv):
        op.matvec_begin_dev(b, a, True, s)
    for r, op, a in zip(ranks, ops, du):
        ghosts = np.concatenate([r.rank_to_global_map[:r.n_pre], r.rank_to_global_map[r.n_pre + r.n_owned:]])
        op.halo_import_dev(a, dev(torch, u[ghosts.astype(np.int64)]), s)
    for op, a, b in zip(ops, du, dv):
        op.matvec_end_dev(b, a, True, s)
    sends = []
    for r, op, b in zip(ranks, ops, dv):
        h = torch.zeros(max(r.n_pre + r.n_post, 1), dtype=torch.float64, device="cuda")
        op.halo_export_dev(b, h, s)
        sends.append(h.cpu().numpy()[:r.n_pre + r.n_post])
    for i, (r, op, b, a) in enumerate(zip(ranks, ops, dv, du)):
        recv = []
        for j, mj in enumerate(maps):
            cnt, off = mj.get_RecvDofCounts(), mj.get_RecvDofOffset()
            recv.append(sends[j][off[i]:off[i] + cnt[i]])
        rv = np.concatenate(recv) if recv else np.zeros(0)
        if len(rv):
            op.halo_accumulate_dev(b, dev(torch, rv), s)
        op.bc_fixup_dev(b, a, s)
    full = orc.matmult_mf(gp.gmap, gp.element_matrices(), u, gp.constrained)
    for r, b in zip(ranks, dv):
        got = b.cpu().numpy()[r.n_pre:r.n_pre + r.n_owned]
        assert rel_l2(got, full[r.owned_begin:r.owned_end + 1]) <= 1e-13


def test_free_operators_served(torch_gpu):
    """Method-2 operators in the continuous-batching server (FP64 CG): each
    job's solution is bit-identical to the single solve on the same operator,
    alone and in a mix with a stored-matrix job."""
    torch = torch_gpu
    from amat_amd import serve
    cases = [serve.build_problem("hex8", 6), serve.build_problem("hex20", 3), serve.build_problem("tet10", 3)]
    ops = [serve.build_operator(p, free=True) for p in cases]
    rhs = [serve.rhs_for(p, 40 + i) for i, p in enumerate(cases)]
    single = []
    for p, op, b in zip(cases, ops, rhs):
        db = torch.from_numpy(b).cuda()
        x = torch.zeros_like(db)
        st = op.cg(db, x, rtol=1e-10, max_it=20000)
        assert st.converged
        single.append((x.cpu().numpy(), st.iterations))
    stored_p = serve.build_problem("hex8", 5)
    stored_op = serve.build_operator(stored_p)
    srv = Server(max_jobs=8)
    ids = [srv.submit(op, PREC.FP64, b) for op, b in zip(ops, rhs)]
    sid = srv.submit(stored_op, PREC.FP64, serve.rhs_for(stored_p, 7))
    sizes = {j: p.n_dofs for j, p in zip(ids, cases)}
    sizes[sid] = stored_p.n_dofs
    out = {}
    while len(out) < len(sizes):
        srv.step(8)
        for j in srv.poll():
            out[j] = srv.result(j, sizes[j])
    for j, (x1, it1) in zip(ids, single):
        x, st = out[j]
        assert st.converged and st.rel_residual <= 1e-10
        np.testing.assert_array_equal(x, x1)
        assert st.iterations == it1
    with pytest.raises(AmatError):
        srv.submit(ops[0], PREC.BF16, rhs[0])   # FREE has no BF16 form


@pytest.mark.parametrize("which", list(CASES))
def test_free_fp32_form(torch_gpu, which):
    """The FP32 form of the FREE kernels (FP32 arithmetic and vectors, the
    inner operator of the mixed-precision solvers) against the FP64 form:
    within FP32 rounding of the element algebra."""
    torch = torch_gpu
    p = CASES[which]()
    op = p.build_free_operator(aMatFree, Maps, bcop=BCOP.IDENTITY)
    u = p.random_u()
    v64 = run(torch, op.matmult_dev, u)
    du = torch.from_numpy(u).float().cuda()
    dv = torch.zeros_like(du)
    op.matmult_f32_dev(PREC.FP32, dv, du)
    torch.cuda.synchronize()
    err = rel_l2(dv.cpu().numpy().astype(np.float64), v64)
    assert 1e-9 < err <= 2e-5, err   # genuinely FP32, and as accurate as FP32 storage


def test_free_fp32_reliable_updates(torch_gpu):
    """FP32 FREE inner solves with FP64 reliable updates reach 1e-10 on hex20
    elasticity, single solve and served (bit-identical alone and in a mix)."""
    torch = torch_gpu
    from amat_amd import serve
    from amat_amd.amat import IR_RELIABLE
    p = serve.build_problem("hex20", 4)
    op = serve.build_operator(p, free=True)
    b = serve.rhs_for(p, 3)
    db = torch.from_numpy(b).cuda()
    x64 = torch.zeros_like(db)
    st64 = op.cg(db, x64, rtol=1e-10, max_it=20000)
    method = int(PREC.FP32) | IR_RELIABLE
    x = torch.zeros_like(db)
    st = op.ir_cg(method, db, x, rtol=1e-10, inner_rtol=0.1, max_outer=200, max_inner=20000)
    assert st.converged and st.rel_residual <= 1e-10
    assert st.iterations <= 1.5 * st64.iterations
    assert rel_l2(x.cpu().numpy(), x64.cpu().numpy()) <= 1e-8
    srv = Server(max_jobs=4)
    j1 = srv.submit(op, method, b, inner_rtol=0.1, max_outer=200)
    other = serve.build_problem("tet10", 3)
    j2 = srv.submit(serve.build_operator(other, free=True), PREC.FP64, serve.rhs_for(other, 5))
    out = {}
    while len(out) < 2:
        srv.step(8)
        for j in srv.poll():
            out[j] = srv.result(j, p.n_dofs if j == j1 else other.n_dofs)
    xs, sts = out[j1]
    assert sts.converged and sts.iterations == st.iterations
    np.testing.assert_array_equal(xs, x.cpu().numpy())


def test_free_rejections(torch_gpu):
    torch = torch_gpu
    p = problems.poisson_hex8(3)
    op = p.build_free_operator(aMatFree, Maps)
    x = torch.zeros(p.n_dofs, dtype=torch.float32, device="cuda")
    with pytest.raises(AmatError):
        op.matmult_f32_dev(PREC.BF16, x, x)      # no BF16 form
    with pytest.raises(AmatError):
        MultiRHS([op], PREC.FP32)
    with pytest.raises(AmatError):
        Server(max_jobs=2).submit(op, PREC.BF16, np.zeros(p.n_dofs))
    # wrong element kind for the DoF count; FREE finalize without coordinates
    o2 = aMatFree(p.build_maps(Maps))
    with pytest.raises(AmatError):
        o2.set_element_function(ELEM.TET10_LAPLACE, np.zeros((p.n_elem, 10, 3)), None)
    o2.set_matfree_type(2)
    with pytest.raises(AmatError):
        o2.finalize()
    with pytest.raises(AmatError):
        o2.set_element_function(ELEM.HEX8_LAPLACE, p.element_coords(), None, ngt=0)
    # an inverted (mirrored) element: non-positive Jacobians, rejected at finalize
    # as the element routines reject them
    for kind_p, kind, perm in ((problems.poisson_hex8(3), ELEM.HEX8_LAPLACE, [1, 0, 3, 2, 5, 4, 7, 6]),
                               (problems.elasticity_hex20(2, ngt=2), ELEM.HEX20_ELASTIC,
                                [1, 0, 3, 2, 5, 4, 7, 6, 8, 11, 10, 9, 12, 15, 14, 13, 17, 16, 19, 18])):
        xe = kind_p.element_coords().copy()
        xe[1] = xe[1][perm]
        o3 = aMatFree(kind_p.build_maps(Maps))
        o3.set_element_function(kind, xe, None, ngt=2)
        with pytest.raises(AmatError) as e:
            o3.finalize()
        assert e.value.name == "INVALID_ARGUMENT" and "element 1" in str(e.value)


def test_free_jobs_on_side_streams_bit_identical(torch_gpu):
    """More method-2 jobs than the server's side streams (FP64 and FP32
    reliable-update jobs of three kinds, different sizes, joining late): every
    served solution equals its single solve bit for bit."""
    torch = torch_gpu
    from amat_amd import serve
    from amat_amd.amat import IR_RELIABLE
    specs = [("hex8", 5, PREC.FP64), ("hex8", 7, int(PREC.FP32) | IR_RELIABLE), ("hex20", 2, PREC.FP64),
             ("hex20", 3, int(PREC.FP32) | IR_RELIABLE), ("tet10", 2, PREC.FP64), ("tet10", 3, int(PREC.FP32) | IR_RELIABLE),
             ("hex8", 4, int(PREC.FP32) | IR_RELIABLE)]
    cases = []
    for i, (kind, n, method) in enumerate(specs):
        p = serve.build_problem(kind, n)
        op = serve.build_operator(p, free=True)
        b = serve.rhs_for(p, 60 + i)
        db = torch.from_numpy(b).cuda()
        x = torch.zeros_like(db)
        if method == PREC.FP64:
            st = op.cg(db, x, rtol=1e-10, max_it=20000)
        else:
            st = op.ir_cg(method, db, x, rtol=1e-10, inner_rtol=0.1, max_outer=60, max_inner=20000)
        assert st.converged
        cases.append((p, op, b, method, x.cpu().numpy(), st.iterations))
    srv = Server(max_jobs=16)
    ids = {}
    for k, (p, op, b, method, _, _) in enumerate(cases):
        ids[srv.submit(op, method, b, inner_rtol=0.1)] = k
        if k == 3:
            srv.step(8)   # the later jobs join while the first ones iterate
    out = {}
    while len(out) < len(cases):
        srv.step(8)
        for j in srv.poll():
            out[j] = srv.result(j, cases[ids[j]][0].n_dofs)
    for j, k in ids.items():
        x, st = out[j]
        assert st.converged and st.iterations == cases[k][5]
        np.testing.assert_array_equal(x, cases[k][4])
