"""GPU parity of the EbE operator (HIP kernels through the C ABI) against the
oracle on the same seeded inputs.

Bar: the deterministic FP64 path is bit-exact against the oracle's
sequential matvec_ghosted_OMP / MatMult_mf (same products, same summation
order, no FMA contraction); the atomic-scatter and penalty paths agree to
1e-13 relative L2 (the stated FP64 tolerance); FP32/BF16 storage errors are
bounded and reported.
"""
import numpy as np
import pytest

import oracle as orc
from amat_amd import BCOP, MATRIX_TYPE, PREC, SCATTER, Maps, aMatFree, mesh, problems
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(torch, fn, u, dtype=np.float64):
    """v = fn(v, u, stream) on torch's current stream; returns numpy."""
    du = dev(torch, u.astype(dtype))
    dv = torch.empty_like(du)
    fn(dv, du, torch.cuda.current_stream())
    return dv.cpu().numpy()


def build(p, bcop=BCOP.IDENTITY, precision=PREC.FP64, scatter=SCATTER.DETERMINISTIC, penalty=None, bc_meth=0):
    return p.build_operator(aMatFree, Maps, bc_meth=bc_meth, precision=precision, scatter=scatter, bcop=bcop,
                            penalty=penalty)


def test_poisson_hex8_32_bitexact(torch_gpu):
    """Full size: 32^3 hex8 Poisson, identity BCs on the boundary."""
    p = problems.poisson_hex8(32)
    op = build(p)
    u = p.random_u()
    K = p.element_matrices()
    v = run(torch_gpu, op.matvec_dev, u)
    np.testing.assert_array_equal(v, orc.matvec(p.gmap, K, u))
    vm = run(torch_gpu, op.matmult_dev, u)
    np.testing.assert_array_equal(vm, orc.matmult_mf(p.gmap, K, u, p.constrained))
    # host-array forms (aMatFree::matvec / MatMult_mf signatures)
    vh = np.zeros_like(u)
    op.matvec(vh, u, True)
    np.testing.assert_array_equal(vh, v)
    np.testing.assert_array_equal(op.MatMult_mf(u), vm)
    out = np.full_like(u, np.nan)
    assert op.MatMult_mf(u, out=out) is out
    np.testing.assert_array_equal(out, vm)
    with pytest.raises(ValueError):
        op.MatMult_mf(u, out=np.zeros(u.size + 1))
    # aMatFree::get_timer: the host-array calls above (matvec, 2 x MatMult_mf) accumulated
    sc, ga, mv, tot = op.get_timer()
    assert sc > 0 and ga > 0 and mv > 0 and tot >= sc + mv + ga > 0
    op.MatMult_mf(u, out=out)
    sc2, ga2, mv2, tot2 = op.get_timer()
    assert sc2 > sc and ga2 > ga and mv2 > mv and tot2 > tot
    # device forms are not counted
    run(torch_gpu, op.matmult_dev, u)
    assert op.get_timer() == (sc2, ga2, mv2, tot2)


def test_dump_mat_matches_assembled_matrix(torch_gpu, tmp_path):
    """aMatFree::dump_mat: probing matvec with unit vectors gives the matrix
    assembled from the element matrices (scipy COO, duplicates summed), with
    boundary rows untouched (matvec applies no BCs)."""
    import scipy.sparse as sp
    p = problems.poisson_tet10(2, with_rhs=False)
    op = build(p)
    op.finalize_end()
    assert op.get_matrix_type() == MATRIX_TYPE.MATRIX_FREE
    K = p.element_matrices()
    trip = op.dump_mat(tmp_path / "m.out")
    n = p.n_dofs
    gm = p.gmap.astype(np.int64)
    E, ne = gm.shape
    R = np.broadcast_to(gm[:, :, None], (E, ne, ne)).ravel()
    Cc = np.broadcast_to(gm[:, None, :], (E, ne, ne)).ravel()
    A = sp.coo_matrix((K.reshape(-1), (R, Cc)), shape=(n, n)).tocsr()
    D = sp.coo_matrix((trip.value, (trip.row, trip.col)), shape=(n, n)).tocsr()
    assert abs(A - D).max() <= 1e-14 * abs(A).max()
    assert (D != D.T).nnz == 0          # symmetric K_e, same element order: symmetric bit for bit
    assert np.all(np.diff(trip.row) >= 0)
    lines = (tmp_path / "m.out").read_text().splitlines()
    assert len(lines) == len(trip)
    r0, c0, v0 = lines[0].split()
    assert (int(r0), int(c0), float(v0)) == (int(trip.row[0]), int(trip.col[0]), float(trip.value[0]))


def test_hex20_elasticity_bitexact_and_penalty(torch_gpu):
    p = problems.elasticity_hex20(6)
    op = build(p, bcop=BCOP.PENALTY_DIAG, penalty=1e8)
    K = p.element_matrices()
    u = p.random_u()
    np.testing.assert_array_equal(run(torch_gpu, op.matvec_dev, u), orc.matvec(p.gmap, K, u))
    v = run(torch_gpu, op.matmult_dev, u)
    maps = orc.set_bdr_map(0, 1, [p.n_dofs], p.gmap, p.constrained, p.prescribed)
    ref = orc.matmult_penalty(p.gmap, K, maps["bdr_map"], 1e8, u)
    free = np.setdiff1d(np.arange(p.n_dofs), p.constrained.astype(np.int64))
    assert rel_l2(v, ref) <= 1e-13
    assert rel_l2(v[free], ref[free]) <= 1e-13


def test_tet10_bitexact(torch_gpu):
    p = problems.poisson_tet10(4, with_rhs=False)
    op = build(p)
    K = p.element_matrices()
    u = p.random_u()
    np.testing.assert_array_equal(run(torch_gpu, op.matvec_dev, u), orc.matvec(p.gmap, K, u))
    np.testing.assert_array_equal(run(torch_gpu, op.matmult_dev, u), orc.matmult_mf(p.gmap, K, u, p.constrained))


@pytest.mark.parametrize("which", ["hex8", "hex20", "tet10"])
def test_atomic_scatter_and_determinism(torch_gpu, which):
    p = {"hex8": lambda: problems.poisson_hex8(10), "hex20": lambda: problems.elasticity_hex20(4),
         "tet10": lambda: problems.poisson_tet10(3, with_rhs=False)}[which]()
    det = build(p)
    ato = build(p, scatter=SCATTER.ATOMIC)
    u = p.random_u()
    v1 = run(torch_gpu, det.matmult_dev, u)
    v2 = run(torch_gpu, det.matmult_dev, u)
    np.testing.assert_array_equal(v1, v2)   # deterministic order: bitwise repeatable
    va = run(torch_gpu, ato.matmult_dev, u)
    assert rel_l2(va, v1) <= 1e-13


# about twice the measured errors (FP32 8.7e-8, BF16 9.3e-4)
@pytest.mark.parametrize("prec,bound", [(PREC.FP32, 2e-7), (PREC.BF16, 2e-3)])
def test_low_precision_storage_error(torch_gpu, prec, bound):
    p = problems.elasticity_hex20(4)
    op = build(p, precision=prec)
    u = p.random_u()
    ref = run(torch_gpu, op.matmult_dev, u)
    v32 = run(torch_gpu, lambda v, uu, s: op.matmult_f32_dev(prec, v, uu, s), u, np.float32)
    err = rel_l2(v32, ref)
    print(f"storage {prec.name}: rel L2 error vs FP64 = {err:.3e}")
    assert err <= bound


def flat_problem(seed=0):
    """Elements of three sizes sharing DoFs: n = 8 and 60 (tile kernels) and
    n = 36 (the generic kernel)."""
    rng = np.random.default_rng(seed)
    N = 400
    sizes = [8] * 50 + [60] * 12 + [36] * 9
    rng.shuffle(sizes)
    maps, mats = [], []
    for n in sizes:
        maps.append(rng.choice(N, n, replace=False).astype(np.uint32))
        A = rng.normal(size=(n, n))
        mats.append(A)
    return N, sizes, maps, mats


def test_mixed_element_sizes_and_generic_kernel(torch_gpu):
    N, sizes, maps_l, mats = flat_problem()
    m = Maps()
    m.set_map(len(sizes), maps_l, np.array(sizes, np.uint32), N, np.arange(N), 0, N - 1, N)
    op = aMatFree(m)
    op.set_matfree_type(3)
    for e, A in enumerate(mats):
        op.set_element_matrix(e, A)
    op.set_bc_operator(BCOP.NONE)
    op.finalize()
    assert op.info().n_groups == 3
    u = mesh.uniform(1, 2, N, -1, 1)
    v = run(torch_gpu, op.matvec_dev, u)
    ref = np.zeros(N)
    for mp, A in zip(maps_l, mats):   # sequential element order, separate multiply and add
        ue = u[mp]
        ve = np.zeros(len(mp))
        for c in range(len(mp)):
            ve = ve + A[:, c] * ue[c]
        np.add.at(ref, mp, 0)
        for r in range(len(mp)):
            ref[mp[r]] += ve[r]
    assert rel_l2(v, ref) <= 1e-14
    K_flat = np.concatenate([A.reshape(-1) for A in mats])
    import ctypes as C
    dpe = np.array(sizes, np.uint32)
    lm = np.concatenate(maps_l).astype(np.uint32)
    vo = np.zeros(N)
    order = np.arange(len(sizes), dtype=np.uint32)
    orc.lib().orc_matvec_ghosted_omp(C.c_uint32(len(sizes)), orc._p(dpe), orc._p(lm), orc._p(K_flat), C.c_int(0),
                                     orc._p(order), C.c_uint32(len(sizes)), None, C.c_uint32(0), orc._p(u),
                                     orc._p(vo), C.c_uint32(N), C.c_int(1))
    np.testing.assert_array_equal(v, vo)


def test_long_lists_span_gather_windows(torch_gpu):
    """The gather reads a block's list range in windows of 2,048 entries: DoFs
    shared by hundreds to thousands of elements make a block's range several
    windows long and single lists cross window boundaries; bit-exact against
    the oracle's sequential matvec_ghosted_OMP (lists summed in element order)."""
    import ctypes as C
    rng = np.random.default_rng(11)
    N, E, n = 1200, 4000, 8
    hubs = [0, 3, 255, 256, 700]                 # within, at the end of and after the first block
    maps_l = []
    for e in range(E):
        h = 0 if e % 2 == 0 else hubs[(e // 2) % len(hubs)]
        rest = rng.choice(np.setdiff1d(np.arange(N), [h]), n - 1, replace=False)
        mp = np.concatenate([[h], rest]).astype(np.uint32)
        rng.shuffle(mp)
        maps_l.append(mp)
    mats = [rng.normal(size=(n, n)) for _ in range(E)]
    counts = np.bincount(np.concatenate(maps_l), minlength=N)
    assert counts[:256].sum() > 3 * 2048 and counts[0] > 2048
    m = Maps()
    m.set_map(E, maps_l, np.full(E, n, np.uint32), N, np.arange(N), 0, N - 1, N)
    op = aMatFree(m)
    op.set_matfree_type(3)
    for e, A in enumerate(mats):
        op.set_element_matrix(e, A)
    op.set_bc_operator(BCOP.NONE)
    op.finalize()
    u = mesh.uniform(2, 2, N, -1, 1)
    v = run(torch_gpu, op.matvec_dev, u)
    K_flat = np.concatenate([A.reshape(-1) for A in mats])
    lm = np.concatenate(maps_l).astype(np.uint32)
    dpe = np.full(E, n, np.uint32)
    vo = np.zeros(N)
    order = np.arange(E, dtype=np.uint32)
    orc.lib().orc_matvec_ghosted_omp(C.c_uint32(E), orc._p(dpe), orc._p(lm), orc._p(K_flat), C.c_int(0),
                                     orc._p(order), C.c_uint32(E), None, C.c_uint32(0), orc._p(u),
                                     orc._p(vo), C.c_uint32(N), C.c_int(1))
    np.testing.assert_array_equal(v, vo)


def test_operator_without_elements(torch_gpu):
    """A rank that owns DoFs but no elements: K = 0, so matvec gives 0 and
    MatMult_mf only its identity rows v_c = u_c (include/aMatFree.hpp:3916-3933)."""
    m = Maps()
    m.set_map(0, np.zeros((0, 8), np.uint32), np.zeros(0, np.uint32), 4, np.arange(4), 0, 3, 4)
    m.set_bdr_map([1], [0.0])
    op = aMatFree(m)
    op.set_matfree_type(3)
    op.finalize()
    assert op.info().n_elems == 0
    u = np.array([1.0, 2.0, 3.0, 4.0])
    np.testing.assert_array_equal(run(torch_gpu, op.matvec_dev, u), np.zeros(4))
    np.testing.assert_array_equal(run(torch_gpu, op.matmult_dev, u), [0.0, 2.0, 0.0, 0.0])
    np.testing.assert_array_equal(op.MatMult_mf(u), [0.0, 2.0, 0.0, 0.0])
    # apply_bc_diagonal walks the elements' BC masks (include/aMatFree.hpp:4309-4337):
    # a constrained DoF that no element touches keeps its assembled diagonal, 0
    np.testing.assert_array_equal(op.get_diagonal(), np.zeros(4))


def test_block_element_matrices(torch_gpu):
    rng = np.random.default_rng(4)
    E, bd, nb, N = 40, 2, 4, 90
    n = bd * nb
    lmap = np.stack([rng.choice(N, n, replace=False) for _ in range(E)]).astype(np.uint32)
    present = rng.random((E, bd * bd)) < 0.75
    m = Maps()
    m.set_map(E, lmap, None, N, np.arange(N), 0, N - 1, N)
    op = aMatFree(m)
    flat = []
    for e in range(E):
        for bi in range(bd):
            for bj in range(bd):
                if present[e, bi * bd + bj]:
                    B = rng.normal(size=(nb, nb))
                    flat.append(B.reshape(-1))
                    op.set_element_matrix(e, B, bi, bj, bd)
    op.set_bc_operator(BCOP.NONE)
    op.finalize()
    u = rng.normal(size=N)
    v = run(torch_gpu, op.matvec_dev, u)
    ref = orc.matvec_blocks(np.full(E, n), np.full(E, bd), present.reshape(-1), lmap.reshape(-1),
                            np.concatenate(flat), u, N)
    assert rel_l2(v, ref) <= 1e-14


def test_diagonal_trace_and_rhs_bc(torch_gpu):
    p = problems.poisson_tet10(3, with_rhs=True)
    K = p.element_matrices()
    maps = orc.set_bdr_map(0, 1, [p.n_dofs], p.gmap, p.constrained, p.prescribed)
    for bc_meth in (0, 1):
        op = build(p, bc_meth=bc_meth)
        assert abs(op.info().trace - orc.trace(K)) == 0.0
        d = op.get_diagonal()
        np.testing.assert_array_equal(d, orc.get_diagonal(p.gmap, K, maps["bdr_map"], bc_meth, p.n_dofs))
    # nonzero prescribed values: rhs_f - K_fc U_c, rhs_c = U_c
    p.prescribed = mesh.uniform(4, 4, len(p.constrained), -1, 1)
    op = build(p)
    r = op.apply_bc(p.rhs)
    uc = np.zeros(p.n_dofs)
    c = p.constrained.astype(np.int64)
    uc[c] = p.prescribed
    w = orc.matvec(p.gmap, K, uc)
    exp = p.rhs - w
    exp[c] = p.prescribed
    assert rel_l2(r, exp) <= 1e-14


def test_two_rank_partition_with_halo_on_one_gpu(torch_gpu):
    """One rank's share with ghost layers: both ranks of a z-slab partition run
    on the one GPU, halos move through device buffers, and each rank's owned
    result equals the single-rank MatMult_mf."""
    torch = torch_gpu
    m = mesh.hex8(6, 5, 8)
    gp = problems.poisson_hex8(6, 5, 8)
    ranks = mesh.zslab_partition(m, 2)
    K_ref = gp.k_ref
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
        op.set_matfree_type(4)
        op.set_element_matrices_scaled(K_ref, np.ones(len(r.elements)))
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
    for op, a, b in zip(ops, du, dv):   # independent elements while halos move
        op.matvec_begin_dev(b, a, True, s)
    for r, op, a in zip(ranks, ops, du):   # owner -> ghost: values of u at the ghost ids
        ghosts = np.concatenate([r.rank_to_global_map[:r.n_pre], r.rank_to_global_map[r.n_pre + r.n_owned:]])
        op.halo_import_dev(a, dev(torch, u[ghosts.astype(np.int64)]), s)
    for op, a, b in zip(ops, du, dv):
        op.matvec_end_dev(b, a, True, s)
    # ghost -> owner accumulation through the send lists
    sends = []
    for r, op, b in zip(ranks, ops, dv):
        h = torch.zeros(max(r.n_pre + r.n_post, 1), dtype=torch.float64, device="cuda")
        op.halo_export_dev(b, h, s)
        sends.append(h.cpu().numpy()[:r.n_pre + r.n_post])
    for i, (r, mp, op, b, a) in enumerate(zip(ranks, maps, ops, dv, du)):
        recv = []
        for j, (rj, mj) in enumerate(zip(ranks, maps)):
            cnt, off = mj.get_RecvDofCounts(), mj.get_RecvDofOffset()
            recv.append(sends[j][off[i]:off[i] + cnt[i]])
        rv = np.concatenate(recv) if recv else np.zeros(0)
        if len(rv):
            op.halo_accumulate_dev(b, dev(torch, rv), s)
        op.bc_fixup_dev(b, a, s)
    full = orc.matmult_mf(gp.gmap, gp.element_matrices(), u, gp.constrained)
    for r, b in zip(ranks, dv):
        got = b.cpu().numpy()[r.n_pre:r.n_pre + r.n_owned]
        assert rel_l2(got, full[r.owned_begin:r.owned_end + 1]) <= 1e-14


def test_elasticity_hex20_64_bitexact(torch_gpu):
    """Full size: 64^3 hex20 elasticity (262,144 elements,
    3,257,475 DoFs), K_e = E_e K_ref; the plain product is bit-exact against
    the oracle, the penalty operator's free rows equal it bit for bit and its
    constrained rows agree through the diagonal identity, and symmetry holds."""
    p = problems.elasticity_hex20(64)
    assert p.n_dofs == 3_257_475 and p.n_elem == 262_144
    op = build(p, bcop=BCOP.PENALTY_DIAG, penalty=1e8)
    u = p.random_u()
    v = run(torch_gpu, op.matvec_dev, u)
    K = p.element_matrices()               # 7.55 GB on the host
    vo = orc.matvec(p.gmap, K, u)
    np.testing.assert_array_equal(v, vo)
    vp = run(torch_gpu, op.matmult_dev, u)
    d = orc.get_diagonal(p.gmap, K, None, 0, p.n_dofs)
    del K
    c = p.constrained.astype(np.int64)
    free = np.ones(p.n_dofs, dtype=bool)
    free[c] = False
    np.testing.assert_array_equal(vp[free], vo[free])   # free rows: the plain product, bit for bit
    ref = vo.copy()
    ref[c] += (1e8 - d[c]) * u[c]
    assert rel_l2(vp[c], ref[c]) <= 1e-13
    w = mesh.uniform(77, 0, p.n_dofs, -1, 1)
    kw = run(torch_gpu, op.matvec_dev, w)
    assert abs(np.dot(w, v) - np.dot(u, kw)) <= 1e-12 * np.linalg.norm(w) * np.linalg.norm(v)


def test_measured_read_bandwidth(torch_gpu):
    """The measured HBM read stream used beside the nominal 8 TB/s peak: a
    plausible figure for one MI355X, and argument checks."""
    from amat_amd.amat import measure_read_bandwidth
    from amat_amd import AmatError
    bw = measure_read_bandwidth(torch_gpu.cuda.current_device(), 1 << 30, 3)
    assert 1e12 < bw < 1e13
    with pytest.raises(AmatError):
        measure_read_bandwidth(0, 1024, 1)


@pytest.mark.parametrize("which", ["hex8", "tet10"])
def test_packed_symmetric_storage(torch_gpu, which):
    """8- and 10-DoF element matrices that are bitwise symmetric are stored as
    packed upper triangles (about half the bytes); the kernels read the same
    values in the same order, so every product is bit-identical to the whole
    storage (FP64, FP32, BF16; deterministic scatter) and the FP64 one to the
    oracle.  One asymmetric entry makes the group fall back to whole matrices."""
    p = {"hex8": lambda: problems.poisson_hex8(7, 6, 5), "tet10": lambda: problems.poisson_tet10(3)}[which]()
    K = p.element_matrices().copy()

    def make(Kset, mode, scatter=SCATTER.DETERMINISTIC):
        op = aMatFree(p.build_maps(Maps))
        op.set_matfree_type(3)
        op.set_element_matrices(0, Kset)
        op.set_options(PREC.FP64 | PREC.FP32 | PREC.BF16, scatter)
        op.set_symmetric_storage(mode)
        op.finalize()
        return op

    packed, whole = make(K, 1), make(K, 0)
    ip, iw = packed.info(), whole.info()
    assert ip.packed_symmetric == 1 and iw.packed_symmetric == 0
    n = p.dofs_per_elem
    assert ip.k_bytes_fp64 * n * n == iw.k_bytes_fp64 * n * (n + 1) // 2
    u = p.random_u()
    for fn in ("matvec_dev", "matmult_dev"):
        a = run(torch_gpu, getattr(packed, fn), u)
        np.testing.assert_array_equal(a, run(torch_gpu, getattr(whole, fn), u))
    np.testing.assert_array_equal(run(torch_gpu, packed.matvec_dev, u), orc.matvec(p.gmap, K, u))
    for prec in (PREC.FP32, PREC.BF16):
        a = run(torch_gpu, lambda v, uu, s: packed.matmult_f32_dev(prec, v, uu, s), u, np.float32)
        b = run(torch_gpu, lambda v, uu, s: whole.matmult_f32_dev(prec, v, uu, s), u, np.float32)
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(packed.get_diagonal(), whole.get_diagonal())
    ato = make(K, 1, SCATTER.ATOMIC)
    assert rel_l2(run(torch_gpu, ato.matmult_dev, u), run(torch_gpu, whole.matmult_dev, u)) <= 1e-13
    # one entry off symmetry: whole matrices, still the oracle's product
    K2 = K.copy()
    K2[1, 0, n - 1] = np.nextafter(K2[1, 0, n - 1], 1.0)
    fb = make(K2, 1)
    assert fb.info().packed_symmetric == 0
    np.testing.assert_array_equal(run(torch_gpu, fb.matvec_dev, u), orc.matvec(p.gmap, K2, u))
