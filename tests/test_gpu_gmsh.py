"""Gmsh-partitioned meshes through the HIP operator: the set-up of
examples/src/example05a.cpp:135-163 (GmshMesh -> Maps -> aMatFree), every
rank's share on the one GPU with its halo moved through device buffers.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from amat_amd import BCOP, ELEM, Maps, aMatFree, fem, gmsh, mesh
from conftest import rel_l2
from test_gpu_matvec import dev, run
from test_maps import alltoallv

pytestmark = pytest.mark.gpu


def _ranks(path, P, dpn):
    ms = [gmsh.GmshMesh(path, r, P, dpn) for r in range(P)]
    maps = []
    for g in ms:
        m = Maps(g.rank, g.size)
        m.set_map(*g.set_map_args())
        maps.append(m)
    if P > 1:
        for m, (c, gids) in zip(maps, alltoallv([m.get_ghost_requests() for m in maps])):
            m.set_peer_requests(c, gids)
    for g, m in zip(ms, maps):
        m.set_bdr_map(g.get_constrained_dofs(), np.zeros(g.get_n_constraints()))
    return ms, maps


def test_gmsh_tet10_partitions_with_halos(torch_gpu, tmp_path):
    """A tet10 Laplace mesh in 3 Gmsh partitions: each rank's owned part of
    MatMult_mf equals the one-rank product on the same global numbering."""
    torch = torch_gpu
    m = mesh.tet10(4, seed=6)
    P = 3
    parts = (np.arange(m.n_elem) * P) // m.n_elem
    path = tmp_path / "t10.msh"
    gmsh.write_msh(path, m, parts=parts)
    ms, maps = _ranks(path, P, 1)
    ops, K_all, g_all, cons = [], [], [], []
    for g, mp in zip(ms, maps):
        K = fem.ke_tet10_batch(g.element_coordinates(), 2)
        op = aMatFree(mp)
        op.set_matfree_type(4)
        op.set_element_matrices(0, K)
        op.finalize()
        ops.append(op)
        K_all.append(K)
        g_all.append(np.asarray(g.get_globalDofMap()))
        cons.append(g.get_constrained_dofs())
    N = ms[0].get_nDofsTotal()
    gmap, K = np.concatenate(g_all), np.concatenate(K_all)
    cd = np.unique(np.concatenate(cons))
    np.testing.assert_array_equal(cd.size, mesh.boundary_nodes(m).size)
    u = np.random.default_rng(1).uniform(-1, 1, N)
    full = orc.matmult_mf(gmap, K, u, cd)
    s = torch.cuda.current_stream()
    du, dv = [], []
    for g in ms:
        l2g = g.get_local2GlobalDofMap().astype(np.int64)
        inf = g.info
        owned = np.zeros(g.get_nLocalDofs())
        sl = slice(inf.n_pre_ghost_nodes, inf.n_pre_ghost_nodes + inf.n_owned_nodes)
        owned[sl] = u[l2g[sl]]
        du.append(dev(torch, owned))
        dv.append(torch.zeros(g.get_nLocalDofs(), dtype=torch.float64, device="cuda"))
    for op, a, b in zip(ops, du, dv):
        op.matvec_begin_dev(b, a, True, s)
    for g, op, a in zip(ms, ops, du):
        inf = g.info
        l2g = g.get_local2GlobalDofMap().astype(np.int64)
        ghosts = np.concatenate([l2g[:inf.n_pre_ghost_nodes], l2g[inf.n_pre_ghost_nodes + inf.n_owned_nodes:]])
        if len(ghosts):
            op.halo_import_dev(a, dev(torch, u[ghosts]), s)
    for op, a, b in zip(ops, du, dv):
        op.matvec_end_dev(b, a, True, s)
    sends = []
    for g, op, b in zip(ms, ops, dv):
        inf = g.info
        ng = inf.n_pre_ghost_nodes + inf.n_post_ghost_nodes
        h = torch.zeros(max(ng, 1), dtype=torch.float64, device="cuda")
        op.halo_export_dev(b, h, s)
        sends.append(h.cpu().numpy()[:ng])
    for i, (g, op, b, a) in enumerate(zip(ms, ops, dv, du)):
        recv = []
        for j, mj in enumerate(maps):
            cnt, off = mj.get_RecvDofCounts(), mj.get_RecvDofOffset()
            recv.append(sends[j][off[i]:off[i] + cnt[i]])
        rv = np.concatenate(recv)
        if len(rv):
            op.halo_accumulate_dev(b, dev(torch, rv), s)
        op.bc_fixup_dev(b, a, s)
    for g, b in zip(ms, dv):
        inf = g.info
        lo = g.get_startGlobalDof()
        got = b.cpu().numpy()[inf.n_pre_ghost_nodes:inf.n_pre_ghost_nodes + inf.n_owned_nodes]
        assert rel_l2(got, full[lo:lo + inf.n_owned_nodes]) <= 1e-14


def test_gmsh_one_rank_stored_and_free(torch_gpu, tmp_path):
    """One rank of a Gmsh tet10 mesh: the stored-matrix product is bit-exact
    against the oracle, and method 2 (FREE) from the mesh's coordinates agrees
    with it; a Gmsh hex20 elasticity mesh likewise."""
    torch = torch_gpu
    m = mesh.tet10(3, seed=8)
    path = tmp_path / "one.msh"
    gmsh.write_msh(path, m)
    (g,), (mp,) = _ranks(path, 1, 1)
    xe = g.element_coordinates()
    K = fem.ke_tet10_batch(xe, 2)
    op = aMatFree(mp)
    op.set_matfree_type(4)
    op.set_element_matrices(0, K)
    op.finalize()
    u = np.random.default_rng(2).uniform(-1, 1, g.get_nDofsTotal())
    gmap = np.asarray(g.get_globalDofMap())
    v = run(torch, op.matvec_dev, u)
    np.testing.assert_array_equal(v, orc.matvec(gmap, K, u))
    fr = aMatFree(mp)
    fr.set_matfree_type(2)
    fr.set_element_function(ELEM.TET10_LAPLACE, xe, None, ngt=2)
    fr.finalize()
    assert rel_l2(run(torch, fr.matvec_dev, u), v) <= 1e-13
    # hex20, 3 DoFs per node (Gmsh type 17)
    h = mesh.hex20(2, 3, 2)
    hp = tmp_path / "h20.msh"
    gmsh.write_msh(hp, h)
    (gh,), (mh,) = _ranks(hp, 1, 3)
    xh = gh.element_coordinates()
    Kh = np.stack([fem.ke_hex20_iso(x, 1.0, 0.3, 2) for x in xh])
    oh = aMatFree(mh)
    oh.set_matfree_type(4)
    oh.set_element_matrices(0, Kh)
    oh.finalize()
    uh = np.random.default_rng(3).uniform(-1, 1, gh.get_nDofsTotal())
    vh = run(torch, oh.matvec_dev, uh)
    np.testing.assert_array_equal(vh, orc.matvec(np.asarray(gh.get_globalDofMap()), Kh, uh))
    fh = aMatFree(mh)
    fh.set_matfree_type(2)
    fh.set_element_function(ELEM.HEX20_ELASTIC, xh, None, E=1.0, nu=0.3, ngt=2)
    fh.finalize()
    assert rel_l2(run(torch, fh.matvec_dev, uh), vh) <= 1e-13
