"""A rank's share with the halo exchange inside the operator's own calls.

The reference's matvec, MatMult_mf, MatGetDiagonal_mf and apply_bc_rhs on
a rank's share exchange ghost values through MPI point-to-point messages
(include/aMatFree.hpp:1426-1679, 1722-1768, 3893-3941, 1051-1053,
4640-4646).  Here the caller registers the transport
(amx_op_set_halo_transport); the tests run every rank of a z-slab partition
in a thread of this process on the one GPU, with a transport that moves the
buffers between the threads exactly as the reference's Isend/Irecv pairs do
(peer i's part at the send/recv offsets of the maps), in both buffer modes.

Bar: each rank's owned rows equal the one-rank oracle's rows to 1e-14
relative L2 (the summation order differs: a rank adds its ghosts'
contributions after its own elements), 1e-13 for the atomic scatter and the
penalty operator; the diagonal with identity BC rows to 1e-14.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
import threading

import numpy as np
import pytest

import oracle as orc
from amat_amd import (BCOP, GHOST_TO_OWNER, HALO, OWNER_TO_GHOST, SCATTER, AmatError, Maps, aMatFree, mesh,
                      problems)
from conftest import rel_l2

pytestmark = pytest.mark.gpu


class ThreadTransport:
    """Moves halo buffers between ranks running in threads of one process,
    with the pairing of the reference's messages: what rank i sends to rank j
    sits at i's send offset j (count j) going out, and lands at j's recv
    offset i; the return trip uses the same lists the other way round."""

    def __init__(self, maps: list[Maps]):
        self.size = len(maps)
        self.send = [m._send()[:2] for m in maps]   # counts, offsets
        self.recv = [m._recv() for m in maps]
        self.barrier = threading.Barrier(self.size, timeout=60)
        self.posted = [None] * self.size
        self.calls = [0] * self.size
        self._hip = None

    def _exchange(self, rank, direction, out: np.ndarray, into: np.ndarray):
        self.posted[rank] = out.copy()
        self.barrier.wait()
        for i in range(self.size):
            if i == rank:
                continue
            if direction == OWNER_TO_GHOST:   # my ghosts owned by i
                rc, ro = self.recv[rank]
                sc, so = self.send[i]
                assert sc[rank] == rc[i]
                into[ro[i]:ro[i] + rc[i]] = self.posted[i][so[rank]:so[rank] + sc[rank]]
            else:                             # i's ghosts that I own
                sc, so = self.send[rank]
                rc, ro = self.recv[i]
                assert rc[rank] == sc[i]
                into[so[i]:so[i] + sc[i]] = self.posted[i][ro[rank]:ro[rank] + rc[rank]]
        self.calls[rank] += 1
        self.barrier.wait()

    def host_fn(self, rank):
        def fn(direction, send, recv, stream):
            self._exchange(rank, direction, send, recv)
        return fn

    def device_fn(self, rank, sizes):
        """Device buffers: stream-synchronize, copy through host memory."""
        if self._hip is None:
            self._hip = C.CDLL("libamdhip64.so")
        hip = self._hip
        n_ghost, n_send = sizes

        def fn(direction, send, recv, stream):
            assert hip.hipStreamSynchronize(C.c_void_p(stream)) == 0
            n_out, n_in = (n_send, n_ghost) if direction == OWNER_TO_GHOST else (n_ghost, n_send)
            out, into = np.zeros(n_out), np.zeros(n_in)
            if n_out:
                assert hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(send), C.c_size_t(8 * n_out), 2) == 0
            self._exchange(rank, direction, out, into)
            if n_in:
                assert hip.hipMemcpy(C.c_void_p(recv), C.c_void_p(into.ctypes.data), C.c_size_t(8 * n_in), 1) == 0
        return fn


def partition(gp, m, size, prescribed=None, bc_meth=0, scatter=SCATTER.DETERMINISTIC, bcop=BCOP.IDENTITY,
              penalty=None, mode=HALO.HOST, finalize_first=True):
    """Operators of every rank of a z-slab partition with the transport set."""
    ranks = mesh.zslab_partition(m, size)
    maps = []
    for r in ranks:
        mp = Maps(r.rank, r.size)
        mp.set_map(len(r.elements), r.element_to_rank_map, None, r.n_all_dofs, r.rank_to_global_map,
                   r.owned_begin, r.owned_end, r.n_global, r.rank_dof_counts)
        maps.append(mp)
    from test_maps import alltoallv
    for mp, (c, g) in zip(maps, alltoallv([mp.get_ghost_requests() for mp in maps])):
        mp.set_peer_requests(c, g)
    tr = ThreadTransport(maps)
    ops = []
    for r, mp in zip(ranks, maps):
        cd = np.intersect1d(gp.constrained, np.unique(r.rank_to_global_map)).astype(np.uint64)
        cv = np.zeros(len(cd)) if prescribed is None else prescribed[cd.astype(np.int64)]
        mp.set_bdr_map(cd, cv)
        op = aMatFree(mp, bc_meth)
        op.set_matfree_type(3)
        op.set_element_matrices_scaled(gp.k_ref, gp.scale[r.elements])
        op.set_options(1, scatter)
        op.set_bc_operator(bcop)
        if penalty is not None:
            op.set_penalty(penalty)
        ops.append(op)
    fns = [tr.host_fn(r.rank) if mode == HALO.HOST else
           tr.device_fn(r.rank, (r.n_pre + r.n_post, int(mp.info().n_send))) for r, mp in zip(ranks, maps)]

    def setup(i):
        if finalize_first:   # the diagonal is assembled inside set_halo_transport
            ops[i].finalize()
            ops[i].set_halo_transport(fns[i], mode)
        else:                # ... or inside finalize
            ops[i].set_halo_transport(fns[i], mode)
            ops[i].finalize()
    on_ranks(len(ranks), setup)
    return ranks, maps, ops, tr


def on_ranks(size, fn):
    """fn(rank) in one thread per rank; the first failure is raised."""
    with cf.ThreadPoolExecutor(max_workers=size) as ex:
        return [f.result(timeout=120) for f in [ex.submit(fn, i) for i in range(size)]]


def owned(r, full):
    return full[r.owned_begin:r.owned_end + 1]


def test_host_forms_on_two_ranks():
    """matvec (owned-length and ghosted), MatMult_mf, the diagonal and
    apply_bc_rhs of each rank against the one-rank oracle, host buffers."""
    m = mesh.hex8(6, 5, 8)
    gp = problems.poisson_hex8(6, 5, 8)
    gp.scale = mesh.uniform(9, 1, gp.n_elem, 0.5, 2.0)
    pres = np.sin(0.37 * np.arange(gp.n_dofs))
    ranks, maps, ops, tr = partition(gp, m, 2, prescribed=pres)
    assert sum(r.n_pre + r.n_post for r in ranks) > 0
    K = gp.element_matrices()
    u = gp.random_u()
    kv = orc.matvec(gp.gmap, K, u)
    mm = orc.matmult_mf(gp.gmap, K, u, gp.constrained)
    bmaps = orc.set_bdr_map(0, 1, [gp.n_dofs], gp.gmap, gp.constrained, pres[gp.constrained.astype(np.int64)])
    diag = orc.get_diagonal(gp.gmap, K, bmaps["bdr_map"], 0, gp.n_dofs)
    c = gp.constrained.astype(np.int64)
    uc = np.zeros(gp.n_dofs)
    uc[c] = pres[c]
    rhs = mesh.uniform(5, 0, gp.n_dofs, -1, 1)
    exp_rhs = rhs - orc.matvec(gp.gmap, K, uc)
    exp_rhs[c] = pres[c]

    def work(i):
        r, op = ranks[i], ops[i]
        uo = owned(r, u).copy()
        v = np.zeros(r.n_owned)
        op.matvec(v, uo, False)                       # local_to_ghost: ghost slots from the owners
        assert rel_l2(v, owned(r, kv)) <= 1e-14
        ug = np.full(r.n_all_dofs, 1e30)              # ghost slots are overwritten by the exchange
        ug[r.n_pre:r.n_pre + r.n_owned] = uo
        vg = np.zeros(r.n_all_dofs)
        op.matvec(vg, ug, True)
        np.testing.assert_array_equal(vg[r.n_pre:r.n_pre + r.n_owned], v)
        assert rel_l2(op.MatMult_mf(uo), owned(r, mm)) <= 1e-14
        assert rel_l2(op.get_diagonal(), owned(r, diag)) <= 1e-14
        assert rel_l2(op.apply_bc(owned(r, rhs).copy()), owned(r, exp_rhs)) <= 1e-14
        return True
    assert all(on_ranks(2, work))
    # 1 exchange (diagonal) + 2 per matvec x 3 + 1 (rhs)
    assert tr.calls == [8, 8]


@pytest.mark.parametrize("scatter", [SCATTER.DETERMINISTIC, SCATTER.ATOMIC])
def test_device_forms_on_three_ranks(torch_gpu, scatter):
    """matmult_dev of a hex20 elasticity share with device buffers, both
    scatter modes (the atomic one zero-fills v itself), identity and penalty
    rows; the penalty coefficients use the diagonal assembled across ranks."""
    torch = torch_gpu
    m = mesh.hex20(3, 3, 4)
    gp = problems.elasticity_hex20(3, 3, 4)
    K = gp.element_matrices()
    u = gp.random_u()
    for bcop, penalty, want, tol in (
            (BCOP.IDENTITY, None, orc.matmult_mf(gp.gmap, K, u, gp.constrained), 1e-13),
            (BCOP.PENALTY_DIAG, 1e8, None, 1e-13)):
        if want is None:
            bmaps = orc.set_bdr_map(0, 1, [gp.n_dofs], gp.gmap, gp.constrained, gp.prescribed)
            want = orc.matmult_penalty(gp.gmap, K, bmaps["bdr_map"], penalty, u)
        ranks, maps, ops, tr = partition(gp, m, 3, scatter=scatter, bcop=bcop, penalty=penalty, mode=HALO.DEVICE,
                                         finalize_first=bcop == BCOP.IDENTITY)
        s = [torch.cuda.Stream() for _ in ranks]
        outs = [None] * 3

        def work(i):
            r = ranks[i]
            ug = np.zeros(r.n_all_dofs)
            ug[r.n_pre:r.n_pre + r.n_owned] = owned(r, u)
            with torch.cuda.stream(s[i]):
                du = torch.from_numpy(ug).cuda()
                dv = torch.full_like(du, 7.0)          # not zeroed by the caller
                ops[i].matmult_dev(dv, du, s[i])
                s[i].synchronize()
                outs[i] = dv.cpu().numpy()[r.n_pre:r.n_pre + r.n_owned]
        on_ranks(3, work)
        got = np.concatenate(outs)
        assert rel_l2(got, want) <= tol, (bcop, scatter)


def test_exchanging_forms_refuse_without_transport(torch_gpu):
    """On a rank's share without a transport the forms that need the halo
    refuse (the reference always exchanges); the two-phase pieces still run."""
    torch = torch_gpu
    m = mesh.hex8(4, 4, 4)
    gp = problems.poisson_hex8(4)
    r = mesh.zslab_partition(m, 2)[0]
    mp = Maps(0, 2)
    mp.set_map(len(r.elements), r.element_to_rank_map, None, r.n_all_dofs, r.rank_to_global_map, r.owned_begin,
               r.owned_end, r.n_global, r.rank_dof_counts)
    mp.set_bdr_map(np.zeros(0, np.uint64), np.zeros(0))
    op = aMatFree(mp)
    op.set_element_matrices_scaled(gp.k_ref, gp.scale[r.elements])
    op.finalize()
    du = torch.zeros(r.n_all_dofs, dtype=torch.float64, device="cuda")
    dv = torch.zeros_like(du)
    for fn in (lambda: op.matvec(np.zeros(r.n_owned), np.zeros(r.n_owned)),
               lambda: op.MatMult_mf(np.zeros(r.n_owned)),
               lambda: op.get_diagonal(),
               lambda: op.matmult_dev(dv, du),
               lambda: op.matvec_dev(dv, du)):
        with pytest.raises(AmatError) as e:
            fn()
        assert e.value.name == "NOT_IMPLEMENTED"
    with pytest.raises(AmatError) as e:            # no peer lists yet
        op.set_halo_transport(lambda *a: None)
    assert e.value.name == "INVALID_STATE"
    op.matvec_begin_dev(dv, du)
    op.matvec_end_dev(dv, du)
    torch.cuda.synchronize()


def test_one_rank_has_no_transport():
    gp = problems.poisson_hex8(3)
    op = gp.build_operator(aMatFree, Maps)
    with pytest.raises(AmatError) as e:
        op.set_halo_transport(lambda *a: None)
    assert e.value.name == "INVALID_STATE"


def test_transport_failure_is_reported():
    """A transport that fails makes the calling entry point return
    WRONG_COMMUNICATION (par::Error) instead of a partial result."""
    m = mesh.hex8(4, 4, 4)
    gp = problems.poisson_hex8(4)
    ranks, maps, ops, tr = partition(gp, m, 2)

    def bad(direction, send, recv, stream):
        raise RuntimeError("link down")
    for op in ops:
        op.set_halo_transport(bad)
    with pytest.raises(AmatError) as e:
        ops[0].matvec(np.zeros(ranks[0].n_owned), np.zeros(ranks[0].n_owned))
    assert e.value.name == "WRONG_COMMUNICATION"


def test_device_arrays_are_checked_before_launch(torch_gpu):
    """CPU tensors, numpy arrays, float32 tensors and short tensors are
    rejected by the Python binding before any kernel sees their address."""
    torch = torch_gpu
    gp = problems.poisson_hex8(3)
    op = gp.build_operator(aMatFree, Maps)
    n = gp.n_dofs
    good = torch.zeros(n, dtype=torch.float64, device="cuda")
    for bad, exc in ((torch.zeros(n, dtype=torch.float64), TypeError),
                     (np.zeros(n), TypeError),
                     (torch.zeros(n, dtype=torch.float32, device="cuda"), TypeError),
                     (torch.zeros(n - 1, dtype=torch.float64, device="cuda"), ValueError)):
        with pytest.raises(exc):
            op.matvec_dev(good, bad)
        with pytest.raises(exc):
            op.cg(bad, good)
    with pytest.raises(ValueError):
        op.matvec(np.zeros(n - 1), np.zeros(n), True)
    with pytest.raises(TypeError):
        op.matmult_f32_dev(2, good, good)
