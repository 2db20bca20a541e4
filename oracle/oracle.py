"""numpy wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU restatement of aMat's matvec path (amx_oracle.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
Element maps are (E, n) arrays (one DoF count) unless stated otherwise.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB = _HERE / "liboracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
        _lib = C.CDLL(str(_LIB))
        _lib.orc_trace.restype = C.c_double
        _lib.orc_pcg.restype = C.c_int
    return _lib


def native_library(outdir) -> C.CDLL:
    """The restatement built with -O3 -march=native for this host into
    `outdir` (the CPU baseline's vectorized leg)."""
    outdir = pathlib.Path(outdir)
    subprocess.run(["make", "-s", "-C", str(_HERE), "native", f"NATIVE_OUT={outdir}"], check=True)
    return C.CDLL(str(outdir / "liboracle_native.so"))


def _p(a):
    # data_as keeps a reference to the array for as long as the pointer lives,
    # so temporaries stay alive during the call
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _u32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.uint32)


def _u64(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.uint64)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.float64)


def _dpe(lmap):
    lmap = np.asarray(lmap)
    return np.full(lmap.shape[0], lmap.shape[1], dtype=np.uint32)


def build_scatter_map(rank, size, rank_dof_counts, element_to_rank_map, n_all_dofs, rank_to_global, owned_begin,
                      owned_end, n_global, dofs_per_elem=None):
    e2r = _u32(element_to_rank_map).reshape(-1)
    dpe = _u32(dofs_per_elem) if dofs_per_elem is not None else _dpe(element_to_rank_map)
    tot = int(dpe.sum())
    gm = np.zeros(tot, np.uint64)
    lm = np.zeros(tot, np.uint32)
    l2g = np.zeros(n_all_dofs, np.uint64)
    l2r = np.zeros(n_all_dofs, np.uint32)
    npre, npost = C.c_uint32(), C.c_uint32()
    rc_ = np.zeros(size, np.uint32)
    req = np.zeros(max(n_all_dofs, 1), np.uint64)
    counts = _u32(rank_dof_counts)
    r2g = _u64(rank_to_global)
    err = lib().orc_build_scatter_map(
        C.c_uint32(rank), C.c_uint32(size), _p(counts), C.c_uint32(len(dpe)), _p(dpe), _p(e2r), C.c_uint32(n_all_dofs),
        _p(r2g), C.c_uint64(owned_begin), C.c_uint64(owned_end), C.c_uint64(n_global), _p(gm), _p(lm), _p(l2g),
        _p(l2r), C.byref(npre), C.byref(npost), _p(rc_), _p(req))
    nreq = npre.value + npost.value
    return dict(err=err, global_map=gm, local_map=lm, local2global=l2g, localid2rank=l2r, n_pre=npre.value,
                n_post=npost.value, ghost_req_counts=rc_, ghost_req_gids=req[:nreq])


def set_send_ids(rank, size, rank_dof_counts, n_pre, recv_gids):
    g = _u64(recv_gids)
    out = np.zeros(max(len(g), 1), np.uint32)
    err = lib().orc_set_send_ids(C.c_uint32(rank), C.c_uint32(size), _p(_u32(rank_dof_counts)), C.c_uint32(n_pre),
                                 C.c_uint32(len(g)), _p(g) if len(g) else None, _p(out))
    return err, out[:len(g)]


def set_bdr_map(rank, size, rank_dof_counts, global_map_2d, constrained, values):
    gm = _u64(global_map_2d).reshape(-1)
    dpe = _dpe(global_map_2d)
    cd, cv = _u64(constrained), _f64(values)
    n = len(cd)
    oc, ov = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1))
    bm, pv = np.zeros(len(gm), np.uint32), np.zeros(len(gm))
    of = np.zeros(max(len(gm), 1), np.uint64)
    noc, nof = C.c_uint32(), C.c_uint32()
    lib().orc_set_bdr_map(C.c_uint32(rank), C.c_uint32(size), _p(_u32(rank_dof_counts)), C.c_uint32(len(dpe)), _p(dpe),
                          _p(gm), _p(cd) if n else None, _p(cv) if n else None, C.c_uint32(n), _p(oc), _p(ov),
                          C.byref(noc), _p(bm), _p(pv), _p(of), C.byref(nof))
    return dict(owned_cdofs=oc[:noc.value], owned_cvals=ov[:noc.value], bdr_map=bm, presval_map=pv,
                owned_free=of[:nof.value])


def identify_independent(rank, size, rank_dof_counts, global_map_2d):
    gm = _u64(global_map_2d).reshape(-1)
    dpe = _dpe(global_map_2d)
    E = len(dpe)
    a, b = np.zeros(max(E, 1), np.uint32), np.zeros(max(E, 1), np.uint32)
    na, nb = C.c_uint32(), C.c_uint32()
    lib().orc_identify_independent(C.c_uint32(rank), C.c_uint32(size), _p(_u32(rank_dof_counts)), C.c_uint32(E),
                                   _p(dpe), _p(gm), _p(a), C.byref(na), _p(b), C.byref(nb))
    return a[:na.value], b[:nb.value]


def matvec(local_map, K, u, n_total=None, layout=0, nthreads=1, indep=None, dep=None, library=None):
    """matvec_ghosted_OMP; K is (E, n, n) row-major (layout 0) or the same
    matrices stored column-major per element (layout 1: pass K transposed).
    library: another build of the restatement (native_library)."""
    lm = _u32(local_map)
    E, n = lm.shape
    Kf = _f64(K).reshape(-1)
    uu = _f64(u)
    nt = len(uu) if n_total is None else n_total
    v = np.zeros(nt)
    ind = _u32(np.arange(E) if indep is None else indep)
    de = _u32(np.zeros(0) if dep is None else dep)
    (library or lib()).orc_matvec_ghosted_omp(C.c_uint32(E), _p(_dpe(lm)), _p(lm), _p(Kf), C.c_int(layout), _p(ind),
                                 C.c_uint32(len(ind)), _p(de) if len(de) else None, C.c_uint32(len(de)), _p(uu), _p(v),
                                 C.c_uint32(nt), C.c_int(nthreads))
    return v


def matvec_blocks(dofs_per_elem, blocks_dim, block_present, local_map_flat, K_flat, u, n_total):
    v = np.zeros(n_total)
    lib().orc_matvec_blocks(C.c_uint32(len(dofs_per_elem)), _p(_u32(dofs_per_elem)), _p(_u32(blocks_dim)),
                            _p(np.ascontiguousarray(block_present, dtype=np.uint8)), _p(_u32(local_map_flat)),
                            _p(_f64(K_flat)), _p(_f64(u)), _p(v), C.c_uint32(n_total))
    return v


def matmult_mf(local_map, K, uu, owned_cdofs, n_pre=0, n_owned=None, n_total=None, dof_scan_rank=0, nthreads=1):
    lm = _u32(local_map)
    E, n = lm.shape
    uu = _f64(uu)
    no = len(uu) if n_owned is None else n_owned
    nt = no if n_total is None else n_total
    vv = np.zeros(no)
    oc = _u64(owned_cdofs)
    lib().orc_matmult_mf(C.c_uint32(E), _p(_dpe(lm)), _p(lm), _p(_f64(K).reshape(-1)), C.c_uint32(nt),
                         C.c_uint32(n_pre), C.c_uint32(no), _p(oc) if len(oc) else None, C.c_uint32(len(oc)),
                         C.c_uint64(dof_scan_rank), _p(uu), _p(vv), C.c_int(nthreads))
    return vv


def matmult_penalty(local_map, K, bdr_map, penalty, u):
    lm = _u32(local_map)
    E, n = lm.shape
    u = _f64(u)
    v = np.zeros(len(u))
    lib().orc_matmult_penalty(C.c_uint32(E), _p(_dpe(lm)), _p(lm), _p(_f64(K).reshape(-1)),
                              _p(_u32(bdr_map).reshape(-1)), C.c_double(penalty), C.c_uint32(len(u)), _p(u), _p(v))
    return v


def trace(K):
    K = _f64(K)
    E, n, _ = K.shape
    return lib().orc_trace(C.c_uint32(E), _p(np.full(E, n, np.uint32)), _p(K.reshape(-1)))


def get_diagonal(local_map, K, bdr_map, bc_meth, n_total, penalty_override=0.0):
    lm = _u32(local_map)
    E, n = lm.shape
    d = np.zeros(n_total)
    bm = None if bdr_map is None else _u32(bdr_map).reshape(-1)
    lib().orc_get_diagonal(C.c_uint32(E), _p(_dpe(lm)), _p(lm), _p(_f64(K).reshape(-1)), _p(bm), C.c_int(bc_meth),
                           C.c_double(penalty_override), C.c_uint32(n_total), _p(d))
    return d


def build_ew_maps(local_map, n_total):
    lm = _u32(local_map)
    E, n = lm.shape
    off = np.zeros(n_total + 1, np.uint32)
    pos = np.zeros(E * n, np.uint32)
    lib().orc_build_ew_maps(C.c_uint32(E), _p(_dpe(lm)), _p(lm), C.c_uint32(n_total), _p(off), _p(pos))
    return off, pos


def gather_ew(off, pos, ve, n_total):
    v = np.zeros(n_total)
    lib().orc_gather_ew(C.c_uint32(n_total), _p(_u32(off)), _p(_u32(pos)), _p(_f64(ve)), _p(v))
    return v


def element_products(local_map, K, u):
    lm = _u32(local_map)
    E, n = lm.shape
    ve = np.zeros(E * n)
    lib().orc_element_products(C.c_uint32(E), _p(_dpe(lm)), _p(lm), _p(_f64(K).reshape(-1)), _p(_f64(u)), _p(ve))
    return ve


def pcg(op, local_map, K, bdr_map, owned_cdofs, penalty, b, rtol=1e-10, max_it=10000):
    lm = _u32(local_map)
    E, n = lm.shape
    b = _f64(b)
    x = np.zeros(len(b))
    rr = C.c_double()
    oc = _u64(owned_cdofs)
    it = lib().orc_pcg(C.c_int(op), C.c_uint32(E), _p(_dpe(lm)), _p(lm), _p(_f64(K).reshape(-1)),
                       _p(_u32(bdr_map).reshape(-1)), _p(oc) if len(oc) else None, C.c_uint32(len(oc)),
                       C.c_double(penalty), C.c_uint32(len(b)), _p(b), _p(x), C.c_double(rtol), C.c_int(max_it),
                       C.byref(rr))
    return x, it, rr.value
