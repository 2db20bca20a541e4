"""Python mirror of aMat's Maps / aMatFree interface over the C ABI.

Names, argument meanings and error behaviour follow the reference classes
(include/maps.hpp, include/aMatFree.hpp, include/aMat.hpp of paralab/aMat) so
that code written against them reads the same; every call goes to
amat_amd/libamat_amd.so and raises AmatError with the par::Error name on a
non-zero status.  Device arrays are torch tensors on a HIP device; host arrays
are numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import enum
import sys
import traceback
from typing import Callable, Sequence

import numpy as np

from . import _capi
from ._capi import AmatError, call


class BC_METH(enum.IntEnum):
    """par::BC_METH (include/enums.hpp:20-22)."""
    BC_IMATRIX = 0
    BC_PENALTY = 1


class MATFREE_TYPE(enum.IntEnum):
    """par::MATFREE_TYPE (include/enums.hpp:56)."""
    UNDEFINED = 0
    QUASI_FREE = 1
    FREE = 2
    GPU_PURE = 3
    GPU_OVER_CPU = 4
    GPU_OVER_GPU = 5


class Error(enum.IntEnum):
    """par::Error (include/enums.hpp:24-37) plus this library's additions."""
    SUCCESS = 0
    INDEX_OUT_OF_BOUNDS = 1
    UNKNOWN_ELEMENT_TYPE = 2
    UNKNOWN_ELEMENT_STATUS = 3
    NULL_L2G_MAP = 4
    GHOST_NODE_NOT_FOUND = 5
    UNKNOWN_MAT_TYPE = 6
    WRONG_COMMUNICATION = 7
    UNKNOWN_DOF_TYPE = 8
    UNKNOWN_CONSTRAINT = 9
    UNKNOWN_BC_METH = 10
    NOT_IMPLEMENTED = 11
    GLOBAL_DOF_ID_NOT_FOUND = 12
    INVALID_ARGUMENT = 100
    DEVICE_ERROR = 101
    OUT_OF_MEMORY = 102
    INVALID_STATE = 103
    NOT_CONVERGED = 104
    REJECTED = 105


class PREC(enum.IntFlag):
    FP64 = 1
    FP32 = 2
    BF16 = 4


# refinement options (or-ed into ir_cg's precision): inner CG vectors in FP64;
# the reliable-update form (one unrestarted low-precision recurrence whose
# residual is replaced by the FP64 true residual; inner_rtol = threshold)
IR_F64_VECTORS = 0x100
IR_RELIABLE = 0x200


class MATRIX_TYPE(enum.IntEnum):
    """par::MATRIX_TYPE (include/enums.hpp:52-53)."""
    MATRIX_BASED = 0
    MATRIX_FREE = 1


class ELEM(enum.IntEnum):
    """Element kinds of method 2 (FREE), computed in-kernel from coordinates."""
    HEX8_LAPLACE = 1
    HEX20_ELASTIC = 2
    TET10_LAPLACE = 3


class SCATTER(enum.IntEnum):
    DETERMINISTIC = 0
    ATOMIC = 1


class BCOP(enum.IntEnum):
    NONE = 0
    IDENTITY = 1       # aMatFree::MatMult_mf
    PENALTY_DIAG = 2   # aMatBased::apply_bc_mat penalty branch


class HALO(enum.IntEnum):
    """Buffers handed to a halo transport (amx_halo_mode)."""
    HOST = 0     # numpy views of pinned host buffers (an MPI caller)
    DEVICE = 1   # device addresses and the stream (an RCCL caller)


NOALLOC_MAX_DOFS = 1 << 20   # AMX_NOALLOC_MAX_DOFS: largest operator of cg_noalloc

OWNER_TO_GHOST = 0   # ghost_receive_begin/end (include/aMatFree.hpp:1426-1549)
GHOST_TO_OWNER = 1   # ghost_send_begin/end (:1552-1679)


def _ptr(a) -> int | None:
    """Address of a contiguous numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    raise TypeError(f"unsupported array type {type(a)}")


def _dev(t, n: int, name: str, device: int, dtype: str = "float64") -> int:
    """Address of a contiguous torch tensor on HIP device `device` with dtype
    `dtype` and at least n entries; raises before anything is launched."""
    if t is None or not hasattr(t, "data_ptr"):
        raise TypeError(f"{name} must be a torch tensor on a HIP device")
    if not getattr(t, "is_cuda", False):
        raise TypeError(f"{name} is on {t.device}; it must be on a HIP device")
    if t.device.index is not None and t.device.index != device:
        raise ValueError(f"{name} is on device {t.device.index}; the operator is on device {device}")
    if str(t.dtype) != "torch." + dtype:
        raise TypeError(f"{name} is {t.dtype}; expected torch.{dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < n:
        raise ValueError(f"{name} has {t.numel()} entries; the operator needs {n}")
    return t.data_ptr()


def _host(a, n: int, name: str) -> int:
    """Address of a contiguous float64 numpy array with at least n entries."""
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a contiguous float64 numpy array")
    if a.size < n:
        raise ValueError(f"{name} has {a.size} entries; the operator needs {n}")
    return a.ctypes.data


def _arr(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def _flatten_rows(rows, dtype) -> tuple[np.ndarray, np.ndarray]:
    """A 2-D array or a sequence of 1-D rows -> (flat values, row lengths)."""
    if isinstance(rows, np.ndarray) and rows.ndim == 2:
        return _arr(rows.reshape(-1), dtype), np.full(rows.shape[0], rows.shape[1], dtype=np.uint32)
    rows = [np.asarray(r) for r in rows]
    lens = np.array([len(r) for r in rows], dtype=np.uint32)
    flat = _arr(np.concatenate(rows) if rows else np.zeros(0), dtype)
    return flat, lens


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


class Maps:
    """par::Maps (include/maps.hpp) for one rank."""

    def __init__(self, rank: int = 0, size: int = 1):
        h = C.c_void_p()
        call("amx_maps_create", rank, size, C.byref(h))
        self._h = h
        self.rank, self.size = rank, size
        self._dpe = np.zeros(0, dtype=np.uint32)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().amx_maps_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    # -- set-up --------------------------------------------------------
    def set_map(self, n_elements_on_rank, element_to_rank_map, dofs_per_element, n_all_dofs_on_rank,
                rank_to_global_map, owned_global_dof_range_begin, owned_global_dof_range_end, n_global_dofs,
                rank_dof_counts=None):
        """Maps::set_map (include/maps.hpp:398-451); element_to_rank_map is a
        2-D array or a list of per-element rows."""
        flat, lens = _flatten_rows(element_to_rank_map, np.uint32)
        dpe = _arr(dofs_per_element, np.uint32) if dofs_per_element is not None else lens
        if len(dpe) != n_elements_on_rank:
            raise ValueError("dofs_per_element length differs from n_elements_on_rank")
        r2g = _arr(rank_to_global_map, np.uint64)
        counts = None if rank_dof_counts is None else _arr(rank_dof_counts, np.uint32)
        call("amx_maps_set_map", self._h, n_elements_on_rank, _ptr(flat), _ptr(dpe), n_all_dofs_on_rank, _ptr(r2g),
             int(owned_global_dof_range_begin), int(owned_global_dof_range_end), int(n_global_dofs), _ptr(counts))
        self._dpe = dpe.copy()

    def get_ghost_requests(self) -> tuple[np.ndarray, np.ndarray]:
        """(counts per owner rank, ghost global ids) -- what buildScatterMap's
        Alltoallv sends (include/maps.hpp:716-739)."""
        info = self.info()
        counts = np.zeros(self.size, dtype=np.uint32)
        gids = np.zeros(max(info.n_recv, 1), dtype=np.uint64)
        call("amx_maps_get_ghost_requests", self._h, _ptr(counts), _ptr(gids))
        return counts, gids[:info.n_recv]

    def set_peer_requests(self, counts, gids):
        counts = _arr(counts, np.uint32)
        gids = _arr(gids, np.uint64) if len(gids) else np.zeros(1, dtype=np.uint64)
        call("amx_maps_set_peer_requests", self._h, _ptr(counts), _ptr(gids))

    def set_bdr_map(self, constrainedDofs, prescribedValues, numConstraints=None):
        """Maps::set_bdr_map (include/maps.hpp:828-880)."""
        cd = _arr(constrainedDofs, np.uint64)
        pv = _arr(prescribedValues, np.float64)
        n = len(cd) if numConstraints is None else int(numConstraints)
        if len(cd) == 0:
            cd, pv = np.zeros(1, np.uint64), np.zeros(1)
        call("amx_maps_set_bdr_map", self._h, _ptr(cd), _ptr(pv), n)

    # -- getters (include/maps.hpp:85-276) ------------------------------
    def info(self) -> _capi.MapsInfo:
        inf = _capi.MapsInfo()
        call("amx_maps_get_info", self._h, C.byref(inf))
        return inf

    def _get(self, fn: str, n: int, dtype):
        out = np.zeros(max(n, 1), dtype=dtype)
        call(fn, self._h, _ptr(out))
        return out[:n]

    def get_NumElems(self): return self.info().n_elems
    def get_NumDofs(self): return self.info().n_owned
    def get_NumDofsTotal(self): return self.info().n_total
    def get_NumPreGhostDofs(self): return self.info().n_pre_ghost
    def get_NumPostGhostDofs(self): return self.info().n_post_ghost
    def get_DofsPerElem(self): return self._dpe.copy()
    def get_LocalMap(self): return self._get("amx_maps_get_local_map", self.info().n_entries, np.uint32)
    def get_Map(self): return self._get("amx_maps_get_global_map", self.info().n_entries, np.uint64)
    def get_Local2Global(self): return self._get("amx_maps_get_local2global", self.info().n_total, np.uint64)
    def get_localID2Rank(self): return self._get("amx_maps_get_localid2rank", self.info().n_total, np.uint32)
    def get_ownedFreeDofs(self): return self._get("amx_maps_get_owned_free", self.info().n_owned_free, np.uint64)
    def get_LocalDofScan(self): return self._get("amx_maps_get_dof_scan", self.size, np.uint64)

    def get_BdrMap(self):
        n = self.info().n_entries
        m, v = np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1))
        call("amx_maps_get_bdr_map", self._h, _ptr(m), _ptr(v))
        return m[:n]

    def get_PresValMap(self):
        n = self.info().n_entries
        m, v = np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1))
        call("amx_maps_get_bdr_map", self._h, _ptr(m), _ptr(v))
        return v[:n]

    def _constraints(self):
        n = self.info().n_owned_constraints
        d, v = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1))
        call("amx_maps_get_owned_constraints", self._h, _ptr(d), _ptr(v))
        return d[:n], v[:n]

    def get_ownedConstrainedDofs(self): return self._constraints()[0]
    def get_ownedPrescribedValues(self): return self._constraints()[1]
    def get_n_owned_constraints(self): return self.info().n_owned_constraints

    def _elem_lists(self):
        inf = self.info()
        a, b = np.zeros(max(inf.n_independent, 1), np.uint32), np.zeros(max(inf.n_dependent, 1), np.uint32)
        call("amx_maps_get_element_lists", self._h, _ptr(a), _ptr(b))
        return a[:inf.n_independent], b[:inf.n_dependent]

    def get_independentElem(self): return self._elem_lists()[0]
    def get_dependentElem(self): return self._elem_lists()[1]

    def _send(self):
        inf = self.info()
        c, o = np.zeros(self.size, np.uint32), np.zeros(self.size, np.uint32)
        ids = np.zeros(max(inf.n_send, 1), np.uint32)
        call("amx_maps_get_send_lists", self._h, _ptr(c), _ptr(o), _ptr(ids))
        return c, o, ids[:inf.n_send]

    def _recv(self):
        c, o = np.zeros(self.size, np.uint32), np.zeros(self.size, np.uint32)
        call("amx_maps_get_recv_lists", self._h, _ptr(c), _ptr(o))
        return c, o

    def get_SendDofCounts(self): return self._send()[0]
    def get_SendDofOffset(self): return self._send()[1]
    def get_SendDofIds(self): return self._send()[2]
    def get_SendRankIds(self): return np.nonzero(self._send()[0])[0].astype(np.uint32)
    def get_RecvDofCounts(self): return self._recv()[0]
    def get_RecvDofOffset(self): return self._recv()[1]
    def get_RecvRankIds(self): return np.nonzero(self._recv()[0])[0].astype(np.uint32)


def measure_read_bandwidth(device: int = 0, nbytes: int = 1 << 30, iters: int = 10) -> float:
    """Achievable HBM read bandwidth of the device in bytes/s (amx_measure_read_bandwidth)."""
    v = C.c_double()
    call("amx_measure_read_bandwidth", int(device), int(nbytes), int(iters), C.byref(v))
    return v.value


class aMatFree:
    """par::aMatFree (include/aMatFree.hpp) running MATFREE_TYPE 2 (built-in element
    kinds) and 3/4/5 on one GPU."""

    def __init__(self, mesh_maps: Maps, bcType: BC_METH = BC_METH.BC_IMATRIX, device: int = 0):
        h = C.c_void_p()
        call("amx_op_create", mesh_maps.handle, int(bcType), int(device), C.byref(h))
        self._h = h
        self.maps = mesh_maps
        self.device = device
        # the operator copied these from the maps at creation; the argument
        # checks of every call use them
        inf = mesh_maps.info()
        self._size_t = (inf.n_total, inf.n_owned, inf.n_pre_ghost + inf.n_post_ghost, inf.n_send)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().amx_op_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    # -- set-up --------------------------------------------------------
    def set_matfree_type(self, t: MATFREE_TYPE):
        call("amx_op_set_matfree_type", self._h, int(t))

    def set_num_streams(self, n: int):
        call("amx_op_set_num_streams", self._h, int(n))

    def set_element_matrix(self, eid: int, e_mat, block_i: int = 0, block_j: int = 0, blocks_dim: int = 1):
        """aMatFree::set_element_matrix (include/aMatFree.hpp:854-880); e_mat is an
        n_b x n_b array (row-major numpy)."""
        m = _arr(e_mat, np.float64)
        call("amx_op_set_element_matrix", self._h, eid, _ptr(m), 0, block_i, block_j, blocks_dim)

    def set_element_matrices(self, first_eid: int, e_mats):
        m = _arr(e_mats, np.float64)
        count = m.shape[0] if m.ndim == 3 else None
        if count is None:
            raise ValueError("e_mats must be (count, n, n)")
        call("amx_op_set_element_matrices", self._h, first_eid, count, _ptr(m), 0)

    def set_element_matrices_scaled(self, k_ref, scale):
        kr = _arr(k_ref, np.float64)
        sc = _arr(scale, np.float64)
        call("amx_op_set_element_matrices_scaled", self._h, _ptr(kr), 0, _ptr(sc))

    def set_element_function(self, kind: ELEM, xe, scale=None, E: float = 1.0, nu: float = 0.3, ngt: int = 2):
        """aMat::set_element_matrix_function (include/aMat.hpp:114-116) for method 2
        (FREE): xe is (n_elems, nodes, 3) node coordinates in element DoF-node
        order; K_e = scale[e] * ke(xe[e]) is applied in-kernel at every matvec."""
        x = _arr(xe, np.float64)
        sc = None if scale is None else _arr(scale, np.float64)
        call("amx_op_set_element_function", self._h, int(kind), _ptr(x), _ptr(sc), float(E), float(nu), int(ngt))

    def set_options(self, precision=PREC.FP64, scatter=SCATTER.DETERMINISTIC):
        call("amx_op_set_options", self._h, int(precision), int(scatter))

    def set_symmetric_storage(self, mode: int):
        """1 (default): pack the upper triangles of 8-/10-DoF element matrices
        when all are bitwise symmetric, and keep 60-DoF FP64 matrices that are
        symmetric up to low-word round-off as near-symmetric records (every
        entry rebuilt bit for bit); 0: whole matrices (next finalize)."""
        call("amx_op_set_symmetric_storage", self._h, int(mode))

    def set_penalty(self, p: float):
        call("amx_op_set_penalty", self._h, float(p))

    def set_bc_operator(self, mode: BCOP):
        call("amx_op_set_bc_operator", self._h, int(mode))

    def finalize(self):
        call("amx_op_finalize", self._h)

    def finalize_begin(self):
        """aMatFree::finalize_begin (include/aMatFree.hpp:656-745): the whole set-up."""
        self.finalize()

    def finalize_end(self):
        """aMatFree::finalize_end (include/aMatFree.hpp:748-785): the trace is
        computed in finalize_begin; checks that it ran."""
        if not self.info().finalized:
            raise AmatError(Error.INVALID_STATE, "finalize_end: finalize_begin was not called")

    def get_matrix_type(self) -> MATRIX_TYPE:
        return MATRIX_TYPE.MATRIX_FREE

    def dump_mat(self, filename=None) -> np.ndarray:
        """aMatFree::dump_mat (include/aMatFree.hpp:4737-4830) on one rank: the
        matrix of matvec (no boundary conditions) probed column by column with
        unit vectors.  Returns (row, col, value) triplets with |value| > 1e-16
        sorted by row then column, global ids; with a filename also writes them
        as "row col value" lines (the reference prints PETSc's MatView)."""
        mi = self.maps.info()
        if mi.size != 1:
            raise AmatError(Error.NOT_IMPLEMENTED, "dump_mat: one rank")
        n, g0 = mi.n_total, int(mi.owned_begin)
        ui, vi = np.zeros(n), np.zeros(n)
        rows, cols, vals = [], [], []
        for j in range(n):
            ui[j] = 1.0
            self.matvec(vi, ui, True)
            ui[j] = 0.0
            nz = np.nonzero(np.abs(vi) > 1e-16)[0]
            rows.append(nz)
            cols.append(np.full(nz.size, j))
            vals.append(vi[nz].copy())
        r = (np.concatenate(rows) if rows else np.zeros(0, np.int64)) + g0
        c = (np.concatenate(cols) if cols else np.zeros(0, np.int64)) + g0
        v = np.concatenate(vals) if vals else np.zeros(0)
        order = np.lexsort((c, r))
        trip = np.rec.fromarrays([r[order], c[order], v[order]], names="row,col,value")
        if filename is not None:
            with open(filename, "w") as f:
                for t in trip:
                    f.write(f"{int(t.row)} {int(t.col)} {float(t.value)!r}\n")
        return trip

    def info(self) -> _capi.OpInfo:
        inf = _capi.OpInfo()
        call("amx_op_get_info", self._h, C.byref(inf))
        return inf

    # -- host-array forms ----------------------------------------------
    def _sizes(self):
        """(n_total, n_owned, ghosts, send entries) of the operator's share."""
        return self._size_t

    def matvec(self, v: np.ndarray, u: np.ndarray, isGhosted: bool = False):
        """aMatFree::matvec (include/aMatFree.hpp:1681-1775): v = K u; length
        n_total when isGhosted, else n_owned.  More than one rank: with the
        halo transport (set_halo_transport)."""
        n_total, n_owned, _, _ = self._sizes()
        n = n_total if isGhosted else n_owned
        uu = _arr(u, np.float64)
        pu, pv = _host(uu, n, "u"), _host(v, n, "v")
        call("amx_op_matvec", self._h, pv, pu, int(bool(isGhosted)))
        return v

    def MatMult_mf(self, u: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        """aMatFree::MatMult_mf (include/aMatFree.hpp:3893-3941) on owned-length arrays
        (`out`: a contiguous float64 array of u's length to write v into)."""
        n_owned = self._sizes()[1]
        uu = _arr(u, np.float64)
        if out is None:
            out = np.zeros_like(uu)
        elif out.dtype != np.float64 or not out.flags.c_contiguous or out.shape != uu.shape:
            raise ValueError("out must be a contiguous float64 array of u's shape")
        call("amx_op_matmult_mf", self._h, _host(out, n_owned, "v"), _host(uu, n_owned, "u"))
        return out

    def get_diagonal(self) -> np.ndarray:
        """MatGetDiagonal_mf (include/aMatFree.hpp:3944-3974), owned length."""
        d = np.zeros(max(self.info().n_owned, 1))
        call("amx_op_get_diagonal", self._h, _ptr(d))
        return d[:self.info().n_owned]

    def apply_bc(self, rhs: np.ndarray) -> np.ndarray:
        """aMatFree::apply_bc -> apply_bc_rhs (include/aMatFree.hpp:957-971, 4536-4700)."""
        r = _arr(rhs, np.float64).copy()
        call("amx_op_apply_bc_rhs", self._h, _host(r, self._sizes()[1], "rhs"))
        return r

    # -- device forms (torch tensors) -----------------------------------
    # vectors: float64 (float32 for the FP32 forms) tensors on the operator's
    # device with at least n_total entries; checked before any launch
    def _vec(self, t, name, dtype="float64"):
        return _dev(t, self._sizes()[0], name, self.device, dtype)

    def matvec_dev(self, v, u, stream=None):
        call("amx_op_matvec_dev", self._h, self._vec(v, "v"), self._vec(u, "u"), _stream_ptr(stream))

    def matmult_dev(self, v, u, stream=None):
        call("amx_op_matmult_dev", self._h, self._vec(v, "v"), self._vec(u, "u"), _stream_ptr(stream))

    def matmult_f32_dev(self, precision: PREC, v, u, stream=None):
        call("amx_op_matmult_f32_dev", self._h, int(precision), self._vec(v, "v", "float32"),
             self._vec(u, "u", "float32"), _stream_ptr(stream))

    def matvec_begin_dev(self, v, u, apply_bc=True, stream=None):
        call("amx_op_matvec_begin_dev", self._h, self._vec(v, "v"), self._vec(u, "u"), int(apply_bc),
             _stream_ptr(stream))

    def matvec_end_dev(self, v, u, apply_bc=True, stream=None):
        call("amx_op_matvec_end_dev", self._h, self._vec(v, "v"), self._vec(u, "u"), int(apply_bc),
             _stream_ptr(stream))

    def halo_pack_dev(self, u, send, stream=None):
        """ghost_receive_begin's pack (include/aMatFree.hpp:1452-1454): send[i] = u[send_ids[i]]."""
        n_send = self._sizes()[3]
        call("amx_op_halo_pack_dev", self._h, self._vec(u, "u"),
             _dev(send, n_send, "send", self.device) if n_send else None, _stream_ptr(stream))

    def halo_import_dev(self, u, halo, stream=None):
        n_ghost = self._sizes()[2]
        call("amx_op_halo_import_dev", self._h, self._vec(u, "u"),
             _dev(halo, n_ghost, "halo", self.device) if n_ghost else None, _stream_ptr(stream))

    def halo_export_dev(self, v, halo, stream=None):
        n_ghost = self._sizes()[2]
        call("amx_op_halo_export_dev", self._h, self._vec(v, "v"), _dev(halo, n_ghost, "halo", self.device),
             _stream_ptr(stream))

    def halo_accumulate_dev(self, v, recv, stream=None):
        n_send = self._sizes()[3]
        call("amx_op_halo_accumulate_dev", self._h, self._vec(v, "v"),
             _dev(recv, n_send, "recv", self.device) if n_send else None, _stream_ptr(stream))

    def bc_fixup_dev(self, v, u, stream=None):
        call("amx_op_bc_fixup_dev", self._h, self._vec(v, "v"), self._vec(u, "u"), _stream_ptr(stream))

    def get_diagonal_dev(self, diag, stream=None):
        call("amx_op_get_diagonal_dev", self._h, self._vec(diag, "diag"), _stream_ptr(stream))

    # -- halo transport -------------------------------------------------
    def set_halo_transport(self, fn: Callable | None, mode: HALO = HALO.HOST):
        """The exchange of a rank's share (amx_op_set_halo_transport):
        fn(direction, send, recv, stream) moves send to the peers and fills
        recv.  direction OWNER_TO_GHOST: send = u at the send ids (n_send),
        recv = ghost values (n_pre + n_post, recv offsets); GHOST_TO_OWNER:
        send = ghost slots of v (n_pre + n_post), recv = owners' contributions
        (n_send, send offsets).  HOST: numpy views of pinned host buffers;
        DEVICE: device addresses (ints) and the stream address.  None removes
        the transport.  The call is collective: every rank makes it."""
        if fn is None:
            call("amx_op_set_halo_transport", self._h, _capi.HALO_FN(), None, int(mode))
            self._halo_cb = None
            return
        _, _, n_ghost, n_send = self._sizes()
        mode = HALO(mode)

        def cb(ctx, direction, send, recv, stream):
            try:
                n_out, n_in = (n_send, n_ghost) if direction == OWNER_TO_GHOST else (n_ghost, n_send)
                if mode == HALO.HOST:
                    so = np.ctypeslib.as_array(send, shape=(n_out,)) if n_out else np.zeros(0)
                    ri = np.ctypeslib.as_array(recv, shape=(n_in,)) if n_in else np.zeros(0)
                    fn(direction, so, ri, stream)
                else:
                    fn(direction, C.cast(send, C.c_void_p).value, C.cast(recv, C.c_void_p).value, stream)
                return 0
            except Exception:  # reported through the calling entry point's status
                traceback.print_exc(file=sys.stderr)
                return 1

        self._halo_cb = _capi.HALO_FN(cb)   # kept alive with the operator
        call("amx_op_set_halo_transport", self._h, self._halo_cb, None, int(mode))

    def get_timer(self) -> tuple:
        """aMatFree::get_timer (include/aMatFree.hpp:239-245): seconds
        (scatter, gather, mv, total) accumulated since finalize by the
        host-array forms (matvec, MatMult_mf): u copied in, v copied out, the
        device application, the whole calls."""
        s, g, m, t = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        call("amx_op_get_timer", self._h, C.byref(s), C.byref(g), C.byref(m), C.byref(t))
        return s.value, g.value, m.value, t.value

    def time_kernels_dev(self, n_iter, v, u, stream=None):
        """Mean ms of the element kernels, the gather and the whole application
        over n_iter runs (hipEvents on the launch stream)."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        call("amx_op_time_kernels_dev", self._h, int(n_iter), self._vec(v, "v"), self._vec(u, "u"), C.byref(a),
             C.byref(b), C.byref(c), _stream_ptr(stream))
        return a.value, b.value, c.value

    def stream(self) -> int:
        return _capi.lib().amx_op_stream(self._h) or 0

    def synchronize(self):
        call("amx_op_synchronize", self._h)

    # -- solvers ---------------------------------------------------------
    def cg(self, b, x, rtol=1e-10, max_it=10000, stream=None) -> _capi.SolveStats:
        st = _capi.SolveStats()
        call("amx_cg_solve_dev", self._h, self._vec(b, "b"), self._vec(x, "x"), float(rtol), int(max_it),
             C.byref(st), _stream_ptr(stream))
        return st

    def cg_noalloc(self, b, x, rtol=1e-10, max_it=10000, stream=None) -> _capi.SolveStats:
        """cg() with its workspace in the library's fixed device array
        (amx_cg_solve_noalloc_dev): no device allocation; at most
        NOALLOC_MAX_DOFS DoFs; calls overlapping on one GPU race."""
        st = _capi.SolveStats()
        call("amx_cg_solve_noalloc_dev", self._h, self._vec(b, "b"), self._vec(x, "x"), float(rtol), int(max_it),
             C.byref(st), _stream_ptr(stream))
        return st

    def ir_cg(self, precision: PREC, b, x, rtol=1e-10, inner_rtol=1e-2, max_outer=50, max_inner=10000,
              stream=None) -> _capi.SolveStats:
        st = _capi.SolveStats()
        call("amx_ir_cg_solve_dev", self._h, int(precision), self._vec(b, "b"), self._vec(x, "x"), float(rtol),
             float(inner_rtol),
             int(max_outer), int(max_inner), C.byref(st), _stream_ptr(stream))
        return st


class MultiRHS:
    """Multi-RHS element products V_i = sum_e P_e^T K_e P_e U_i of several
    finalized operators in one persistent MFMA launch (FP32 or BF16 storage,
    FP32 accumulation).  U_i, V_i: float32 device tensors of n_total x nrhs."""

    TILE_SHAPES = [(rt, w, c) for rt in (32, 64, 128) for w in (1, 2, 4) for c in (0, 1)]

    def __init__(self, ops: Sequence[aMatFree], precision: PREC = PREC.FP32):
        arr = (C.c_void_p * len(ops))(*[o.handle.value for o in ops])
        h = C.c_void_p()
        call("amx_mrhs_create", arr, len(ops), int(precision), C.byref(h))
        self._h = h
        self.ops = list(ops)  # keeps the operators alive
        self.precision = PREC(precision)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().amx_mrhs_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @staticmethod
    def _lists(U, V):
        assert len(U) == len(V)
        return (C.c_void_p * len(U))(*[_ptr(u) for u in U]), (C.c_void_p * len(V))(*[_ptr(v) for v in V])

    def set_tiles(self, rhs_tile: int, waves_per_block: int, work_split: int = 1):
        call("amx_mrhs_set_tiles", self._h, int(rhs_tile), int(waves_per_block), int(work_split))

    def set_scatter(self, mode: SCATTER):
        call("amx_mrhs_set_scatter", self._h, int(mode))

    def tiles(self) -> tuple[int, int, int]:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        call("amx_mrhs_get_tiles", self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def apply_dev(self, nrhs: int, U, V, stream=None):
        pu, pv = self._lists(U, V)
        call("amx_mrhs_apply_dev", self._h, int(nrhs), pu, pv, _stream_ptr(stream))

    def autotune(self, nrhs: int, U, V, reps: int = 5, stream=None) -> dict:
        """Times every launch shape and keeps the fastest; returns {(rhs_tile, waves, split): ms}."""
        pu, pv = self._lists(U, V)
        res = (C.c_double * len(self.TILE_SHAPES))()
        call("amx_mrhs_autotune", self._h, int(nrhs), pu, pv, int(reps), res, _stream_ptr(stream))
        return {shape: res[i] for i, shape in enumerate(self.TILE_SHAPES) if res[i] < 1e29}

    def graph_capture(self, nrhs: int, U, V, stream):
        pu, pv = self._lists(U, V)
        call("amx_mrhs_graph_capture", self._h, int(nrhs), pu, pv, _stream_ptr(stream))

    def graph_launch(self, stream):
        call("amx_mrhs_graph_launch", self._h, _stream_ptr(stream))


def device_count() -> int:
    return int(_capi.lib().amx_device_count())


def version() -> str:
    return _capi.lib().amx_version().decode()


__all__ = ["AmatError", "BC_METH", "BCOP", "Error", "GHOST_TO_OWNER", "HALO", "MATFREE_TYPE", "Maps", "MultiRHS",
           "OWNER_TO_GHOST", "PREC", "SCATTER", "aMatFree", "device_count", "version"]


def flatten_rows(rows: Sequence) -> tuple[np.ndarray, np.ndarray]:
    return _flatten_rows(rows, np.uint32)
