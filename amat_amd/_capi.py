"""ctypes binding of the C ABI in include/amat_amd.h.

The library is loaded from this package directory (amat_amd/libamat_amd.so,
built in-tree by ``python -m amat_amd.build``).  There is no fallback: if the
library is missing, importing the package's GPU classes raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libamat_amd.so"

# status codes (amx_status); values below 100 mirror par::Error
SUCCESS = 0
STATUS_NAMES = {
    0: "SUCCESS", 1: "INDEX_OUT_OF_BOUNDS", 2: "UNKNOWN_ELEMENT_TYPE", 3: "UNKNOWN_ELEMENT_STATUS",
    4: "NULL_L2G_MAP", 5: "GHOST_NODE_NOT_FOUND", 6: "UNKNOWN_MAT_TYPE", 7: "WRONG_COMMUNICATION",
    8: "UNKNOWN_DOF_TYPE", 9: "UNKNOWN_CONSTRAINT", 10: "UNKNOWN_BC_METH", 11: "NOT_IMPLEMENTED",
    12: "GLOBAL_DOF_ID_NOT_FOUND", 100: "INVALID_ARGUMENT", 101: "DEVICE_ERROR", 102: "OUT_OF_MEMORY",
    103: "INVALID_STATE", 104: "NOT_CONVERGED", 105: "REJECTED",
}


class MapsInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "rank", "size", "n_elems", "n_owned", "n_pre_ghost", "n_post_ghost", "n_total", "n_entries",
        "max_dofs_per_elem", "n_independent", "n_dependent", "n_owned_constraints", "n_owned_free",
        "n_send", "n_recv", "has_bdr_map", "has_peer_lists")] + [
        ("owned_begin", C.c_uint64), ("owned_end", C.c_uint64), ("n_global", C.c_uint64)]


class GmshInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "rank", "size", "dofs_per_node", "n_owned_elements", "n_owned_nodes", "n_pre_ghost_nodes",
        "n_post_ghost_nodes", "n_local_nodes", "n_entries", "max_nodes_per_element", "n_boundary_nodes",
        "n_partitions")] + [(n, C.c_uint64) for n in (
            "n_nodes", "n_file_nodes", "n_volume_elements", "n_surface_elements", "n_line_elements",
            "n_point_elements", "node_offset", "n_dofs_total", "owned_dof_begin", "owned_dof_end")]


# displacement(xyz, values) callback of amx_gmsh_get_prescribed_values
DISPLACEMENT_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double))


# fn(ctx, direction, send, recv, stream) of amx_op_set_halo_transport
HALO_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class ServerInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("active_jobs", "held_jobs", "free_slots", "pad")] + [
        (n, C.c_uint64) for n in ("workspace_bytes", "steps", "iterations", "completed")]


class OpInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("n_elems", "n_total", "n_owned", "n_pre_ghost", "n_groups", "finalized")] + [
        (n, C.c_int32) for n in ("matfree_type", "bc_meth", "bc_operator", "scatter_mode", "precision_mask")] + [
        ("trace", C.c_double), ("penalty", C.c_double)] + [
        (n, C.c_uint64) for n in ("device_bytes", "k_bytes_fp64", "k_bytes_fp32", "k_bytes_bf16", "alg_bytes_fp64",
                                  "flops")] + [("packed_symmetric", C.c_uint32), ("near_symmetric", C.c_uint32)]


class SolveStats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("iterations", "outer_iterations", "converged", "precision")] + [
        ("rel_residual", C.c_double), ("seconds", C.c_double)]


P = C.c_void_p
U32, U64, I32, F64 = C.c_uint32, C.c_uint64, C.c_int, C.c_double

# name -> (restype, argtypes); every function of include/amat_amd.h
SIGNATURES = {
    "amx_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "amx_version": (C.c_char_p, []),
    "amx_device_count": (I32, []),
    "amx_maps_create": (I32, [U32, U32, C.POINTER(P)]),
    "amx_maps_destroy": (I32, [P]),
    "amx_maps_set_map": (I32, [P, U32, P, P, U32, P, U64, U64, U64, P]),
    "amx_maps_get_ghost_requests": (I32, [P, P, P]),
    "amx_maps_set_peer_requests": (I32, [P, P, P]),
    "amx_maps_set_bdr_map": (I32, [P, P, P, U32]),
    "amx_maps_get_info": (I32, [P, C.POINTER(MapsInfo)]),
    "amx_maps_get_local_map": (I32, [P, P]),
    "amx_maps_get_global_map": (I32, [P, P]),
    "amx_maps_get_local2global": (I32, [P, P]),
    "amx_maps_get_localid2rank": (I32, [P, P]),
    "amx_maps_get_bdr_map": (I32, [P, P, P]),
    "amx_maps_get_owned_constraints": (I32, [P, P, P]),
    "amx_maps_get_owned_free": (I32, [P, P]),
    "amx_maps_get_element_lists": (I32, [P, P, P]),
    "amx_maps_get_send_lists": (I32, [P, P, P, P]),
    "amx_maps_get_recv_lists": (I32, [P, P, P]),
    "amx_maps_get_dof_scan": (I32, [P, P]),
    "amx_op_create": (I32, [P, I32, I32, C.POINTER(P)]),
    "amx_op_destroy": (I32, [P]),
    "amx_op_set_matfree_type": (I32, [P, I32]),
    "amx_op_set_num_streams": (I32, [P, U32]),
    "amx_op_set_element_matrix": (I32, [P, U32, P, I32, U32, U32, U32]),
    "amx_op_set_element_matrices": (I32, [P, U32, U32, P, I32]),
    "amx_op_set_element_matrices_scaled": (I32, [P, P, I32, P]),
    "amx_op_set_element_function": (I32, [P, I32, P, P, F64, F64, U32]),
    "amx_measure_read_bandwidth": (I32, [I32, U64, I32, C.POINTER(F64)]),
    "amx_op_set_options": (I32, [P, I32, I32]),
    "amx_op_set_penalty": (I32, [P, F64]),
    "amx_op_set_symmetric_storage": (I32, [P, I32]),
    "amx_op_set_bc_operator": (I32, [P, I32]),
    "amx_op_finalize": (I32, [P]),
    "amx_op_get_info": (I32, [P, C.POINTER(OpInfo)]),
    "amx_op_set_halo_transport": (I32, [P, HALO_FN, P, I32]),
    "amx_op_matvec": (I32, [P, P, P, I32]),
    "amx_op_matmult_mf": (I32, [P, P, P]),
    "amx_op_matvec_dev": (I32, [P, P, P, P]),
    "amx_op_matmult_dev": (I32, [P, P, P, P]),
    "amx_op_matvec_begin_dev": (I32, [P, P, P, I32, P]),
    "amx_op_matvec_end_dev": (I32, [P, P, P, I32, P]),
    "amx_op_halo_pack_dev": (I32, [P, P, P, P]),
    "amx_op_halo_import_dev": (I32, [P, P, P, P]),
    "amx_op_halo_export_dev": (I32, [P, P, P, P]),
    "amx_op_halo_accumulate_dev": (I32, [P, P, P, P]),
    "amx_op_bc_fixup_dev": (I32, [P, P, P, P]),
    "amx_op_matmult_f32_dev": (I32, [P, I32, P, P, P]),
    "amx_op_get_diagonal": (I32, [P, P]),
    "amx_op_get_diagonal_dev": (I32, [P, P, P]),
    "amx_op_apply_bc_rhs": (I32, [P, P]),
    "amx_op_get_timer": (I32, [P, P, P, P, P]),
    "amx_op_time_kernels_dev": (I32, [P, I32, P, P, P, P, P, P]),
    "amx_op_stream": (P, [P]),
    "amx_op_synchronize": (I32, [P]),
    "amx_cg_solve_dev": (I32, [P, P, P, F64, I32, C.POINTER(SolveStats), P]),
    "amx_cg_solve_noalloc_dev": (I32, [P, P, P, F64, I32, C.POINTER(SolveStats), P]),
    "amx_dev_dot2": (I32, [P, P, U64, U32, U32, P, I32, P]),
    "amx_ir_cg_solve_dev": (I32, [P, I32, P, P, F64, F64, I32, I32, C.POINTER(SolveStats), P]),
    "amx_mrhs_create": (I32, [P, I32, I32, P]),
    "amx_mrhs_destroy": (I32, [P]),
    "amx_mrhs_set_tiles": (I32, [P, I32, I32, I32]),
    "amx_mrhs_get_tiles": (I32, [P, P, P, P]),
    "amx_mrhs_set_scatter": (I32, [P, I32]),
    "amx_mrhs_apply_dev": (I32, [P, I32, P, P, P]),
    "amx_mrhs_autotune": (I32, [P, I32, P, P, I32, P, P]),
    "amx_mrhs_graph_capture": (I32, [P, I32, P, P, P]),
    "amx_mrhs_graph_launch": (I32, [P, P]),
    "amx_server_create": (I32, [I32, I32, P]),
    "amx_server_destroy": (I32, [P]),
    "amx_server_job_bytes": (I32, [P, I32, P]),
    "amx_server_submit": (I32, [P, P, I32, P, F64, F64, I32, I32, P]),
    "amx_server_step": (I32, [P, I32, P]),
    "amx_server_poll": (I32, [P, P, I32, P]),
    "amx_server_result": (I32, [P, C.c_int64, P, P]),
    "amx_server_get_info": (I32, [P, P]),
    "amx_fem_gauss": (I32, [U32, P]),
    "amx_fem_ke_hex8": (I32, [P, U32, P]),
    "amx_fem_ke_hex20_iso": (I32, [P, F64, F64, U32, P]),
    "amx_fem_ke_tet10": (I32, [P, U32, P]),
    "amx_fem_fe_tet10": (I32, [P, P, U32, P]),
    "amx_fem_ke_tet10_batch": (I32, [U32, P, U32, P]),
    "amx_fem_fe_tet10_batch": (I32, [U32, P, P, U32, P]),
    "amx_gmsh_read": (I32, [C.c_char_p, U32, U32, U32, C.POINTER(P)]),
    "amx_gmsh_destroy": (I32, [P]),
    "amx_gmsh_get_info": (I32, [P, C.POINTER(GmshInfo)]),
    "amx_gmsh_get_element_sizes": (I32, [P, P, P, P]),
    "amx_gmsh_get_maps": (I32, [P, P, P, P, P]),
    "amx_gmsh_get_local2global": (I32, [P, P, P]),
    "amx_gmsh_get_rank_node_counts": (I32, [P, P]),
    "amx_gmsh_get_coordinates": (I32, [P, P]),
    "amx_gmsh_get_constraints": (I32, [P, P, P]),
    "amx_gmsh_get_prescribed_values": (I32, [P, DISPLACEMENT_FN, P]),
}

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """The loaded C library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m amat_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class AmatError(RuntimeError):
    """A non-zero amx_status; .code is the status, .name its par::Error name."""

    def __init__(self, code: int, message: str):
        self.code = int(code)
        self.name = STATUS_NAMES.get(self.code, str(code))
        super().__init__(f"[{self.name}] {message}")


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().amx_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code: int) -> None:
    if code != SUCCESS:
        raise AmatError(code, last_error())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
