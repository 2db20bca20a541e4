/*
 * amat_amd.h -- C ABI of the MI355X-native assembly-free element-by-element
 * (EbE) operator  v = sum_e P_e^T K_e P_e u.
 *
 * This is the drop-in boundary for aMat's GPU methods (MATFREE_TYPE 3/4/5,
 * include/enums.hpp:56 of paralab/aMat).  Every entry point names the
 * reference interface it replaces (file:line in the paralab/aMat tree).  Plain
 * pointers and sizes only; no C++ or torch types cross this boundary.
 *
 * Types follow the reference instantiation <DT,GI,LI> =
 * <double, unsigned long, unsigned int> (examples/src/example03a.cpp:423-426):
 * global ids are uint64_t, local ids uint32_t, values double.
 *
 * Ownership: every set_* call copies its inputs; no pointer passed in is kept
 * after the call returns.  Device buffers belong to the handle.
 * Threading: a handle is used by one host thread at a time; distinct handles
 * are independent.  Device entry points (*_dev) are asynchronous on the given
 * HIP stream; NULL is the HIP default (null) stream, as everywhere in HIP.
 * amx_op_stream() returns the handle's own stream for callers that want it.
 * Host-array entry points run on the handle's stream and return synchronized.
 * Errors: every call returns an int status (AMX_SUCCESS = 0); the values below
 * AMX_INVALID_ARGUMENT mirror par::Error (include/enums.hpp:24-37).  The
 * message of the last failure on the calling thread is in amx_last_error().
 * Nothing here calls exit().
 */
#ifndef AMAT_AMD_H
#define AMAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
/* par::Error, include/enums.hpp:24-37 */
enum amx_status {
    AMX_SUCCESS = 0,
    AMX_INDEX_OUT_OF_BOUNDS = 1,
    AMX_UNKNOWN_ELEMENT_TYPE = 2,
    AMX_UNKNOWN_ELEMENT_STATUS = 3,
    AMX_NULL_L2G_MAP = 4,
    AMX_GHOST_NODE_NOT_FOUND = 5,
    AMX_UNKNOWN_MAT_TYPE = 6,
    AMX_WRONG_COMMUNICATION = 7,
    AMX_UNKNOWN_DOF_TYPE = 8,
    AMX_UNKNOWN_CONSTRAINT = 9,
    AMX_UNKNOWN_BC_METH = 10,
    AMX_NOT_IMPLEMENTED = 11,
    AMX_GLOBAL_DOF_ID_NOT_FOUND = 12,
    /* additions of this library */
    AMX_INVALID_ARGUMENT = 100,
    AMX_DEVICE_ERROR = 101,
    AMX_OUT_OF_MEMORY = 102,
    AMX_INVALID_STATE = 103,
    AMX_NOT_CONVERGED = 104,
    AMX_REJECTED = 105
};

/* par::BC_METH, include/enums.hpp:20-22 */
enum amx_bc_meth { AMX_BC_IMATRIX = 0, AMX_BC_PENALTY = 1 };

/* par::MATFREE_TYPE, include/enums.hpp:56.  This library runs 2 (FREE, with
 * the built-in element kinds of amx_op_set_element_function) and 3/4/5 on the
 * GPU; 1 is the reference's CPU loop and returns AMX_NOT_IMPLEMENTED. */
enum amx_matfree_type {
    AMX_MATFREE_UNDEFINED = 0,
    AMX_MATFREE_QUASI_FREE = 1,
    AMX_MATFREE_FREE = 2,
    AMX_MATFREE_GPU_PURE = 3,
    AMX_MATFREE_GPU_OVER_CPU = 4,
    AMX_MATFREE_GPU_OVER_GPU = 5
};

/* element kinds computed in-kernel by method 2 (FREE): the element matrices
 * of the examples, ke_hex8 / ke_hex20_iso / ke_tet10
 * (examples/src/ke_matrix.cpp:436-526, 1573-1740, 329-433) */
enum amx_element_kind {
    AMX_ELEM_HEX8_LAPLACE = 1,   /* 8 nodes, 1 DoF per node */
    AMX_ELEM_HEX20_ELASTIC = 2,  /* 20 nodes, 3 DoFs per node (DoF 3a+i), isotropic */
    AMX_ELEM_TET10_LAPLACE = 3   /* 10 nodes, 1 DoF per node */
};

/* storage precision of the element matrices (bit mask, amx_op_set_options) */
enum amx_precision { AMX_PREC_FP64 = 1, AMX_PREC_FP32 = 2, AMX_PREC_BF16 = 4 };
/* refinement options (or-ed into the precision argument of the refinement
 * entry points): AMX_IR_F64_VECTORS, inner CG vectors in FP64 instead of
 * FP32; AMX_IR_RELIABLE, one low-precision CG recurrence never restarted,
 * whose residual is replaced by the FP64 true residual b - A x (and its
 * increment flushed into the FP64 solution) each time it has fallen by
 * inner_rtol (then the replacement threshold delta) since the last
 * replacement; max_outer bounds the replacements, max_inner the iterations */
enum amx_ir_flags { AMX_IR_F64_VECTORS = 0x100, AMX_IR_RELIABLE = 0x200 };

/* how element vectors are summed into v */
enum amx_scatter {
    AMX_SCATTER_DETERMINISTIC = 0, /* node -> (element, local id) transposed gather, fixed order */
    AMX_SCATTER_ATOMIC = 1         /* FP64/FP32 hardware atomics, order not fixed */
};

/* boundary-condition form applied by amx_op_matmult_* */
enum amx_bc_operator {
    AMX_BCOP_NONE = 0,        /* plain K u */
    AMX_BCOP_IDENTITY = 1,    /* aMatFree::MatMult_mf identity rows, include/aMatFree.hpp:3893-3941 */
    AMX_BCOP_PENALTY_DIAG = 2 /* aMatBased::apply_bc_mat penalty: K_cc := p, include/aMatBased.hpp:332-347 */
};

/* element-matrix storage order of the caller's arrays */
enum amx_layout { AMX_ROW_MAJOR = 0, AMX_COL_MAJOR = 1 };

/* Message of the last failed call on this thread ("" if none).  Copies at most
 * len-1 bytes plus a terminating NUL; returns the full message length. */
size_t amx_last_error(char* buf, size_t len);

/* Library version string, e.g. "amat_amd 0.1.0 gfx950". */
const char* amx_version(void);

/* Number of visible HIP devices (0 when none); never fails. */
int amx_device_count(void);

/* Measurement aid, no counterpart in the reference: the achievable HBM read
 * bandwidth of `device` (16-byte streaming loads over a `bytes` buffer, `iters`
 * timed passes, hipEvents), the measured peak beside the nominal one. */
int amx_measure_read_bandwidth(int device, uint64_t bytes, int iters, double* bytes_per_s);

/* ------------------------------------------------------------------ maps
 * par::Maps<DT,GI,LI> (include/maps.hpp).  One handle per rank.  In the
 * reference the counts of all ranks come from MPI_Allgather (maps.hpp:547) and
 * the ghost-id exchange is an MPI_Alltoallv (maps.hpp:738); here the caller
 * brings the counts (rank_dof_counts) and moves the request lists with its own
 * transport (amx_maps_get_ghost_requests / amx_maps_set_peer_requests).  With
 * one rank no exchange is needed. */
typedef struct amx_maps amx_maps;

typedef struct amx_maps_info {
    uint32_t rank, size;
    uint32_t n_elems;         /* Maps::get_NumElems */
    uint32_t n_owned;         /* Maps::get_NumDofs */
    uint32_t n_pre_ghost;     /* Maps::get_NumPreGhostDofs */
    uint32_t n_post_ghost;    /* Maps::get_NumPostGhostDofs */
    uint32_t n_total;         /* Maps::get_NumDofsTotal */
    uint32_t n_entries;       /* sum of dofs_per_element */
    uint32_t max_dofs_per_elem;
    uint32_t n_independent, n_dependent;
    uint32_t n_owned_constraints, n_owned_free;
    uint32_t n_send, n_recv;  /* lengths of the send/recv id lists */
    uint32_t has_bdr_map, has_peer_lists;
    uint64_t owned_begin, owned_end, n_global;
} amx_maps_info;

/* Maps::Maps(MPI_Comm) (include/maps.hpp:80) */
int amx_maps_create(uint32_t rank, uint32_t size, amx_maps** out);
int amx_maps_destroy(amx_maps* maps);

/* Maps::set_map (include/maps.hpp:398-451) + buildScatterMap (:519-826).
 * element_to_rank_map: the LI** map flattened element after element
 * (element e holds dofs_per_element[e] entries).  owned range inclusive.
 * rank_dof_counts[size]: owned DoF count of every rank (NULL when size == 1). */
int amx_maps_set_map(amx_maps* maps, uint32_t n_elements_on_rank,
                     const uint32_t* element_to_rank_map, const uint32_t* dofs_per_element,
                     uint32_t n_all_dofs_on_rank, const uint64_t* rank_to_global_map,
                     uint64_t owned_global_dof_range_begin, uint64_t owned_global_dof_range_end,
                     uint64_t n_global_dofs, const uint32_t* rank_dof_counts);

/* sendBuf of buildScatterMap's Alltoallv (include/maps.hpp:716-739): ghost
 * global ids grouped by owner rank.  counts[size]; gids[n_pre+n_post]. */
int amx_maps_get_ghost_requests(const amx_maps* maps, uint32_t* counts, uint64_t* gids);

/* recvBuf of that Alltoallv: the ids peers requested from this rank, grouped
 * by requesting rank (counts[size]).  Builds m_uivSendDofIds and the
 * send/recv counts and offsets (include/maps.hpp:741-785). */
int amx_maps_set_peer_requests(amx_maps* maps, const uint32_t* counts, const uint64_t* gids);

/* Maps::set_bdr_map (include/maps.hpp:828-880).  Same outputs as the
 * reference's linear scan, found through a hash of the constraint list (first
 * occurrence wins, as the reference's `break`). */
int amx_maps_set_bdr_map(amx_maps* maps, const uint64_t* constrained_dofs,
                         const double* prescribed_values, uint32_t n_constraints);

int amx_maps_get_info(const amx_maps* maps, amx_maps_info* info);
/* getters (include/maps.hpp:85-276); arrays are sized from amx_maps_info */
int amx_maps_get_local_map(const amx_maps* maps, uint32_t* out);      /* n_entries, get_LocalMap */
int amx_maps_get_global_map(const amx_maps* maps, uint64_t* out);     /* n_entries, get_Map */
int amx_maps_get_local2global(const amx_maps* maps, uint64_t* out);   /* n_total, get_Local2Global */
int amx_maps_get_localid2rank(const amx_maps* maps, uint32_t* out);   /* n_total, get_localID2Rank */
int amx_maps_get_bdr_map(const amx_maps* maps, uint32_t* mask, double* presval); /* n_entries each */
int amx_maps_get_owned_constraints(const amx_maps* maps, uint64_t* dofs, double* values);
int amx_maps_get_owned_free(const amx_maps* maps, uint64_t* dofs);
int amx_maps_get_element_lists(const amx_maps* maps, uint32_t* independent, uint32_t* dependent);
/* per-rank counts/offsets (size entries each) and id lists (n_send / n_recv) */
int amx_maps_get_send_lists(const amx_maps* maps, uint32_t* counts, uint32_t* offsets, uint32_t* ids);
int amx_maps_get_recv_lists(const amx_maps* maps, uint32_t* counts, uint32_t* offsets);
int amx_maps_get_dof_scan(const amx_maps* maps, uint64_t* scan);     /* size, get_LocalDofScan */

/* ---------------------------------------------------------------- operator
 * par::aMatFree<DT,GI,LI> (include/aMatFree.hpp:25-353) for methods 3/4/5,
 * replacing aMatGpu (include/aMatGpu.hpp:21-591). */
typedef struct amx_op amx_op;

typedef struct amx_op_info {
    uint32_t n_elems, n_total, n_owned, n_pre_ghost;
    uint32_t n_groups;          /* element groups by DoFs-per-element */
    uint32_t finalized;
    int32_t matfree_type, bc_meth, bc_operator, scatter_mode, precision_mask;
    double trace;               /* m_dtTraceK, include/aMatFree.hpp:748-785 */
    double penalty;             /* value placed on constrained diagonals */
    uint64_t device_bytes;      /* device memory held by the handle */
    uint64_t k_bytes_fp64, k_bytes_fp32, k_bytes_bf16; /* element-matrix storage */
    uint64_t alg_bytes_fp64;    /* E(b + 4n) + 16 N, b = FP64 bytes stored per element (8 n^2, 8 n(n+1)/2
                                   packed, the record size of near-symmetric records): one FP64 matvec's
                                   algorithmic HBM bytes */
    uint64_t flops;             /* 2 sum_e n_e^2 per matvec */
    uint32_t packed_symmetric;  /* element groups stored as packed symmetric tiles */
    uint32_t near_symmetric;    /* element groups stored as near-symmetric records (FP64) */
} amx_op_info;

/* aMatFree(Maps&, BC_METH) (include/aMatFree.hpp:405-439).  Copies the maps it
 * needs; `maps` may be destroyed afterwards.  device = HIP ordinal. */
int amx_op_create(const amx_maps* maps, int bc_meth, int device, amx_op** out);
int amx_op_destroy(amx_op* op);

/* aMat::set_matfree_type (include/aMat.hpp:184) */
int amx_op_set_matfree_type(amx_op* op, int matfree_type);
/* aMat::set_element_matrix_function (include/aMat.hpp:114-116) for method 2
 * (FREE, include/aMatFree.hpp:2240-2744): instead of a callback that fills
 * K_e from the element's coordinates (examples/src/example07a.cpp:54-70),
 * the kernel computes s_e K_e u_e from the coordinates at every matvec.
 * xe: n_elems x (3 * nodes) node coordinates, node-major in the element's DoF
 * order (x0,y0,z0,x1,...); scale: per-element factor s_e (NULL: 1); E, nu:
 * isotropic material (hex20); ngt: Gauss points per direction, as the
 * examples' NGT (hex20: ngt^3 <= 64).  Every element must have the kind's DoF
 * count.  Sets the matfree type to FREE; finalize() builds no element
 * matrices.  FP64, and an FP32 form (FP32 arithmetic and vectors, nothing
 * stored) for amx_op_matmult_f32_dev(op, AMX_PREC_FP32, ...), the
 * mixed-precision solvers and served FP32 jobs; no BF16 form; the multi-RHS
 * entry points reject a FREE operator. */
int amx_op_set_element_function(amx_op* op, int kind, const double* xe, const double* scale,
                                double E, double nu, uint32_t ngt);
/* aMatFree::set_num_streams (include/aMatFree.hpp:230-237): number of HIP
 * streams used by the two-phase (independent/dependent) launch; >= 1 */
int amx_op_set_num_streams(amx_op* op, uint32_t n_streams);

/* Storage of the element matrices (no counterpart in the reference, which
 * stores every matrix whole, include/aMatFree.hpp:882-953).  Mode 1 (default):
 * 8- and 10-DoF groups pack the upper triangles when every matrix of the group
 * is bitwise symmetric -- the kernels then read K[r][c] = K[min(r,c)][max(r,c)]:
 * the same values in the same order, about half the bytes; 60-DoF groups keep
 * their FP64 matrices as near-symmetric records when every matrix is symmetric
 * up to low-word round-off (upper triangle, one byte per lower entry coding
 * its low word against the transposed entry, full values for at most a few
 * hundred positions of the group that the byte cannot code): every entry is
 * rebuilt bit for bit, about 0.6 of the bytes.  Mode 0 keeps whole matrices.
 * Takes effect at the next finalize. */
int amx_op_set_symmetric_storage(amx_op* op, int mode);

/* aMatFree::set_element_matrix -> copy_element_matrix
 * (include/aMatFree.hpp:854-953): block (block_i, block_j) of element eid,
 * n_b x n_b with n_b = dofs_per_element[eid] / blocks_dim, in `layout` order. */
int amx_op_set_element_matrix(amx_op* op, uint32_t eid, const double* e_mat, int layout,
                              uint32_t block_i, uint32_t block_j, uint32_t blocks_dim);
/* Bulk form of the above for count consecutive elements with equal DoF
 * counts (one n x n matrix each, back to back). */
int amx_op_set_element_matrices(amx_op* op, uint32_t first_eid, uint32_t count,
                                const double* e_mats, int layout);
/* K_e = scale[e] * k_ref for every element (all elements share n): the
 * congruent structured meshes of the examples with per-element stiffness. */
int amx_op_set_element_matrices_scaled(amx_op* op, const double* k_ref, int layout,
                                       const double* scale);

/* precision_mask: AMX_PREC_* storages to build at finalize (FP64 always kept);
 * scatter_mode: AMX_SCATTER_*. */
int amx_op_set_options(amx_op* op, int precision_mask, int scatter_mode);
/* value placed on constrained diagonals by AMX_BCOP_PENALTY_DIAG; by default
 * PENALTY_FACTOR (100, include/aMat.hpp:68) times the trace. */
int amx_op_set_penalty(amx_op* op, double penalty);
/* which BC form amx_op_matmult_* applies (default AMX_BCOP_IDENTITY, what
 * aMatFree::MatMult_mf does whatever the BC method). */
int amx_op_set_bc_operator(amx_op* op, int bc_operator);

/* aMatFree::finalize (include/aMatFree.hpp:151-155, 656-785): uploads K in
 * the device layout, builds the transposed gather lists, BC tables, trace
 * (element order) and the assembled diagonal. */
int amx_op_finalize(amx_op* op);
int amx_op_get_info(const amx_op* op, amx_op_info* info);

/* ------------------------------------------------------- halo transport
 * The reference's matvec on a rank's share is bracketed by MPI point-to-point
 * exchanges through the maps' send/recv lists: ghost_receive_begin/end
 * (include/aMatFree.hpp:1426-1549) copies owners' u into the ghost slots
 * before the dependent elements, ghost_send_begin/end (:1552-1679) adds the
 * ghost slots of v into their owners afterwards.  This library holds no
 * communicator: the caller registers the transport that moves the buffers
 * (MPI, RCCL, or threads sharing one process), and the library packs,
 * imports and accumulates around it.
 *
 * fn(ctx, direction, send, recv, stream) is called once per exchange with
 *   AMX_HALO_OWNER_TO_GHOST: send[n_send] = u at the send ids (peer i's part
 *     at send offset i, count i of amx_maps_get_send_lists); recv[n_pre +
 *     n_post] receives the ghost values, peer i's part at recv offset i,
 *     count i (amx_maps_get_recv_lists) -- ghost_receive_begin's layout;
 *   AMX_HALO_GHOST_TO_OWNER: send[n_pre + n_post] = the ghost slots of v
 *     (pre then post), peer i's part at recv offset i; recv[n_send] receives
 *     the owners' contributions, peer i's part at send offset i --
 *     ghost_send_begin's layout.
 * mode AMX_HALO_HOST: send/recv are pinned host buffers; the library has
 * synchronized its stream before the call and copies recv back after it (an
 * MPI caller, as the reference).  AMX_HALO_DEVICE: device buffers; the
 * transfers must be complete, or ordered on `stream` (an RCCL caller), when
 * fn returns.  fn returns 0 on success; anything else fails the calling
 * entry point with AMX_WRONG_COMMUNICATION.  Every rank must make the same
 * sequence of exchanging calls, as with the reference's MPI calls. */
enum amx_halo_direction { AMX_HALO_OWNER_TO_GHOST = 0, AMX_HALO_GHOST_TO_OWNER = 1 };
enum amx_halo_mode { AMX_HALO_HOST = 0, AMX_HALO_DEVICE = 1 };
typedef int (*amx_halo_fn)(void* ctx, int direction, const double* send, double* recv, void* stream);

/* Registers (fn != NULL) or removes (fn == NULL) the transport of a handle
 * whose maps have more than one rank and peer lists.  The first exchange with
 * the handle's finalized state adds the ghost rows of the element diagonals
 * into their owners (mat_get_diagonal_ghosted's ghost_send, :1051-1053), so
 * that the diagonal, the penalty coefficients p - K_cc and the Jacobi
 * diagonal are the assembled ones: this happens inside finalize when a
 * transport is registered, else inside this call -- on every rank. */
int amx_op_set_halo_transport(amx_op* op, amx_halo_fn fn, void* ctx, int mode);

/* aMatFree::matvec (include/aMatFree.hpp:1681-1775): v = K u, no BCs, host
 * arrays.  is_ghosted != 0: length n_total; else n_owned (local_to_ghost /
 * ghost_to_local, :824-852: ghost slots start at 0).  More than one rank:
 * needs the halo transport (AMX_NOT_IMPLEMENTED without one); the ghost slots
 * of u are filled from their owners and owners' rows receive the ghost
 * contributions, as the reference's exchanges do. */
int amx_op_matvec(amx_op* op, double* v, const double* u, int is_ghosted);

/* aMatFree::MatMult_mf array semantics (include/aMatFree.hpp:3893-3941) with
 * the configured BC form, owned-length host arrays.  More than one rank: with
 * the halo transport, as amx_op_matvec. */
int amx_op_matmult_mf(amx_op* op, double* vv, const double* uu);

/* Device forms on ghosted-length device arrays, asynchronous on `stream`.
 * matvec_dev: v = K u.  matmult_dev: the configured BC form.  One rank:
 * ghost slots are empty.  More than one rank: the whole exchange runs through
 * the halo transport (the ghost slots of u are overwritten with the owners'
 * values, as ghost_receive_end does); without one they return
 * AMX_NOT_IMPLEMENTED -- the caller then drives the pieces itself
 * (amx_op_matvec_begin_dev and the halo entry points below). */
int amx_op_matvec_dev(amx_op* op, double* v, const double* u, void* stream);
int amx_op_matmult_dev(amx_op* op, double* v, const double* u, void* stream);

/* Two-phase form of matvec_ghosted_gpuOverGpu (include/aMatFree.hpp:3858-3889):
 * begin launches the independent elements (they read owned u only); the caller
 * then fills u's ghost slots (amx_op_halo_import_dev); end launches the
 * dependent elements and the gather (atomic scatter: begin zero-fills v and
 * both add into it; no gather).  bc_operator as amx_op_matmult_dev when
 * apply_bc != 0; ghost slots of v then hold contributions for their owners
 * (amx_op_halo_export_dev / amx_op_halo_accumulate_dev), after which
 * amx_op_bc_fixup_dev applies the constrained rows. */
int amx_op_matvec_begin_dev(amx_op* op, double* v, const double* u, int apply_bc, void* stream);
int amx_op_matvec_end_dev(amx_op* op, double* v, const double* u, int apply_bc, void* stream);
/* ghost_receive_begin's pack (include/aMatFree.hpp:1452-1454):
 * send[i] = u[send_ids[i]], n_send entries (what the owner sends) */
int amx_op_halo_pack_dev(amx_op* op, const double* u, double* send, void* stream);
/* ghost_receive_end memcpy (include/aMatFree.hpp:1536-1539): halo[n_pre+n_post]
 * (pre then post) -> ghost slots of u.  Zeroes constrained ghosts when the
 * identity BC form is active, as the owner's MatMult_mf does. */
int amx_op_halo_import_dev(amx_op* op, double* u, const double* halo, void* stream);
/* ghost slots of v -> halo[n_pre+n_post] (what ghost_send_begin sends) */
int amx_op_halo_export_dev(amx_op* op, const double* v, double* halo, void* stream);
/* ghost_send_end accumulate (include/aMatFree.hpp:1662-1669):
 * v[send_ids[i]] += recv[i] in the reference's loop order (deterministic). */
int amx_op_halo_accumulate_dev(amx_op* op, double* v, const double* recv, void* stream);
/* constrained owned rows after accumulation (MatMult_mf :3929-3933 or penalty) */
int amx_op_bc_fixup_dev(amx_op* op, double* v, const double* u, void* stream);

/* Mixed-precision forms: K in FP32 or BF16 (must have been built at
 * finalize), FP32 vectors, FP32 accumulation, configured BC form. */
int amx_op_matmult_f32_dev(amx_op* op, int precision, float* v, const float* u, void* stream);

/* aMatFree::MatGetDiagonal_mf (include/aMatFree.hpp:3944-3974):
 * mat_get_diagonal_ghosted (:997-1056) + apply_bc_diagonal (:4309-4337).
 * Host array, owned length.  _dev: device array, ghosted length.  More than
 * one rank: needs the halo transport (the owned rows then hold the assembled
 * diagonal). */
int amx_op_get_diagonal(amx_op* op, double* diag);
int amx_op_get_diagonal_dev(amx_op* op, double* diag, void* stream);

/* aMatFree::apply_bc_rhs (include/aMatFree.hpp:4536-4700): rhs (owned length,
 * host) -> rhs with constrained rows set (prescribed value, or penalty *
 * prescribed value) and, for BC_IMATRIX, free rows reduced by K_fc U_c.
 * More than one rank with BC_IMATRIX: K_fc U_c of the ghost rows reaches
 * their owners through the halo transport (the reference's ADD_VALUES into
 * a PETSc vector), which must be registered. */
int amx_op_apply_bc_rhs(amx_op* op, double* rhs);

/* aMatFree::get_timer (include/aMatFree.hpp:239-245): seconds accumulated
 * since finalize by the host-array forms (amx_op_matvec, amx_op_matmult_mf),
 * in the reference's argument order.  scatter = u copied to the device (the
 * reference: pack + host-to-device), mv = the device application (element
 * kernels, gather, BC rows; the reference: the batched GEMV), gather = v
 * copied back (the reference: device-to-host + host accumulation), total =
 * the whole calls on the host clock.  Device-array forms are not counted
 * (amx_op_time_kernels_dev times those). */
int amx_op_get_timer(amx_op* op, double* scatter_s, double* gather_s, double* mv_s, double* total_s);

/* Kernel timers, the GPU counterpart of aMatFree::get_timer
 * (include/aMatFree.hpp:239-245, mv_time / gather_time): runs n_iter
 * applications of the configured BC form on `stream` with hipEvents around
 * the element-kernel launches and around the gather launch; returns the mean
 * milliseconds of each and of the whole application.  Synchronizes. */
int amx_op_time_kernels_dev(amx_op* op, int n_iter, double* v, const double* u, double* ms_element,
                            double* ms_gather, double* ms_total, void* stream);

/* Stream the handle launches on (hipStream_t as void*). */
void* amx_op_stream(amx_op* op);
int amx_op_synchronize(amx_op* op);

/* --------------------------------------------------------------- solvers
 * Device Krylov solvers around the operator (the KSPCG + PCJACOBI of
 * examples/include/solve.hpp:45-65, on the GPU).  b, x: device arrays of
 * length n_total (one rank).  Stop: ||b - A x||_2 <= rtol ||b||_2. */
typedef struct amx_solve_stats {
    int32_t iterations;         /* CG iterations (inner ones for IR) */
    int32_t outer_iterations;   /* refinement steps (IR); FP64 CG: extra residual replacements */
    int32_t converged;
    int32_t precision;          /* AMX_PREC_* of the inner operator */
    double rel_residual;        /* true ||b - A x|| / ||b|| in FP64 at exit */
    double seconds;
} amx_solve_stats;

/* FP64 Jacobi-PCG.  When the recurrence stops, its residual is replaced by
 * the true b - A x; convergence is decided on that, and CG continues the
 * direction from it while it misses rtol (at most 4 replacements). */
int amx_cg_solve_dev(amx_op* op, const double* b, double* x, double rtol, int max_it,
                     amx_solve_stats* stats, void* stream);
/* No counterpart in the reference: amx_cg_solve_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arithmetic and results; the CG
 * vectors, reduction partials and scalars live in a fixed device array that
 * each GPU holds until the process ends, so no call allocates.  Operators of
 * more than AMX_NOALLOC_MAX_DOFS DoFs are rejected (AMX_INVALID_ARGUMENT).
 * All calls on one GPU share that array: calls that overlap on one GPU race. */
#define AMX_NOALLOC_MAX_DOFS 1048576u
int amx_cg_solve_noalloc_dev(amx_op* op, const double* b, double* x, double rtol, int max_it,
                             amx_solve_stats* stats, void* stream);
/* FP64 iterative refinement: true residual, alpha/beta, inner products and the
 * penalty diagonal in FP64; inner Jacobi-PCG on the FP32 or BF16 operator to
 * inner_rtol.  precision = AMX_PREC_FP32 or AMX_PREC_BF16, optionally or-ed
 * with amx_ir_flags (AMX_IR_RELIABLE: the reliable-update form, see above). */
int amx_ir_cg_solve_dev(amx_op* op, int precision, const double* b, double* x, double rtol,
                        double inner_rtol, int max_outer, int max_inner,
                        amx_solve_stats* stats, void* stream);

/* Self-check of the solvers' deterministic inner products (no counterpart in
 * the reference).  out (device, 2 * repeat doubles): out[2k] = x.y and
 * out[2k+1] = y.y in FP64, computed `repeat` times in a row by the reduction
 * every CG kernel uses, over n_blocks (1..1024) blocks of 256 threads: thread
 * t of block b sums elements i = b*256 + t, then i + n_blocks*256, ... in
 * ascending order; a tree over each block's 256 sums (stride 128, 64, ... 1);
 * the blocks in groups of 32 (the same tree over a group's sums, zero-padded);
 * the same tree over the group sums.  x, y: device arrays of n FP64 values on
 * `device`.  Synchronizes `stream`. */
int amx_dev_dot2(const double* x, const double* y, uint64_t n, uint32_t n_blocks, uint32_t repeat,
                 double* out, int device, void* stream);

/* ------------------------------------------------ multi-RHS element products
 * V_i = sum_e P_e^T K_e P_e U_i for several operators i in ONE persistent
 * launch, each U_i / V_i a device array of n_total x nrhs floats (DoF-major,
 * the nrhs values of a DoF contiguous).  The reference applies one vector at a
 * time (aMatFree::matvec, include/aMatFree.hpp:1681-1775; its GPU methods call
 * a batched GEMV per colour, include/aMatGpu.hpp:21-591); this is the same
 * product for nrhs vectors at once, on MFMA: FP32 storage with FP32 MFMA, or
 * BF16 storage and BF16-rounded U with FP32 accumulation.  No boundary
 * conditions (the plain element sum of aMatFree::matvec).  The element rows are
 * summed into V with FP32 atomics (default, order not fixed) or
 * deterministically (rows stored, then summed per DoF in a fixed order:
 * bit-reproducible, independent of the launch shape), amx_mrhs_set_scatter.
 * Element types: 8, 10, 20 and 60 DoFs per element, one type per operator. */
typedef struct amx_mrhs amx_mrhs;
/* Copies the operators' element matrices (from their FP64 storage) into the
 * MFMA layout of precision (AMX_PREC_FP32 or AMX_PREC_BF16).  The operators
 * must be finalized, on one device; at most 8. */
int amx_mrhs_create(amx_op* const* ops, int n_ops, int precision, amx_mrhs** out);
int amx_mrhs_destroy(amx_mrhs* m);
/* Launch shape: rhs_tile = 32, 64 or 128 RHS columns per wave;
 * waves_per_block = 1, 2 or 4; work_split = 1: every wave takes a contiguous
 * run of element tiles of its XCD's share, 0: the waves interleave over it. */
int amx_mrhs_set_tiles(amx_mrhs* m, int rhs_tile, int waves_per_block, int work_split);
int amx_mrhs_get_tiles(const amx_mrhs* m, int* rhs_tile, int* waves_per_block, int* work_split);
/* AMX_SCATTER_ATOMIC (default) or AMX_SCATTER_DETERMINISTIC. */
int amx_mrhs_set_scatter(amx_mrhs* m, int scatter);
/* U[i], V[i]: device arrays of operator i (n_ops of each).  V is overwritten. */
int amx_mrhs_apply_dev(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, void* stream);
/* Times every launch shape (reps launches each) on the given arrays, keeps
 * the fastest; results[18] (may be NULL) receives the mean ms per launch in
 * the order rhs_tile 32/64/128 x waves 1/2/4 x work_split 0/1 (1e30 where a
 * shape was skipped). */
int amx_mrhs_autotune(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, int reps,
                      double* results, void* stream);
/* Captures one apply (zero fill + persistent launch) into a HIP graph bound to
 * these arrays; stream must not be the default stream. */
int amx_mrhs_graph_capture(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, void* stream);
int amx_mrhs_graph_launch(amx_mrhs* m, void* stream);

/* ------------------------------------------------------- serving
 * A continuous-batching CG service on one device (no counterpart in the
 * reference, whose solves go one at a time through par::solve,
 * examples/include/solve.hpp:31-83).  Jobs are CG solves A x = b on finalized
 * one-rank operators with the deterministic scatter; each keeps its vectors and
 * scalars resident on the device.  amx_server_step runs n_iter iterations of
 * every active job, each phase as one launch over all jobs, then advances every
 * job's state (one device-to-host copy of the scalars).  A job's result is
 * bit-identical alone, in any mix, and to amx_cg_solve_dev (method
 * AMX_PREC_FP64) or amx_ir_cg_solve_dev (AMX_PREC_FP32 / AMX_PREC_BF16: FP64
 * iterative refinement, inner Jacobi-PCG on that storage).  Operators must
 * outlive their jobs; one host thread per server. */
typedef struct amx_server amx_server;
typedef struct amx_server_info {
    uint32_t active_jobs;   /* admitted, not finished */
    uint32_t held_jobs;     /* admitted, result not fetched */
    uint32_t free_slots;
    uint32_t pad;
    uint64_t workspace_bytes; /* device bytes of the held jobs (operators not counted) */
    uint64_t steps, iterations, completed;
} amx_server_info;
int amx_server_create(int device, int max_jobs, amx_server** out);
int amx_server_destroy(amx_server* server);
/* Device bytes one job of this method on this operator holds (for admission
 * by memory budget; the operator's own bytes are in amx_op_info). */
int amx_server_job_bytes(const amx_op* op, int method, uint64_t* bytes);
/* Admit a job now: copies b (n_total doubles), allocates its workspace and
 * initializes it; AMX_REJECTED when every slot is taken.  inner_rtol and
 * max_outer apply to refinement; max_it bounds each CG. */
int amx_server_submit(amx_server* server, amx_op* op, int method, const double* b, double rtol,
                      double inner_rtol, int max_it, int max_outer, int64_t* job_id);
/* n_iter batched iterations of all active jobs, then state advance;
 * *n_finished = jobs that finished in this step. */
int amx_server_step(amx_server* server, int n_iter, int* n_finished);
/* ids of finished jobs not yet reported (at most max_ids). */
int amx_server_poll(amx_server* server, int64_t* job_ids, int max_ids, int* n_ids);
/* Copy a finished job's solution (n_total doubles, may be NULL) and stats,
 * then release the job and its device memory. */
int amx_server_result(amx_server* server, int64_t job_id, double* x, amx_solve_stats* stats);
int amx_server_get_info(const amx_server* server, amx_server_info* info);

/* ------------------------------------------------------- element kernels
 * Element matrices of the examples (examples/src/ke_matrix.cpp,
 * examples/src/fe_vector.cpp), FP64 on the host.  Row-major output. */
/* integration<double>::gauss (examples/include/integration.hpp:81-113) */
int amx_fem_gauss(uint32_t n, double* pts_n_wts);
/* ke_hex8 (examples/src/ke_matrix.cpp:436-526) */
int amx_fem_ke_hex8(const double* xe, uint32_t ngt, double* ke);
/* ke_hex20_iso (examples/src/ke_matrix.cpp:1573-1740) */
int amx_fem_ke_hex20_iso(const double* xe, double E, double nu, uint32_t ngt, double* ke);
/* ke_tet10 (examples/src/ke_matrix.cpp:329-433) */
int amx_fem_ke_tet10(const double* xe, uint32_t ngt, double* ke);
/* fe_tet10 (examples/src/fe_vector.cpp:86-165) */
int amx_fem_fe_tet10(const double* xe, const double* be, uint32_t ngt, double* fe);
/* Many tet10 element matrices at once (OpenMP over elements):
 * coords[n_elem*30] -> ke[n_elem*100]. */
int amx_fem_ke_tet10_batch(uint32_t n_elem, const double* coords, uint32_t ngt, double* ke);
int amx_fem_fe_tet10_batch(uint32_t n_elem, const double* coords, const double* be,
                           uint32_t ngt, double* fe);

/* ------------------------------------------------------------ Gmsh ingest
 * GmshMesh (examples/include/GmshMesh.hpp, examples/src/GmshMesh.cpp:133-804):
 * an ASCII Gmsh 2.x mesh, partitioned by Gmsh, as one rank's inputs of
 * amx_maps_set_map / amx_maps_set_bdr_map.  Every rank reads the whole file.
 * Volume elements (Gmsh types 4 tet4, 5 hex8, 11 tet10, 12 hex27, 17 hex20)
 * are the mesh; surface elements (2, 3, 9, 10, 16) give the constrained nodes;
 * points (15) and lines (1, 8) are skipped.  Element nodes come in the aMat
 * element orders (tet10: Gmsh nodes 8 and 9 swapped, GmshMesh.cpp:364-377;
 * hex20: examples/include/shapeFunc.hpp:200-269).  A node is owned by the
 * partition of the first volume element, in file order, that names it
 * (GmshMesh.cpp:379-466); a rank's owned nodes get consecutive global ids in
 * ascending Gmsh id order (:480-495).  Local node ids are
 * [pre ghosts | owned | post ghosts]. */
typedef struct amx_gmsh amx_gmsh;

typedef struct amx_gmsh_info {
    uint32_t rank, size, dofs_per_node;
    uint32_t n_owned_elements;      /* GmshMesh::get_nOwnedElems */
    uint32_t n_owned_nodes;         /* get_nOwnedDofs / dofs_per_node */
    uint32_t n_pre_ghost_nodes, n_post_ghost_nodes;
    uint32_t n_local_nodes;         /* get_nLocalNodes */
    uint32_t n_entries;             /* sum of the owned elements' node counts */
    uint32_t max_nodes_per_element;
    uint32_t n_boundary_nodes;      /* get_n_constraints / dofs_per_node */
    uint32_t n_partitions;          /* highest partition id of the file (1 if none) */
    uint64_t n_nodes;               /* nodes of the volume elements, all ranks */
    uint64_t n_file_nodes;          /* nodes listed in $Nodes */
    uint64_t n_volume_elements, n_surface_elements, n_line_elements, n_point_elements;
    uint64_t node_offset;           /* first global node id of this rank */
    uint64_t n_dofs_total;          /* get_nDofsTotal */
    uint64_t owned_dof_begin, owned_dof_end; /* get_startGlobalDof / get_endGlobalDof (inclusive) */
} amx_gmsh_info;

/* GmshMesh(comm, nDofPerNode) + setGmshMesh (GmshMesh.cpp:133-331).  size > 1
 * needs partition tags on every element; partition ids must be <= size. */
int amx_gmsh_read(const char* path, uint32_t rank, uint32_t size, uint32_t dofs_per_node,
                  amx_gmsh** out);
int amx_gmsh_destroy(amx_gmsh* mesh);
int amx_gmsh_get_info(const amx_gmsh* mesh, amx_gmsh_info* info);
/* per owned element (n_owned_elements): get_nNodesPerElem / get_nDofsPerElem
 * (either may be NULL) and the Gmsh element type */
int amx_gmsh_get_element_sizes(const amx_gmsh* mesh, uint32_t* nodes_per_element,
                               uint32_t* dofs_per_element, uint32_t* gmsh_types);
/* element maps flattened element after element: node level (n_entries,
 * get_localMap / get_globalMap) and DoF level (n_entries * dofs_per_node,
 * get_localDofMap / get_globalDofMap, DoF = node * dofs_per_node + component);
 * any output may be NULL */
int amx_gmsh_get_maps(const amx_gmsh* mesh, uint32_t* local_map, uint64_t* global_map,
                      uint32_t* local_dof_map, uint64_t* global_dof_map);
/* local node -> global node (n_local_nodes, get_local2GlobalMap) and local DoF
 * -> global DoF (n_local_nodes * dofs_per_node, get_local2GlobalDofMap) */
int amx_gmsh_get_local2global(const amx_gmsh* mesh, uint64_t* nodes, uint64_t* dofs);
/* owned node count of every rank (size entries); times dofs_per_node these are
 * the rank_dof_counts of amx_maps_set_map */
int amx_gmsh_get_rank_node_counts(const amx_gmsh* mesh, uint32_t* counts);
/* x, y, z of every local node (n_local_nodes * 3; get_x / get_y / get_z) */
int amx_gmsh_get_coordinates(const amx_gmsh* mesh, double* xyz);
/* boundary nodes (n_boundary_nodes global node ids, ascending) and the
 * constrained DoFs (n_boundary_nodes * dofs_per_node; get_constrained_dofs) */
int amx_gmsh_get_constraints(const amx_gmsh* mesh, uint64_t* boundary_nodes, uint64_t* constrained_dofs);
/* get_prescribed_values (GmshMesh.cpp:752-783): displacement(xyz, values)
 * writes dofs_per_node values for a boundary node at xyz. */
int amx_gmsh_get_prescribed_values(const amx_gmsh* mesh,
                                   void (*displacement)(const double* xyz, double* values),
                                   double* values);

#ifdef __cplusplus
}
#endif
#endif /* AMAT_AMD_H */
