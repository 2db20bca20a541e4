// amx_op.hpp -- the operator handle (internal).
#pragma once

#include <memory>

#include "amx_internal.hpp"
#include "amx_kernels.hpp"
#include "amx_solver.hpp"

namespace amx {

constexpr uint32_t kPadSlot = 0xffffffffu;
constexpr double kPenaltyFactor = 100.0;  // PENALTY_FACTOR, include/aMat.hpp:68

// Elements sharing a DoF count n.  Slots are the positions of the group's
// device arrays: independent elements first (ascending id), padded to a whole
// tile, then the dependent ones, padded again.
struct Group {
    uint32_t n = 0;
    std::vector<uint32_t> elems;      // element ids, ascending (staging order)
    std::vector<double> stage;        // row-major n x n per element (staging order)
    Layout L[3];                      // per KType
    uint32_t G = 1;
    std::vector<uint32_t> slot_elem;  // slot -> index into elems, kPadSlot for padding
    std::vector<uint32_t> elem_slot;  // index into elems -> slot
    uint32_t n_slots = 0, n_tiles = 0, tiles_phase0 = 0;
    uint64_t ve_base = 0;             // first ve row
    // element row (slot, r) at ve_base + r*ve_rs + slot*ve_ss: row-major per element for
    // one-element tiles, row r of consecutive slots together for multi-element tiles (a
    // wave of consecutive DoFs then reads neighbouring rows in the gather)
    uint32_t ve_rs = 1, ve_ss = 1;
    DevBuf K[3];                      // tiled storage per KType (F64 always)
    DevBuf map;                       // [slot][r] local ids | identity-BC flag
    DevBuf kdiag16;                   // BF16 storage: FP32 corrected diagonal [slot][r]
    DevBuf xe, xe32, fscale;          // FREE: coordinates (FP64 and an FP32 copy) and factors per slot
    // near-symmetric records (L[0].nsym): first exception index of each row,
    // exception index of each (r, c) (0xffff: none), columns holding one
    DevBuf nsym_xs, nsym_xidx;
    uint64_t nsym_colany = 0;
    // FP32 / BF16 records (L[1] / L[2].nsym): per lane exception mask low,
    // high and first X index ([kt], kt = 1, 2)
    DevBuf psx_tab[3];
};

struct Op {
    int device = 0;
    hipStream_t stream = nullptr;
    int matfree_type = AMX_MATFREE_GPU_PURE;
    uint32_t n_streams = 1;
    int bc_meth = AMX_BC_IMATRIX;
    int bc_operator = AMX_BCOP_IDENTITY;
    int scatter = AMX_SCATTER_DETERMINISTIC;
    int prec_mask = AMX_PREC_FP64;
    int sym_mode = 1;                    // packed symmetric tiles / near-symmetric records: 0 never, 1 when
                                         // the group's matrices allow them
    bool penalty_set = false;
    double penalty_value = 0.0;
    double trace = 0.0;
    bool finalized = false;

    // copied from the maps
    uint32_t rank = 0, size = 1;
    uint32_t n_elem = 0, n_total = 0, n_owned = 0, n_pre = 0, n_post = 0;
    uint64_t dof_lo = 0;
    std::vector<uint32_t> dpe;
    std::vector<uint64_t> eoff;
    std::vector<uint32_t> local_map;
    std::vector<uint8_t> dependent;      // per element
    std::vector<uint32_t> indep_list, dep_list;
    bool has_bdr = false;
    std::vector<uint32_t> bdr_map;       // per element entry
    std::vector<double> presval_map;
    std::vector<uint32_t> owned_c_local; // local ids of Maps::ownedConstrainedDofs (input order)
    std::vector<double> owned_c_vals;
    std::vector<uint32_t> owned_free_local;
    std::vector<uint32_t> send_ids;      // ghost_send_end targets (concatenated over ranks)
    bool has_peer_lists = false;

    // element matrices
    std::vector<int32_t> elem_group, elem_index;
    std::vector<uint32_t> elem_bdim;     // blocks_dim last used (0 = never set)
    bool scaled = false;
    std::vector<double> kref;            // row-major, scaled mode
    std::vector<double> scale;           // per element, scaled mode
    std::vector<Group> groups;
    // method 2 (FREE): element kind, coordinates [e][3*nodes], factors
    int elem_kind = 0;
    uint32_t elem_ngt = 0, elem_nq = 0;
    double elem_c[3] = {0.0, 0.0, 0.0};
    std::vector<double> elem_xe, elem_scale;
    DevBuf qp;

    // device tables
    uint64_t ve_rows = 0;
    DevBuf ve, off, pos, flag, coef64, cdofs, diag, diag_bc_ids, ghost_c_ids;
    uint32_t n_cdofs = 0, n_diag_bc = 0, n_ghost_c = 0;
    DevBuf halo_dst, halo_off, halo_src;
    uint32_t n_halo_unique = 0;
    // caller's halo transport (amx_op_set_halo_transport) and its staging:
    // device send ids, device buffers of max(n_send, n_ghost) values, pinned
    // host copies for AMX_HALO_HOST
    amx_halo_fn halo_fn = nullptr;
    void* halo_ctx = nullptr;
    int halo_mode = AMX_HALO_HOST;
    bool diag_assembled = false;         // ghost rows of diag added into their owners
    DevBuf halo_send_ids, hx_send, hx_recv;
    double* hh_send = nullptr;
    double* hh_recv = nullptr;
    uint32_t n_ghost() const { return n_pre + n_post; }
    uint32_t n_send() const { return (uint32_t)send_ids.size(); }
    DevBuf scratch_u, scratch_v;         // host-array API staging
    // host-array matvec timers (aMatFree::get_timer): events of the phases
    // (created with the staging buffers) and seconds accumulated since finalize
    hipEvent_t tev[4] = {};
    double t_scatter = 0, t_mv = 0, t_gather = 0, t_total = 0;
    // solver workspace (allocated on first solve)
    DevBuf ws_r, ws_z, ws_p, ws_q, ws_dinv, ws_partials, ws_sc, ws_x32, ws_b32, ws_r32, ws_z32, ws_p32, ws_q32;
    DevBuf ws_ib, ws_ix, ws_ir, ws_iz, ws_ip, ws_iq;  // refinement with FP64 inner vectors
    CgScalars* host_sc = nullptr;        // pinned copy of the device scalars
    uint64_t device_bytes = 0;

    Op() = default;
    Op(const Op&) = delete;
    Op& operator=(const Op&) = delete;
    ~Op()
    {
        if (host_sc) (void)hipHostFree(host_sc);
        if (hh_send) (void)hipHostFree(hh_send);
        if (hh_recv) (void)hipHostFree(hh_recv);
        for (hipEvent_t e : tev)
            if (e) (void)hipEventDestroy(e);
    }
    double penalty() const { return penalty_set ? penalty_value : kPenaltyFactor * trace; }
    // the caller's stream; NULL is the HIP default stream (not the handle's)
    hipStream_t pick(void* s) const { return static_cast<hipStream_t>(s); }
    bool is_free() const { return matfree_type == AMX_MATFREE_FREE; }
    // FREE operators apply in FP32 with nothing stored: FP32 is always available
    bool has(int prec_bit) const { return (prec_mask & prec_bit) != 0 || (prec_bit == AMX_PREC_FP32 && is_free()); }
};

// nodes of a FREE element kind (amx_element_kind)
inline uint32_t free_nodes(int kind)
{
    return kind == AMX_ELEM_HEX8_LAPLACE ? 8u : (kind == AMX_ELEM_HEX20_ELASTIC ? 20u : 10u);
}
// algorithmic FP64 flops of one FREE element application (amx_free.hip):
// Laplace per point 30 n + 85 (Jacobian 18 n, h 6 n, output 6 n, inverse and
// transforms 85); hex20 per point 1260 (Jacobian 360, H 360, output 360,
// inverse, strain, stress and P 180)
inline uint64_t free_flops_per_elem(int kind, uint32_t nq)
{
    return (uint64_t)nq * (kind == AMX_ELEM_HEX20_ELASTIC ? 1260u : 30u * free_nodes(kind) + 85u);
}

// RAII: make the handle's device current, restore the previous one on exit.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) AMX_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// launch pieces shared with the solvers
// bc: 0 none, 1 identity, 2 penalty.  phase: 0 independent tiles, 1 dependent
// tiles, 2 all tiles.  vt_double: FP64 vectors and storage, else FP32 vectors
// with storage kt (F32 or BF16).
// skip: optional device flag; a nonzero value turns the launches into no-ops.
// ve: element-row buffer of the deterministic scatter (NULL: the handle's own;
// callers running several applications of one operator at once bring theirs,
// of ve_rows doubles).
void ebe_launch(Op& op, KType kt, bool vt_double, bool masked, int phase, const void* u, void* v, hipStream_t s,
                const uint32_t* skip = nullptr, void* ve = nullptr);
void gather_launch(Op& op, bool vt_double, int bc, const void* u, void* v, hipStream_t s,
                   const uint32_t* skip = nullptr, const void* ve = nullptr);
// complete one-rank operator application: v = A u with BC form bc
void apply(Op& op, KType kt, bool vt_double, int bc, const void* u, void* v, hipStream_t s,
           const uint32_t* skip = nullptr, void* ve = nullptr);
// Jacobi inverse diagonal of the operator solved (constrained rows: 1 for
// identity rows, the penalty for penalty rows); scratch: n_total doubles.
void build_dinv(Op& op, double* dinv, double* scratch, hipStream_t s);

// Workspace of one FP64 CG solve: five n_total vectors, the reduction
// partials (kPartialsBytes, counters at zero), the device scalars and an
// optional pinned host copy of them (NULL: read through a local)
struct CgWork {
    double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr, *partials = nullptr;
    CgScalars* sc = nullptr;
    CgScalars* host = nullptr;
};
// FP64 Jacobi-PCG with true-residual replacements on workspace w (the body
// of amx_cg_solve_dev); allocates nothing
void fp64_cg(Op& op, const double* b, double* x, double rtol, int max_it, const CgWork& w, amx_solve_stats* stats,
             hipStream_t s);

}  // namespace amx

struct amx_op {
    amx::Op op;
};
