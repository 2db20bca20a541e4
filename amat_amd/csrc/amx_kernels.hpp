// amx_kernels.hpp -- host-side launchers of the gfx950 kernels (amx_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace amx {

// Element-matrix storage types on the device.
enum class KType : int { F64 = 0, F32 = 1, BF16 = 2 };

// Device layout of one element group's matrices ("column-packed tiles"):
// a wave processes a tile of G elements with n DoFs each (lane = g*n + r);
// pack j of the tile holds, for every lane, the P consecutive columns
// j*P .. j*P+P-1 of row r, so one pack is one contiguous G*n*P-value load of
// the whole wave.  Offset of (slot, r, c), slot = t*G + g:
//   ((t*(n/P) + c/P)*(G*n) + g*n + r)*P + c%P.
//
// Packed symmetric tiles (sym): when every element matrix of a group is
// bitwise symmetric, a multi-element tile stores only the upper triangles:
// element g of tile t at t*TE + g*NS, NS = n(n+1)/2 values, entry (a <= b) at
// a*n - a(a-1)/2 + (b - a) (row-major upper triangle); TE = G*NS rounded up
// to 16 bytes.  The kernels read K[r][c] = K[min][max]: the same values in the
// same order as the full layout, about half the bytes.
//
// FP32 / BF16 storage of such a group (nsym in L[1] / L[2]) holds packed
// symmetric records with exceptions: the rounded matrices are bitwise
// symmetric except at the group's exception positions (amx_nsym.hpp).
//
// Near-symmetric records (nsym, one-element tiles of FP64 storage, n = 60):
// when every matrix of the group is symmetric up to small differences in the
// low word of the transposed entries (quadrature round-off) and at most
// nsym_max_exceptions(n) lower positions need their own values, each element
// is one record of nsym_record_bytes(n, nx) bytes: upper triangle, the nx
// exception values and one byte per lower entry (amx_nsym.hpp).  The kernel
// rebuilds every K[r][c] bit for bit and keeps the whole layout's order.
struct Layout {
    uint32_t n = 0, G = 1, P = 1;
    bool specialized = false;  // a template kernel exists for (n, G, P, type)
    bool sym = false;          // packed symmetric tiles
    bool nsym = false;         // near-symmetric records
    uint32_t nx = 0;           // nsym: exception values per record
};

Layout choose_layout(uint32_t n, KType kt);
size_t ktype_size(KType kt);
// packed symmetric tiles exist for these DoF counts (multi-element tiles)
bool sym_supported(uint32_t n);
// near-symmetric records exist for these DoF counts (one-element FP64 tiles)
bool nsym_supported(uint32_t n);
uint32_t nsym_max_exceptions(uint32_t n, KType kt);
uint64_t nsym_record_bytes(uint32_t n, uint32_t nx);
// values per stored tile of layout L in storage kt (padding included)
uint64_t tile_values(const Layout& L, KType kt);

// ---- set-up kernels -----------------------------------------------------
// row-major FP64 element matrices (group element order) -> tiled layout of kt.
// slot_elem[slot] = index into src (UINT32_MAX for padding slots -> zeros).
void launch_relayout_from_rowmajor(const double* src, const uint32_t* slot_elem, uint32_t n_slots,
                                   const Layout& L, KType kt, void* dst, hipStream_t s);
// K_e = scale[elem] * k_ref (row-major n x n) -> tiled layout of kt.
void launch_relayout_scaled(const double* k_ref, const double* scale, const uint32_t* slot_elem,
                            uint32_t n_slots, const Layout& L, KType kt, void* dst, hipStream_t s);
// FP32 defect-corrected diagonal of the BF16 storage, [slot][r]; source either
// row-major matrices (src_rowmajor) or k_ref + scale.
void launch_bf16_diag(const double* src_rowmajor, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                      uint32_t n_slots, uint32_t n, float* out, hipStream_t s);
// near-symmetric records: flags[r*n + c] (r > c) set to 1 where some matrix of
// the group cannot code K[r][c] from K[c][r] (flags zeroed by the caller);
// source either row-major matrices (src) or k_ref + scale
void launch_nsym_scan(const double* src, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                      uint32_t n_slots, uint32_t n, KType kt, uint32_t* flags, hipStream_t s);
// packed symmetric records with exceptions of FP32 / BF16 storage (L.nsym for kt F32 / BF16)
void launch_relayout_psx(const double* src, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                         uint32_t n_slots, const Layout& L, KType kt, const uint16_t* xrc, void* dst, hipStream_t s);
// the records of layout L (L.nsym, L.nx): xrc[x] = r*n + c of exception x in
// row-major order, xidx[r*n + c] = x (0xffff elsewhere)
void launch_relayout_nsym(const double* src, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                          uint32_t n_slots, const Layout& L, const uint16_t* xrc, const uint16_t* xidx, void* dst,
                          hipStream_t s);
// diagonal entries K_e(r,r) of every slot row from the tiled FP64 storage into ve
// (row (slot, r) at ve[r*rs + slot*ss]).
void launch_extract_diag(const double* K, const Layout& L, uint32_t n_slots, uint32_t rs, uint32_t ss, double* ve,
                         hipStream_t s);

// ---- matvec kernels ----------------------------------------------------
struct EbeArgs {
    const void* K;          // tiled storage
    const uint32_t* map;    // [slot][r] local DoF ids, bit 31 = zero on gather (identity BC)
    const void* u;
    void* out;              // ve (deterministic) or v (atomic)
    uint64_t ve_base;       // first ve row of this group
    uint32_t ve_rs, ve_ss;  // element row (slot, r) at ve_base + r*ve_rs + slot*ve_ss
    uint32_t tile_begin, n_tiles;
    const uint32_t* skip = nullptr;  // device flag: nonzero -> kernel exits (converged solver)
    const float* kdiag = nullptr;    // BF16 storage: FP32 diagonal [slot][r]
    const uint32_t* nsym_xs = nullptr;  // near-symmetric records: first X index of each row (FP64); FP32 /
                                        // BF16: per lane exception mask low, high, first X index
    uint64_t nsym_colany = 0;           // FP64: bit c: some row has an exception in column c
};
// vt_double: vectors/accumulation in FP64 (only with KType::F64), else FP32.
void launch_ebe(const Layout& L, KType kt, bool vt_double, bool masked, bool atomic, const EbeArgs& a,
                hipStream_t s);

// v[d] = sum_{q in [off[d], off[d+1])} ve[pos[q]] (fixed order), then the BC
// form: bc 0 none; 1 identity (flag[d] -> v[d] = u[d]); 2 penalty
// (flag[d] -> v[d] += coef[d] * u[d]).
void launch_gather(bool vt_double, int bc, const uint32_t* off, const uint32_t* pos, const void* ve,
                   const void* u, const uint8_t* flag, const double* coef, void* v, uint32_t n, hipStream_t s,
                   const uint32_t* skip = nullptr);

// ---- batched (many operators per launch) -------------------------------
// One entry per (job, element group); work items are tiles (EbE) or blocks of
// 256 DoFs (gather), prefix[i] = first item of entry i, prefix[n_jobs] = total.
struct EbeJob {
    const void* K;
    const uint32_t* map;
    const void* u;
    void* out;              // the job's ve array
    uint64_t ve_base;
    uint32_t ve_rs, ve_ss;  // as EbeArgs
    uint32_t tile_begin, n_tiles;
    const uint32_t* skip;   // the job's converged flag
    const float* kdiag;
    uint32_t masked, sym;   // sym: packed symmetric tiles
    const uint32_t* nsym_xs;  // 60-DoF records (EbeArgs), nsym_nx exceptions
    uint64_t nsym_colany;
    uint32_t nsym_nx;
};
struct GatherJob {
    const uint32_t* off;
    const uint32_t* pos;
    const void* ve;
    const void* u;
    const uint8_t* flag;
    const double* coef;
    void* v;
    uint32_t n, bc;
    const uint32_t* skip;
};
bool batched_n(uint32_t n);
// records: every job's group holds near-symmetric records (FP64) or packed
// records with exceptions (FP32 / BF16) of 60-DoF elements, at most max_nx
// exception values each, and every job's masked flag equals `masked` (no job
// of a launch without `records` holds records)
void launch_ebe_batched(uint32_t n, KType kt, bool vt_double, const EbeJob* jobs, const uint32_t* prefix,
                        uint32_t n_jobs, uint32_t total_tiles, hipStream_t s, bool records, uint32_t max_nx,
                        bool masked);
void launch_gather_batched(bool vt_double, const GatherJob* jobs, const uint32_t* prefix, uint32_t n_jobs,
                           uint32_t total_blocks, hipStream_t s);

// ---- method 2 (FREE): element kernels from coordinates (amx_free.hip) ---
struct FreeArgs {
    const double* xe;        // Laplace kinds: [3*nodes][stride]; hex20: [slot][60]
    const float* xe32;       // the same coordinates rounded to FP32 (the FP32 form reads these)
    const double* scale;     // [slot] stiffness factor (0 on padding slots)
    const uint32_t* map;     // as EbeArgs::map
    const void* u;           // FP64, or FP32 with f32
    void* out;               // ve (deterministic) or v (atomic), same type as u
    uint64_t ve_base;
    uint32_t ve_rs, ve_ss;   // as EbeArgs
    uint32_t slot_begin, n_slots, stride;
    const double* qp;        // [points][3*nodes+1]: dN/dxi per node, weight
    uint32_t nq;             // quadrature points (hex20: table padded to free_padded_points rows)
    double c1, c2, c3;       // isotropic constitutive entries (hex20)
    const uint32_t* skip = nullptr;
    bool f32 = false;        // compute and vectors in FP32 (inner solves); FP64 otherwise
};
// kind: amx_element_kind.  diag: write diag(s_e K_e) into ve rows instead (FP64 only).
void launch_free(int kind, bool masked, bool atomic, bool diag, const FreeArgs& a, hipStream_t s);
uint32_t free_padded_points(int kind, uint32_t nq);

// constrained rows after an atomic scatter or a halo accumulation
void launch_bc_fixup(bool vt_double, int bc, const uint32_t* cdofs, uint32_t nc, const double* coef,
                     const void* u, void* v, hipStream_t s);
void launch_fill(void* p, uint64_t n, bool vt_double, double value, hipStream_t s);
void launch_coef(const double* diag, const uint32_t* cdofs, uint32_t nc, double penalty, double* coef64,
                 hipStream_t s);
void launch_set_entries(double* u, const uint32_t* ids, uint32_t n, double value, hipStream_t s);
// dst[i] = src[ids[i]], i < n (the owner's halo pack)
void launch_pack(double* dst, const double* src, const uint32_t* ids, uint32_t n, hipStream_t s);
// halo accumulate: v[dst[i]] += sum_{q in [off[i], off[i+1])} recv[src[q]], i < n_unique
void launch_halo_accumulate(double* v, const double* recv, const uint32_t* dst, const uint32_t* off,
                            const uint32_t* src, uint32_t n_unique, hipStream_t s);
void launch_d2f(const double* a, float* b, uint64_t n, hipStream_t s);
void launch_f2d(const float* a, double* b, uint64_t n, hipStream_t s);

}  // namespace amx
