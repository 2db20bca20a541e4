// amx_op.cpp -- the aMatFree operator on one GPU (methods 3/4/5 of
// include/enums.hpp:56 in paralab/aMat), C ABI.
//
// Set-up copies the maps and element matrices, then finalize builds the
// device-resident form: per element group a column-packed tiled K (FP64, and
// FP32/BF16 on request), the local map with the identity-BC flag folded into
// bit 31, the node -> (element, row) transposed lists of the deterministic
// gather, and the BC tables.  Each matvec is then one EbE launch per group and
// one gather launch; u and v never leave HBM.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "amx_op.hpp"

using amx::DevBuf;
using amx::fail;
using amx::guarded;
using amx::KType;
using amx::Op;

namespace amx {

namespace {

constexpr uint32_t kFlagBit = 0x80000000u;

template <typename T> void upload(DevBuf& b, const std::vector<T>& h, hipStream_t s, uint64_t& counter)
{
    b.alloc(std::max<size_t>(h.size(), 1) * sizeof(T));
    counter += b.bytes;
    if (!h.empty()) AMX_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

void require_finalized(const Op& op)
{
    AMX_REQUIRE(op.finalized, AMX_INVALID_STATE, "finalize() has not been called");
}

KType ktype_of(int precision)
{
    if (precision == AMX_PREC_FP32) return KType::F32;
    if (precision == AMX_PREC_BF16) return KType::BF16;
    if (precision == AMX_PREC_FP64) return KType::F64;
    fail(AMX_INVALID_ARGUMENT, "unknown precision " + std::to_string(precision));
}

// FREE: coordinates and factors in slot order (hex20 [slot][60], the Laplace
// kinds [coordinate][slot]); padding slots repeat the group's first element
// with factor 0, so that they add exact zeros.
void upload_free(Op& op, Group& g, hipStream_t s)
{
    const uint32_t nx = 3 * free_nodes(op.elem_kind);
    const bool aos = op.elem_kind == AMX_ELEM_HEX20_ELASTIC;
    std::vector<double> hx((size_t)g.n_slots * nx), hs(g.n_slots, 0.0);
    for (uint32_t sl = 0; sl < g.n_slots; sl++) {
        const uint32_t i = g.slot_elem[sl];
        const uint32_t e = g.elems[i == kPadSlot ? 0 : i];
        const double* src = op.elem_xe.data() + (size_t)e * nx;
        if (i != kPadSlot) hs[sl] = op.elem_scale[e];
        for (uint32_t c = 0; c < nx; c++) {
            if (aos) hx[(size_t)sl * nx + c] = src[c];
            else hx[(size_t)c * g.n_slots + sl] = src[c];
        }
    }
    upload(g.xe, hx, s, op.device_bytes);
    const std::vector<float> hx32(hx.begin(), hx.end());
    upload(g.xe32, hx32, s, op.device_bytes);
    upload(g.fscale, hs, s, op.device_bytes);
    AMX_HIP(hipStreamSynchronize(s));  // host vectors leave scope
}

FreeArgs free_args(const Op& op, const Group& g, const void* u, void* out, uint32_t slot_begin, uint32_t n_slots,
                   const uint32_t* skip, bool f32 = false)
{
    FreeArgs a;
    a.xe = g.xe.as<double>();
    a.xe32 = g.xe32.as<float>();
    a.scale = g.fscale.as<double>();
    a.map = g.map.as<uint32_t>();
    a.u = u;
    a.out = out;
    a.f32 = f32;
    a.ve_base = g.ve_base;
    a.ve_rs = g.ve_rs;
    a.ve_ss = g.ve_ss;
    a.slot_begin = slot_begin;
    a.n_slots = n_slots;
    a.stride = g.n_slots;
    a.qp = op.qp.as<double>();
    a.nq = op.elem_nq;
    a.c1 = op.elem_c[0];
    a.c2 = op.elem_c[1];
    a.c3 = op.elem_c[2];
    a.skip = skip;
    return a;
}

// ---- halo exchange through the caller's transport (ghost_receive_begin/end
// and ghost_send_begin/end, include/aMatFree.hpp:1426-1679)

// hx_send holds n_out values; the transport fills n_in values of hx_recv
void transport(Op& op, int dir, uint32_t n_out, uint32_t n_in, hipStream_t s)
{
    AMX_REQUIRE(op.halo_fn != nullptr, AMX_NOT_IMPLEMENTED, "no halo transport registered");
    int rc = 0;
    if (op.halo_mode == AMX_HALO_HOST) {
        if (n_out) AMX_HIP(hipMemcpyAsync(op.hh_send, op.hx_send.p, n_out * sizeof(double), hipMemcpyDeviceToHost, s));
        // also orders the previous exchange's copy out of hh_recv before fn writes it
        AMX_HIP(hipStreamSynchronize(s));
        rc = op.halo_fn(op.halo_ctx, dir, op.hh_send, op.hh_recv, s);
        if (rc == 0 && n_in)
            AMX_HIP(hipMemcpyAsync(op.hx_recv.p, op.hh_recv, n_in * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        rc = op.halo_fn(op.halo_ctx, dir, op.hx_send.as<double>(), op.hx_recv.as<double>(), s);
    }
    AMX_REQUIRE(rc == 0, AMX_WRONG_COMMUNICATION, "halo transport returned " + std::to_string(rc));
}

// owners' values into the ghost slots of u (pack, transport, import)
void exchange_u(Op& op, double* u, hipStream_t s)
{
    launch_pack(op.hx_send.as<double>(), u, op.halo_send_ids.as<uint32_t>(), op.n_send(), s);
    transport(op, AMX_HALO_OWNER_TO_GHOST, op.n_send(), op.n_ghost(), s);
    const double* r = op.hx_recv.as<double>();
    if (op.n_pre) AMX_HIP(hipMemcpyAsync(u, r, op.n_pre * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (op.n_post)
        AMX_HIP(hipMemcpyAsync(u + op.n_pre + op.n_owned, r + op.n_pre, op.n_post * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
}

// ghost slots of v added into their owners' rows (export, transport, accumulate)
void exchange_v(Op& op, double* v, hipStream_t s)
{
    double* x = op.hx_send.as<double>();
    if (op.n_pre) AMX_HIP(hipMemcpyAsync(x, v, op.n_pre * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (op.n_post)
        AMX_HIP(hipMemcpyAsync(x + op.n_pre, v + op.n_pre + op.n_owned, op.n_post * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
    transport(op, AMX_HALO_GHOST_TO_OWNER, op.n_ghost(), op.n_send(), s);
    launch_halo_accumulate(v, op.hx_recv.as<double>(), op.halo_dst.as<uint32_t>(), op.halo_off.as<uint32_t>(),
                           op.halo_src.as<uint32_t>(), op.n_halo_unique, s);
}

// the assembled diagonal on the owned rows (mat_get_diagonal_ghosted's
// ghost_send_begin/end, :1051-1053) and the penalty coefficients from it
void assemble_diag(Op& op, hipStream_t s)
{
    exchange_v(op, op.diag.as<double>(), s);
    launch_fill(op.coef64.p, op.n_total, true, 0.0, s);
    launch_coef(op.diag.as<double>(), op.cdofs.as<uint32_t>(), op.n_cdofs, op.penalty(), op.coef64.as<double>(), s);
    AMX_HIP(hipStreamSynchronize(s));
    op.diag_assembled = true;
}

// more than one rank: the exchanging forms need the transport
void require_exchange(const Op& op)
{
    if (op.size == 1) return;
    AMX_REQUIRE(op.halo_fn != nullptr, AMX_NOT_IMPLEMENTED,
                "a rank's share exchanges halos: register amx_op_set_halo_transport, or drive "
                "amx_op_matvec_begin_dev / halo_* / matvec_end_dev / bc_fixup_dev");
    AMX_REQUIRE(op.diag_assembled, AMX_INVALID_STATE, "diagonal not assembled across ranks");
}

// Build slots, device K, maps, transposed lists and BC tables.
void finalize_op(Op& op)
{
    DeviceGuard dg(op.device);
    hipStream_t s = op.stream;
    // free previous device state (finalize may be called again after updates)
    op.device_bytes = 0;
    for (auto& g : op.groups) {
        for (auto& k : g.K) k.release();
        g.kdiag16.release();
        g.xe.release();
        g.xe32.release();
        g.fscale.release();
        g.nsym_xs.release();
        g.nsym_xidx.release();
        g.nsym_colany = 0;
        for (auto& t : g.psx_tab) t.release();
    }
    const bool free_m = op.is_free();
    if (free_m) {
        AMX_REQUIRE(op.elem_kind != 0, AMX_INVALID_STATE, "method 2 (FREE) needs set_element_function before finalize");
        AMX_REQUIRE(!op.has(AMX_PREC_BF16), AMX_NOT_IMPLEMENTED, "method 2 (FREE) has no BF16 form");
        uint32_t nq = 0;
        const std::vector<double> tab =
            free_quadrature(op.elem_kind, op.elem_ngt, nq, free_padded_points(op.elem_kind, op.elem_nq));
        upload(op.qp, tab, s, op.device_bytes);
        AMX_HIP(hipStreamSynchronize(s));
    }

    // flagged local ids: zeroed on gather by the identity BC form -- owned
    // constrained DoFs (MatMult_mf :3916-3922) and the element DoFs the BC map
    // marks (ghost copies of constrained DoFs, zeroed by their owner)
    std::vector<uint8_t> zero_on_gather(op.n_total, 0);
    for (uint32_t d : op.owned_c_local) zero_on_gather[d] = 1;
    if (op.has_bdr)
        for (uint64_t k = 0; k < op.local_map.size(); k++)
            if (op.bdr_map[k] == 1) zero_on_gather[op.local_map[k]] = 1;

    // ---- slots and device matrices per group
    uint64_t ve_rows = 0;
    for (auto& g : op.groups) {
        const uint32_t n = g.n;
        g.L[0] = choose_layout(n, KType::F64);
        g.L[1] = choose_layout(n, KType::F32);
        g.L[2] = choose_layout(n, KType::BF16);
        g.G = g.L[0].G;
        // packed symmetric tiles when every stored matrix of the group is
        // bitwise symmetric (then K[r][c] and K[c][r] are the same value)
        if (!free_m && op.sym_mode == 1 && g.G > 1 && sym_supported(n)) {
            bool sym = true;
            auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; };
            if (op.scaled) {
                for (uint32_t r = 0; r < n && sym; r++)
                    for (uint32_t c = r + 1; c < n && sym; c++)
                        sym = same(op.kref[(size_t)r * n + c], op.kref[(size_t)c * n + r]);
            } else if (!g.stage.empty()) {
                for (size_t i = 0; i < g.elems.size() && sym; i++) {
                    const double* ke = g.stage.data() + i * n * n;
                    for (uint32_t r = 0; r < n && sym; r++)
                        for (uint32_t c = r + 1; c < n && sym; c++) sym = same(ke[(size_t)r * n + c], ke[(size_t)c * n + r]);
                }
            }
            for (Layout& L : g.L) L.sym = sym;
        }
        // independent elements first, then dependent, each padded to a tile
        g.slot_elem.clear();
        g.elem_slot.assign(g.elems.size(), kPadSlot);
        for (int phase = 0; phase < 2; phase++) {
            for (uint32_t i = 0; i < g.elems.size(); i++) {
                if (op.dependent[g.elems[i]] != phase) continue;
                g.elem_slot[i] = (uint32_t)g.slot_elem.size();
                g.slot_elem.push_back(i);
            }
            while (g.slot_elem.size() % g.G) g.slot_elem.push_back(kPadSlot);
            if (phase == 0) g.tiles_phase0 = (uint32_t)(g.slot_elem.size() / g.G);
        }
        g.n_slots = (uint32_t)g.slot_elem.size();
        g.n_tiles = g.n_slots / g.G;
        // element rows (ve) of the deterministic scatter; set once the storage
        // is known (near-symmetric records pad each element's rows to 64)
        auto assign_ve = [&](uint32_t pitch) {
            g.ve_base = ve_rows;
            g.ve_rs = g.G > 1 ? g.n_slots : 1u;
            g.ve_ss = g.G > 1 ? 1u : pitch;
            ve_rows += (uint64_t)g.n_slots * (g.G > 1 ? n : pitch);
        };

        // map [slot][r]
        std::vector<uint32_t> hmap((size_t)g.n_slots * n, 0);
        for (uint32_t sl = 0; sl < g.n_slots; sl++) {
            const uint32_t i = g.slot_elem[sl];
            if (i == kPadSlot) continue;
            const uint32_t e = g.elems[i];
            for (uint32_t r = 0; r < n; r++) {
                const uint32_t d = op.local_map[op.eoff[e] + r];
                hmap[(size_t)sl * n + r] = d | (zero_on_gather[d] ? kFlagBit : 0u);
            }
        }
        upload(g.map, hmap, s, op.device_bytes);
        if (free_m) {
            assign_ve(n);
            upload_free(op, g, s);
            continue;
        }

        // element matrices: upload the source once, relayout on the device
        DevBuf src, src_scale, dslot;
        std::vector<uint32_t> slot_src(g.slot_elem);
        if (op.scaled) {
            std::vector<double> sc(g.elems.size());
            for (uint32_t i = 0; i < g.elems.size(); i++) sc[i] = op.scale[g.elems[i]];
            uint64_t tmp = 0;
            upload(src, op.kref, s, tmp);
            upload(src_scale, sc, s, tmp);
        } else {
            if (g.stage.empty()) g.stage.assign((size_t)g.elems.size() * n * n, 0.0);
            uint64_t tmp = 0;
            upload(src, g.stage, s, tmp);
        }
        {
            uint64_t tmp = 0;
            upload(dslot, slot_src, s, tmp);
        }
        const double* src_rm = op.scaled ? nullptr : src.as<double>();
        const double* src_kref = op.scaled ? src.as<double>() : nullptr;
        const double* src_sc = op.scaled ? src_scale.as<double>() : nullptr;
        // near-symmetric records for one-element FP64 tiles: find the lower
        // positions that some matrix cannot code from its transpose; few
        // enough of them -> records, else the whole layout
        // near-symmetric records (FP64) and packed symmetric records with
        // exceptions (FP32 / BF16) for one-element tiles: find the lower
        // positions that some matrix cannot take from its transposed entry;
        // few enough of them -> records, else the whole layout
        const int bits[3] = {AMX_PREC_FP64, AMX_PREC_FP32, AMX_PREC_BF16};
        DevBuf dxrc[3];
        // host sources of the asynchronous uploads below, alive until the
        // stream synchronization at the end of the group
        std::vector<uint16_t> xrc[3], xidx;
        std::vector<uint32_t> xs, ptab[3];
        for (int kt = 0; kt < 3; kt++) {
            if (op.sym_mode != 1 || g.G != 1 || !nsym_supported(n) || g.n_slots == 0) break;
            if (kt > 0 && !op.has(bits[kt])) continue;
            DevBuf flags;
            flags.alloc((size_t)n * n * sizeof(uint32_t));
            AMX_HIP(hipMemsetAsync(flags.p, 0, flags.bytes, s));
            launch_nsym_scan(src_rm, src_kref, src_sc, dslot.as<uint32_t>(), g.n_slots, n, (KType)kt,
                             flags.as<uint32_t>(), s);
            std::vector<uint32_t> hf((size_t)n * n);
            AMX_HIP(hipMemcpyAsync(hf.data(), flags.p, flags.bytes, hipMemcpyDeviceToHost, s));
            AMX_HIP(hipStreamSynchronize(s));
            std::vector<uint32_t> first(n, 0u);
            uint64_t colany = 0;
            std::vector<uint64_t> rowmask(n, 0ull);
            if (kt == 0) xidx.assign((size_t)n * n, 0xffffu);
            for (uint32_t r = 0; r < n; r++) {
                first[r] = (uint32_t)xrc[kt].size();
                for (uint32_t c = 0; c < r; c++) {
                    if (!hf[(size_t)r * n + c]) continue;
                    if (kt == 0) xidx[(size_t)r * n + c] = (uint16_t)xrc[kt].size();
                    xrc[kt].push_back((uint16_t)(r * n + c));
                    colany |= 1ull << c;
                    rowmask[r] |= 1ull << c;
                }
            }
            if (xrc[kt].size() > nsym_max_exceptions(n, (KType)kt)) continue;
            g.L[kt].nsym = true;
            g.L[kt].nx = (uint32_t)xrc[kt].size();
            uint64_t tmp = 0;
            upload(dxrc[kt], xrc[kt], s, tmp);
            if (kt == 0) {
                g.nsym_colany = colany;
                xs = first;
                upload(g.nsym_xs, xs, s, op.device_bytes);
                upload(g.nsym_xidx, xidx, s, op.device_bytes);
            } else {  // per lane (row min(l, n-1)): exception mask low, high, first X index
                ptab[kt].assign(3 * 64, 0u);
                for (uint32_t l = 0; l < 64; l++) {
                    const uint32_t r = std::min(l, n - 1u);
                    ptab[kt][3 * l] = (uint32_t)rowmask[r];
                    ptab[kt][3 * l + 1] = (uint32_t)(rowmask[r] >> 32);
                    ptab[kt][3 * l + 2] = first[r];
                }
                upload(g.psx_tab[kt], ptab[kt], s, op.device_bytes);
            }
        }
        // near-symmetric records: one wave per element stores its n rows; a
        // 512-byte pitch keeps two elements' rows out of one cache line (whole
        // lines reach HBM: 214 MB of writes per hex20 64^3 matvec at 480 bytes)
        assign_ve(g.L[0].nsym ? 64u : n);
        for (int kt = 0; kt < 3; kt++) {
            if (kt > 0 && !op.has(bits[kt])) continue;
            const uint64_t bytes = (uint64_t)g.n_tiles * tile_values(g.L[kt], (KType)kt) * ktype_size((KType)kt);
            g.K[kt].alloc(std::max<uint64_t>(bytes, 16));
            op.device_bytes += g.K[kt].bytes;
            if (kt == 0 && g.L[0].nsym)
                launch_relayout_nsym(src_rm, src_kref, src_sc, dslot.as<uint32_t>(), g.n_slots, g.L[0],
                                     dxrc[0].as<uint16_t>(), g.nsym_xidx.as<uint16_t>(), g.K[0].p, s);
            else if (g.L[kt].nsym)
                launch_relayout_psx(src_rm, src_kref, src_sc, dslot.as<uint32_t>(), g.n_slots, g.L[kt], (KType)kt,
                                    dxrc[kt].as<uint16_t>(), g.K[kt].p, s);
            else if (op.scaled)
                launch_relayout_scaled(src.as<double>(), src_scale.as<double>(), dslot.as<uint32_t>(), g.n_slots,
                                       g.L[kt], (KType)kt, g.K[kt].p, s);
            else
                launch_relayout_from_rowmajor(src.as<double>(), dslot.as<uint32_t>(), g.n_slots, g.L[kt], (KType)kt,
                                              g.K[kt].p, s);
            if (kt == (int)KType::BF16) {
                g.kdiag16.alloc(std::max<uint64_t>((uint64_t)g.n_slots * n, 1) * sizeof(float));
                op.device_bytes += g.kdiag16.bytes;
                launch_bf16_diag(op.scaled ? nullptr : src.as<double>(), op.scaled ? src.as<double>() : nullptr,
                                 op.scaled ? src_scale.as<double>() : nullptr, dslot.as<uint32_t>(), g.n_slots, n,
                                 g.kdiag16.as<float>(), s);
            }
        }
        AMX_HIP(hipStreamSynchronize(s));  // sources are freed at scope exit
    }
    op.ve_rows = ve_rows;
    op.ve.alloc(std::max<uint64_t>(ve_rows, 1) * sizeof(double));
    op.device_bytes += op.ve.bytes;

    // ---- transposed gather lists in the reference's summation order:
    // independent elements then dependent ones, ascending id, rows ascending
    // (matvec_ghosted_OMP :1828-2012 then :2048-2230 with one thread)
    std::vector<uint32_t> hoff(op.n_total + 1, 0);
    for (uint32_t d : op.local_map) hoff[d + 1]++;
    for (uint32_t d = 0; d < op.n_total; d++) hoff[d + 1] += hoff[d];
    std::vector<uint32_t> fill(hoff.begin(), hoff.end() - 1);
    std::vector<uint32_t> hpos(op.local_map.size());
    AMX_REQUIRE(ve_rows < (uint64_t(1) << 32), AMX_NOT_IMPLEMENTED, "more than 2^32 element rows in one operator");
    for (const auto* list : {&op.indep_list, &op.dep_list}) {
        for (uint32_t e : *list) {
            const Group& g = op.groups[op.elem_group[e]];
            const uint64_t base = g.ve_base + (uint64_t)g.elem_slot[op.elem_index[e]] * g.ve_ss;
            for (uint32_t r = 0; r < g.n; r++)
                hpos[fill[op.local_map[op.eoff[e] + r]]++] = (uint32_t)(base + (uint64_t)r * g.ve_rs);
        }
    }
    upload(op.off, hoff, s, op.device_bytes);
    upload(op.pos, hpos, s, op.device_bytes);

    // ---- BC tables
    std::vector<uint8_t> hflag(op.n_total, 0);
    for (uint32_t d : op.owned_c_local) hflag[d] = 1;
    upload(op.flag, hflag, s, op.device_bytes);
    upload(op.cdofs, op.owned_c_local, s, op.device_bytes);
    op.n_cdofs = (uint32_t)op.owned_c_local.size();
    // constrained ghosts (zeroed after a halo import under the identity form)
    std::vector<uint32_t> ghosts_c;
    for (uint32_t d = 0; d < op.n_total; d++)
        if (zero_on_gather[d] && (d < op.n_pre || d >= op.n_pre + op.n_owned)) ghosts_c.push_back(d);
    upload(op.ghost_c_ids, ghosts_c, s, op.device_bytes);
    op.n_ghost_c = (uint32_t)ghosts_c.size();
    // owned rows whose diagonal apply_bc_diagonal overwrites (:4318-4333)
    std::vector<uint32_t> dbc;
    if (op.has_bdr) {
        std::vector<uint8_t> seen(op.n_total, 0);
        for (uint64_t k = 0; k < op.local_map.size(); k++) {
            const uint32_t d = op.local_map[k];
            if (op.bdr_map[k] == 1 && d >= op.n_pre && d < op.n_pre + op.n_owned && !seen[d]) {
                seen[d] = 1;
                dbc.push_back(d);
            }
        }
    }
    upload(op.diag_bc_ids, dbc, s, op.device_bytes);
    op.n_diag_bc = (uint32_t)dbc.size();

    // ---- element diagonals into ve (FREE: computed from the coordinates,
    // mat_get_diagonal_ghosted_free)
    for (auto& g : op.groups) {
        if (free_m)
            launch_free(op.elem_kind, false, false, true, free_args(op, g, nullptr, op.ve.p, 0, g.n_slots, nullptr), s);
        else
            launch_extract_diag(g.K[0].as<double>(), g.L[0], g.n_slots, g.ve_rs, g.ve_ss, op.ve.as<double>() + g.ve_base,
                                s);
    }

    // ---- trace in element order (finalize_end :748-785)
    double tr = 0.0;
    if (free_m) {
        std::vector<double> hve(ve_rows);
        if (ve_rows)
            AMX_HIP(hipMemcpyAsync(hve.data(), op.ve.p, ve_rows * sizeof(double), hipMemcpyDeviceToHost, s));
        AMX_HIP(hipStreamSynchronize(s));
        for (uint32_t e = 0; e < op.n_elem; e++) {
            const Group& g = op.groups[op.elem_group[e]];
            const uint64_t base = g.ve_base + (uint64_t)g.elem_slot[op.elem_index[e]] * g.ve_ss;
            for (uint32_t r = 0; r < g.n; r++) {
                const double dv = hve[base + (uint64_t)r * g.ve_rs];
                // the diagonal kernel marks a non-positive Jacobian with NaN
                AMX_REQUIRE(std::isfinite(dv), AMX_INVALID_ARGUMENT,
                            "element " + std::to_string(e) + ": non-positive Jacobian at a quadrature point");
                tr += dv;
            }
        }
    } else {
        for (uint32_t e = 0; e < op.n_elem; e++) {
            const Group& g = op.groups[op.elem_group[e]];
            const uint32_t n = g.n, i = op.elem_index[e];
            for (uint32_t r = 0; r < n; r++)
                tr += op.scaled ? op.scale[e] * op.kref[(size_t)r * n + r]
                                : g.stage[(size_t)i * n * n + (size_t)r * n + r];
        }
    }
    op.trace = tr;

    // ---- assembled diagonal (mat_get_diagonal_ghosted :997-1056), same
    // element order as the gather, and the penalty coefficients p - diag
    op.diag.alloc(std::max<uint32_t>(op.n_total, 1) * sizeof(double));
    op.device_bytes += op.diag.bytes;
    launch_gather(true, 0, op.off.as<uint32_t>(), op.pos.as<uint32_t>(), op.ve.p, nullptr, nullptr, nullptr,
                  op.diag.p, op.n_total, s);
    op.coef64.alloc(std::max<uint32_t>(op.n_total, 1) * sizeof(double));
    op.device_bytes += op.coef64.bytes;
    launch_fill(op.coef64.p, op.n_total, true, 0.0, s);
    launch_coef(op.diag.as<double>(), op.cdofs.as<uint32_t>(), op.n_cdofs, op.penalty(), op.coef64.as<double>(), s);

    // ---- halo accumulation lists: targets in first-appearance order, each
    // summing its received entries in the reference's loop order (:1662-1669)
    op.n_halo_unique = 0;
    if (!op.send_ids.empty()) {
        std::vector<uint32_t> first(op.n_total, kPadSlot), dst, cnt;
        for (uint32_t i = 0; i < op.send_ids.size(); i++) {
            const uint32_t d = op.send_ids[i];
            if (first[d] == kPadSlot) {
                first[d] = (uint32_t)dst.size();
                dst.push_back(d);
                cnt.push_back(0);
            }
            cnt[first[d]]++;
        }
        std::vector<uint32_t> hoff2(dst.size() + 1, 0), src(op.send_ids.size());
        for (size_t i = 0; i < dst.size(); i++) hoff2[i + 1] = hoff2[i] + cnt[i];
        std::vector<uint32_t> f2(hoff2.begin(), hoff2.end() - 1);
        for (uint32_t i = 0; i < op.send_ids.size(); i++) src[f2[first[op.send_ids[i]]]++] = i;
        upload(op.halo_dst, dst, s, op.device_bytes);
        upload(op.halo_off, hoff2, s, op.device_bytes);
        upload(op.halo_src, src, s, op.device_bytes);
        op.n_halo_unique = (uint32_t)dst.size();
        AMX_HIP(hipStreamSynchronize(s));  // host vectors leave scope
    }
    AMX_HIP(hipStreamSynchronize(s));
    op.diag_assembled = op.size == 1;
    if (op.size > 1 && op.halo_fn) assemble_diag(op, s);
    op.t_scatter = op.t_mv = op.t_gather = op.t_total = 0.0;
    op.finalized = true;
}

void ensure_scratch(Op& op)
{
    const size_t b = std::max<uint32_t>(op.n_total, 1) * sizeof(double);
    if (op.scratch_u.bytes < b) op.scratch_u.alloc(b);
    if (op.scratch_v.bytes < b) op.scratch_v.alloc(b);
    for (hipEvent_t& e : op.tev)
        if (!e) AMX_HIP(hipEventCreate(&e));
}

// Phases of a host-array matvec, accumulated as aMatFree::get_timer reports
// them (include/aMatFree.hpp:239-245; the GPU path's timers in
// matvec_ghosted_gpuPure, :3644-3700): u copied in ~ scatter_time (pack +
// host-to-device), the device application ~ mv_time (the batched GEMV), v
// copied out ~ gather_time (device-to-host + host accumulation), the whole
// call on the host clock ~ mvTotal_time.
struct HostCallTimer {
    Op& op;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit HostCallTimer(Op& o) : op(o) {}
    void mark(int i, hipStream_t s) { AMX_HIP(hipEventRecord(op.tev[i], s)); }
    void finish()  // after the stream has synchronized
    {
        float a = 0, b = 0, c = 0;
        AMX_HIP(hipEventElapsedTime(&a, op.tev[0], op.tev[1]));
        AMX_HIP(hipEventElapsedTime(&b, op.tev[1], op.tev[2]));
        AMX_HIP(hipEventElapsedTime(&c, op.tev[2], op.tev[3]));
        op.t_scatter += 1e-3 * a;
        op.t_mv += 1e-3 * b;
        op.t_gather += 1e-3 * c;
        op.t_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// write one element block into the staging copy (copy_element_matrix :882-953)
void stage_block(Op& op, uint32_t eid, const double* m, int layout, uint32_t bi, uint32_t bj, uint32_t bd)
{
    AMX_REQUIRE(eid < op.n_elem, AMX_INDEX_OUT_OF_BOUNDS, "element id " + std::to_string(eid) + " out of range");
    AMX_REQUIRE(bd >= 1 && bi < bd && bj < bd, AMX_INDEX_OUT_OF_BOUNDS, "block index out of range");
    AMX_REQUIRE(m != nullptr, AMX_INVALID_ARGUMENT, "element matrix is NULL");
    AMX_REQUIRE(layout == AMX_ROW_MAJOR || layout == AMX_COL_MAJOR, AMX_INVALID_ARGUMENT, "unknown layout");
    AMX_REQUIRE(!op.scaled, AMX_INVALID_STATE, "element matrices were given in scaled form");
    Group& g = op.groups[op.elem_group[eid]];
    const uint32_t n = g.n;
    AMX_REQUIRE(n % bd == 0, AMX_INVALID_ARGUMENT, "blocks_dim does not divide the element's DoF count");
    const uint32_t nb = n / bd;
    if (g.stage.empty()) g.stage.assign((size_t)g.elems.size() * n * n, 0.0);
    double* ke = g.stage.data() + (size_t)op.elem_index[eid] * n * n;
    if (op.elem_bdim[eid] != bd) {  // a new block structure drops the old blocks (:890-899)
        std::fill(ke, ke + (size_t)n * n, 0.0);
        op.elem_bdim[eid] = bd;
    }
    for (uint32_t r = 0; r < nb; r++)
        for (uint32_t c = 0; c < nb; c++)
            ke[(size_t)(bi * nb + r) * n + bj * nb + c] = layout == AMX_ROW_MAJOR ? m[(size_t)r * nb + c] : m[(size_t)c * nb + r];
    op.finalized = false;
}

}  // namespace

void ebe_launch(Op& op, KType kt, bool vt_double, bool masked, int phase, const void* u, void* v, hipStream_t s,
                const uint32_t* skip, void* ve)
{
    if (!ve) ve = op.ve.p;
    const bool atomic = op.scatter == AMX_SCATTER_ATOMIC;
    if (op.is_free()) {
        // FP64, or FP32 arithmetic and vectors (the inner solves of the
        // mixed-precision solvers; nothing is stored per precision)
        AMX_REQUIRE((kt == KType::F64 && vt_double) || (kt == KType::F32 && !vt_double), AMX_NOT_IMPLEMENTED,
                    "method 2 (FREE) runs in FP64 or FP32");
        for (auto& g : op.groups) {
            uint32_t t0 = 0, nt = g.n_tiles;
            if (phase == 0) nt = g.tiles_phase0;
            else if (phase == 1) { t0 = g.tiles_phase0; nt = g.n_tiles - t0; }
            launch_free(op.elem_kind, masked, atomic, false,
                        free_args(op, g, u, atomic ? v : ve, t0 * g.G, nt * g.G, skip, !vt_double), s);
        }
        return;
    }
    for (auto& g : op.groups) {
        const int k = (int)kt;
        AMX_REQUIRE(g.K[k].p != nullptr, AMX_INVALID_STATE, "requested precision was not built at finalize");
        EbeArgs a;
        a.K = g.K[k].p;
        a.map = g.map.as<uint32_t>();
        a.u = u;
        a.out = atomic ? v : ve;
        a.ve_base = g.ve_base;
        a.ve_rs = g.ve_rs;
        a.ve_ss = g.ve_ss;
        a.skip = skip;
        a.kdiag = kt == KType::BF16 ? g.kdiag16.as<float>() : nullptr;
        a.nsym_xs = k == 0 ? g.nsym_xs.as<uint32_t>() : g.psx_tab[k].as<uint32_t>();
        a.nsym_colany = g.nsym_colany;
        if (phase == 0) { a.tile_begin = 0; a.n_tiles = g.tiles_phase0; }
        else if (phase == 1) { a.tile_begin = g.tiles_phase0; a.n_tiles = g.n_tiles - g.tiles_phase0; }
        else { a.tile_begin = 0; a.n_tiles = g.n_tiles; }
        launch_ebe(g.L[k], kt, vt_double, masked, atomic, a, s);
    }
}

void gather_launch(Op& op, bool vt_double, int bc, const void* u, void* v, hipStream_t s, const uint32_t* skip,
                   const void* ve)
{
    launch_gather(vt_double, bc, op.off.as<uint32_t>(), op.pos.as<uint32_t>(), ve ? ve : op.ve.p, u,
                  op.flag.as<uint8_t>(), op.coef64.as<double>(), v, op.n_total, s, skip);
}

void apply(Op& op, KType kt, bool vt_double, int bc, const void* u, void* v, hipStream_t s, const uint32_t* skip,
           void* ve)
{
    const bool masked = bc == AMX_BCOP_IDENTITY;
    if (op.scatter == AMX_SCATTER_ATOMIC) {
        launch_fill(v, op.n_total, vt_double, 0.0, s);
        ebe_launch(op, kt, vt_double, masked, 2, u, v, s, skip, ve);
        launch_bc_fixup(vt_double, bc, op.cdofs.as<uint32_t>(), op.n_cdofs, op.coef64.as<double>(), u, v, s);
    } else {
        ebe_launch(op, kt, vt_double, masked, 2, u, v, s, skip, ve);
        gather_launch(op, vt_double, bc, u, v, s, skip, ve);
    }
}

// v = A u with BC form bc (FP64) on the handle's share: one rank as apply();
// several ranks with the halo exchanges around the dependent elements, the
// ghost accumulation, then the constrained rows (matvec_ghosted_OMP's
// ghost_receive / ghost_send bracketing, :1795, 2016, 2234-2235)
void apply_full(Op& op, int bc, double* u, double* v, hipStream_t s)
{
    if (op.size == 1) {
        apply(op, KType::F64, true, bc, u, v, s);
        return;
    }
    require_exchange(op);
    const bool masked = bc == AMX_BCOP_IDENTITY;
    const bool atomic = op.scatter == AMX_SCATTER_ATOMIC;
    if (atomic) launch_fill(v, op.n_total, true, 0.0, s);
    ebe_launch(op, KType::F64, true, masked, 0, u, v, s);
    exchange_u(op, u, s);
    ebe_launch(op, KType::F64, true, masked, 1, u, v, s);
    if (!atomic) gather_launch(op, true, AMX_BCOP_NONE, u, v, s);
    exchange_v(op, v, s);
    launch_bc_fixup(true, bc, op.cdofs.as<uint32_t>(), op.n_cdofs, op.coef64.as<double>(), u, v, s);
}

}  // namespace amx

using amx::DeviceGuard;

extern "C" {

int amx_op_create(const amx_maps* maps, int bc_meth, int device, amx_op** out)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && out, AMX_INVALID_ARGUMENT, "NULL argument");
        AMX_REQUIRE(maps->has_map, AMX_INVALID_STATE, "maps: set_map has not run");
        AMX_REQUIRE(bc_meth == AMX_BC_IMATRIX || bc_meth == AMX_BC_PENALTY, AMX_UNKNOWN_BC_METH, "unknown BC method");
        int ndev = 0;
        AMX_HIP(hipGetDeviceCount(&ndev));
        AMX_REQUIRE(device >= 0 && device < ndev, AMX_INVALID_ARGUMENT,
                    "device " + std::to_string(device) + " not present (" + std::to_string(ndev) + " visible)");
        std::unique_ptr<amx_op> h(new amx_op());
        Op& op = h->op;
        op.device = device;
        op.bc_meth = bc_meth;
        op.rank = maps->rank;
        op.size = maps->size;
        op.n_elem = maps->n_elem;
        op.n_total = maps->n_total;
        op.n_owned = maps->n_owned;
        op.n_pre = maps->n_pre;
        op.n_post = maps->n_post;
        op.dof_lo = maps->dof_scan[maps->rank];
        op.dpe = maps->dpe;
        op.eoff = maps->eoff;
        op.local_map = maps->local_map;
        op.dependent.assign(op.n_elem, 0);
        for (uint32_t e : maps->dependent) op.dependent[e] = 1;
        op.indep_list = maps->independent;
        op.dep_list = maps->dependent;
        op.has_bdr = maps->has_bdr;
        if (op.has_bdr) {
            op.bdr_map = maps->bdr_map;
            op.presval_map = maps->presval_map;
            for (size_t i = 0; i < maps->owned_cdofs.size(); i++) {
                op.owned_c_local.push_back((uint32_t)(maps->owned_cdofs[i] - op.dof_lo) + op.n_pre);
                op.owned_c_vals.push_back(maps->owned_cvals[i]);
            }
            for (uint64_t gid : maps->owned_free) op.owned_free_local.push_back((uint32_t)(gid - op.dof_lo) + op.n_pre);
        }
        op.has_peer_lists = maps->size > 1 && maps->has_peer_lists;
        if (op.has_peer_lists) op.send_ids = maps->send_ids;
        // element groups by DoF count
        op.elem_group.assign(op.n_elem, -1);
        op.elem_index.assign(op.n_elem, -1);
        op.elem_bdim.assign(op.n_elem, 0);
        for (uint32_t e = 0; e < op.n_elem; e++) {
            int gi = -1;
            for (size_t k = 0; k < op.groups.size(); k++)
                if (op.groups[k].n == op.dpe[e]) gi = (int)k;
            if (gi < 0) {
                op.groups.emplace_back();
                op.groups.back().n = op.dpe[e];
                gi = (int)op.groups.size() - 1;
            }
            op.elem_group[e] = gi;
            op.elem_index[e] = (int32_t)op.groups[gi].elems.size();
            op.groups[gi].elems.push_back(e);
        }
        DeviceGuard dg(device);
        AMX_HIP(hipStreamCreateWithFlags(&op.stream, hipStreamNonBlocking));
        *out = h.release();
    });
}

int amx_op_destroy(amx_op* h)
{
    if (!h) return AMX_SUCCESS;
    {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(h->op.device);
        if (h->op.stream) (void)hipStreamSynchronize(h->op.stream);
        if (h->op.stream) (void)hipStreamDestroy(h->op.stream);
        h->op.stream = nullptr;
        delete h;  // device buffers are freed with the device still current
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    return AMX_SUCCESS;
}

int amx_op_set_matfree_type(amx_op* h, int type)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        if (type == AMX_MATFREE_QUASI_FREE)
            fail(AMX_NOT_IMPLEMENTED, "method 1 is the reference's CPU loop; this library runs 2 to 5");
        AMX_REQUIRE(type >= AMX_MATFREE_FREE && type <= AMX_MATFREE_GPU_OVER_GPU, AMX_UNKNOWN_MAT_TYPE,
                    "unknown matrix-free type " + std::to_string(type));
        if (type != h->op.matfree_type && (type == AMX_MATFREE_FREE || h->op.is_free())) h->op.finalized = false;
        h->op.matfree_type = type;
    });
}

int amx_op_set_num_streams(amx_op* h, uint32_t n)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        AMX_REQUIRE(n >= 1, AMX_INVALID_ARGUMENT, "need at least one stream");
        h->op.n_streams = n;
    });
}

int amx_op_set_element_matrix(amx_op* h, uint32_t eid, const double* e_mat, int layout, uint32_t block_i,
                              uint32_t block_j, uint32_t blocks_dim)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        amx::stage_block(h->op, eid, e_mat, layout, block_i, block_j, blocks_dim);
    });
}

int amx_op_set_element_matrices(amx_op* h, uint32_t first_eid, uint32_t count, const double* e_mats, int layout)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        Op& op = h->op;
        AMX_REQUIRE((uint64_t)first_eid + count <= op.n_elem, AMX_INDEX_OUT_OF_BOUNDS, "element range out of bounds");
        uint64_t off = 0;
        for (uint32_t e = first_eid; e < first_eid + count; e++) {
            amx::stage_block(op, e, e_mats + off, layout, 0, 0, 1);
            off += (uint64_t)op.dpe[e] * op.dpe[e];
        }
    });
}

int amx_op_set_element_matrices_scaled(amx_op* h, const double* k_ref, int layout, const double* scale)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && k_ref && scale, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        AMX_REQUIRE(op.groups.size() == 1, AMX_INVALID_ARGUMENT, "scaled form needs one DoF count for all elements");
        const uint32_t n = op.groups[0].n;
        op.kref.resize((size_t)n * n);
        for (uint32_t r = 0; r < n; r++)
            for (uint32_t c = 0; c < n; c++)
                op.kref[(size_t)r * n + c] = layout == AMX_ROW_MAJOR ? k_ref[(size_t)r * n + c] : k_ref[(size_t)c * n + r];
        op.scale.assign(scale, scale + op.n_elem);
        op.scaled = true;
        op.groups[0].stage.clear();
        op.groups[0].stage.shrink_to_fit();
        std::fill(op.elem_bdim.begin(), op.elem_bdim.end(), 1u);
        op.finalized = false;
    });
}

int amx_op_set_element_function(amx_op* h, int kind, const double* xe, const double* scale, double E, double nu,
                                uint32_t ngt)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && xe, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        AMX_REQUIRE(kind == AMX_ELEM_HEX8_LAPLACE || kind == AMX_ELEM_HEX20_ELASTIC || kind == AMX_ELEM_TET10_LAPLACE,
                    AMX_UNKNOWN_ELEMENT_TYPE, "unknown element kind " + std::to_string(kind));
        const uint32_t nodes = amx::free_nodes(kind), ndof = kind == AMX_ELEM_HEX20_ELASTIC ? 60u : nodes;
        for (uint32_t e = 0; e < op.n_elem; e++)
            AMX_REQUIRE(op.dpe[e] == ndof, AMX_INVALID_ARGUMENT,
                        "element " + std::to_string(e) + " has " + std::to_string(op.dpe[e]) + " DoFs; the kind has " +
                            std::to_string(ndof));
        uint32_t nq = 0;
        std::vector<double> tab = amx::free_quadrature(kind, ngt, nq, 0);
        if (kind == AMX_ELEM_HEX20_ELASTIC) {
            AMX_REQUIRE(nq <= 64, AMX_NOT_IMPLEMENTED, "hex20: at most 64 quadrature points (ngt <= 4)");
            AMX_REQUIRE(E > 0.0 && nu > -1.0 && nu < 0.5, AMX_INVALID_ARGUMENT, "need E > 0 and -1 < nu < 0.5");
            op.elem_c[0] = E * (1.0 - nu) / (1.0 + nu) / (1.0 - 2.0 * nu);
            op.elem_c[1] = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);
            op.elem_c[2] = 0.5 * E / (1.0 + nu);
        }
        op.elem_kind = kind;
        op.elem_ngt = ngt;
        op.elem_nq = nq;
        op.elem_xe.assign(xe, xe + (size_t)op.n_elem * 3 * nodes);
        if (scale) op.elem_scale.assign(scale, scale + op.n_elem);
        else op.elem_scale.assign(op.n_elem, 1.0);
        op.matfree_type = AMX_MATFREE_FREE;
        op.finalized = false;
    });
}

int amx_op_set_options(amx_op* h, int precision_mask, int scatter_mode)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        AMX_REQUIRE((precision_mask & ~(AMX_PREC_FP64 | AMX_PREC_FP32 | AMX_PREC_BF16)) == 0, AMX_INVALID_ARGUMENT,
                    "unknown precision bits");
        AMX_REQUIRE(scatter_mode == AMX_SCATTER_DETERMINISTIC || scatter_mode == AMX_SCATTER_ATOMIC,
                    AMX_INVALID_ARGUMENT, "unknown scatter mode");
        const int mask = precision_mask | AMX_PREC_FP64;
        if (mask != h->op.prec_mask) h->op.finalized = false;
        h->op.prec_mask = mask;
        h->op.scatter = scatter_mode;
    });
}

int amx_op_set_symmetric_storage(amx_op* h, int mode)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        AMX_REQUIRE(mode == 0 || mode == 1, AMX_INVALID_ARGUMENT, "mode: 0 (full tiles) or 1 (packed when symmetric)");
        if (mode != h->op.sym_mode) h->op.finalized = false;
        h->op.sym_mode = mode;
    });
}

int amx_op_set_penalty(amx_op* h, double p)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        AMX_REQUIRE(std::isfinite(p) && p > 0.0, AMX_INVALID_ARGUMENT, "penalty must be positive and finite");
        Op& op = h->op;
        op.penalty_set = true;
        op.penalty_value = p;
        if (op.finalized) {
            DeviceGuard dg(op.device);
            amx::launch_coef(op.diag.as<double>(), op.cdofs.as<uint32_t>(), op.n_cdofs, op.penalty(),
                             op.coef64.as<double>(), op.stream);
            AMX_HIP(hipStreamSynchronize(op.stream));
        }
    });
}

int amx_op_set_bc_operator(amx_op* h, int mode)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        AMX_REQUIRE(mode >= AMX_BCOP_NONE && mode <= AMX_BCOP_PENALTY_DIAG, AMX_INVALID_ARGUMENT, "unknown BC form");
        h->op.bc_operator = mode;
    });
}

int amx_op_finalize(amx_op* h)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        amx::finalize_op(h->op);
    });
}

int amx_op_get_info(const amx_op* h, amx_op_info* info)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && info, AMX_INVALID_ARGUMENT, "NULL argument");
        const Op& op = h->op;
        std::memset(info, 0, sizeof(*info));
        info->n_elems = op.n_elem;
        info->n_total = op.n_total;
        info->n_owned = op.n_owned;
        info->n_pre_ghost = op.n_pre;
        info->n_groups = (uint32_t)op.groups.size();
        info->finalized = op.finalized;
        info->matfree_type = op.matfree_type;
        info->bc_meth = op.bc_meth;
        info->bc_operator = op.bc_operator;
        info->scatter_mode = op.scatter;
        info->precision_mask = op.prec_mask;
        info->trace = op.trace;
        info->penalty = op.penalty();
        info->device_bytes = op.device_bytes;
        for (const auto& g : op.groups) {
            const uint64_t ne = g.elems.size(), nn = (uint64_t)g.n * g.n;
            if (op.is_free()) {  // coordinates, factor and map read; no element matrices
                info->alg_bytes_fp64 += ne * (24ull * amx::free_nodes(op.elem_kind) + 8 + 4ull * g.n);
                info->flops += ne * amx::free_flops_per_elem(op.elem_kind, op.elem_nq);
                continue;
            }
            // values stored per element: the upper triangle of packed symmetric
            // tiles; near-symmetric records (FP64 only) by their record size
            const uint64_t ns = g.L[0].sym ? (uint64_t)g.n * (g.n + 1) / 2 : nn;
            const uint64_t b64 = g.L[0].nsym ? amx::nsym_record_bytes(g.n, g.L[0].nx) : ns * 8;
            // FP32 / BF16: the records as stored when built, else what they would hold
            auto bytes_of = [&](int kt) {
                return g.L[kt].nsym ? amx::tile_values(g.L[kt], (KType)kt) * amx::ktype_size((KType)kt)
                                    : ns * amx::ktype_size((KType)kt);
            };
            if (g.L[0].sym) info->packed_symmetric++;
            if (g.L[0].nsym) info->near_symmetric++;
            info->k_bytes_fp64 += ne * b64;
            info->k_bytes_fp32 += ne * bytes_of(1);
            info->k_bytes_bf16 += ne * bytes_of(2);
            info->alg_bytes_fp64 += ne * (b64 + 4ull * g.n);
            info->flops += 2 * ne * nn;
        }
        info->alg_bytes_fp64 += 16ull * op.n_total;
    });
}

int amx_op_set_halo_transport(amx_op* h, amx_halo_fn fn, void* ctx, int mode)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        Op& op = h->op;
        if (!fn) {
            op.halo_fn = nullptr;
            op.halo_ctx = nullptr;
            return;
        }
        AMX_REQUIRE(mode == AMX_HALO_HOST || mode == AMX_HALO_DEVICE, AMX_INVALID_ARGUMENT, "unknown transport mode");
        AMX_REQUIRE(op.size > 1, AMX_INVALID_STATE, "one rank has no halo to exchange");
        AMX_REQUIRE(op.has_peer_lists, AMX_INVALID_STATE, "maps: peer requests have not been set");
        DeviceGuard dg(op.device);
        const size_t b = std::max<uint32_t>(std::max(op.n_send(), op.n_ghost()), 1) * sizeof(double);
        if (!op.halo_send_ids.p) amx::upload(op.halo_send_ids, op.send_ids, op.stream, op.device_bytes);
        if (op.hx_send.bytes < b) {
            op.hx_send.alloc(b);
            op.hx_recv.alloc(b);
            op.device_bytes += 2 * b;
        }
        if (mode == AMX_HALO_HOST && !op.hh_send) {
            AMX_HIP(hipHostMalloc((void**)&op.hh_send, b, hipHostMallocDefault));
            AMX_HIP(hipHostMalloc((void**)&op.hh_recv, b, hipHostMallocDefault));
        }
        AMX_HIP(hipStreamSynchronize(op.stream));
        op.halo_fn = fn;
        op.halo_ctx = ctx;
        op.halo_mode = mode;
        if (op.finalized && !op.diag_assembled) amx::assemble_diag(op, op.stream);
    });
}

int amx_op_matvec_dev(amx_op* h, double* v, const double* u, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        amx::require_exchange(op);
        DeviceGuard dg(op.device);
        amx::apply_full(op, AMX_BCOP_NONE, const_cast<double*>(u), v, op.pick(stream));
    });
}

int amx_op_matmult_dev(amx_op* h, double* v, const double* u, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        amx::require_exchange(op);
        DeviceGuard dg(op.device);
        amx::apply_full(op, op.bc_operator, const_cast<double*>(u), v, op.pick(stream));
    });
}

int amx_op_matmult_f32_dev(amx_op* h, int precision, float* v, const float* u, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        const KType kt = amx::ktype_of(precision);
        AMX_REQUIRE(kt != KType::F64, AMX_INVALID_ARGUMENT, "FP32 vectors take FP32 or BF16 storage");
        AMX_REQUIRE(op.has(precision), AMX_INVALID_STATE, "precision was not requested before finalize");
        AMX_REQUIRE(op.size == 1, AMX_NOT_IMPLEMENTED, "mixed-precision operator is one-rank");
        DeviceGuard dg(op.device);
        amx::apply(op, kt, false, op.bc_operator, u, v, op.pick(stream));
    });
}

int amx_op_matvec_begin_dev(amx_op* h, double* v, const double* u, int apply_bc, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        if (op.scatter == AMX_SCATTER_ATOMIC) amx::launch_fill(v, op.n_total, true, 0.0, s);
        amx::ebe_launch(op, KType::F64, true, apply_bc && op.bc_operator == AMX_BCOP_IDENTITY, 0, u, v, s);
    });
}

int amx_op_matvec_end_dev(amx_op* h, double* v, const double* u, int apply_bc, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        const bool masked = apply_bc && op.bc_operator == AMX_BCOP_IDENTITY;
        amx::ebe_launch(op, KType::F64, true, masked, 1, u, v, s);
        // one rank: BC rows now (fused into the gather); several: after the
        // accumulation (bc_fixup)
        const int bc = (apply_bc && op.size == 1) ? op.bc_operator : AMX_BCOP_NONE;
        if (op.scatter == AMX_SCATTER_ATOMIC)
            amx::launch_bc_fixup(true, bc, op.cdofs.as<uint32_t>(), op.n_cdofs, op.coef64.as<double>(), u, v, s);
        else
            amx::gather_launch(op, true, bc, u, v, s);
    });
}

int amx_op_halo_pack_dev(amx_op* h, const double* u, double* send, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        AMX_REQUIRE(op.size == 1 || op.has_peer_lists, AMX_INVALID_STATE, "maps: peer requests have not been set");
        AMX_REQUIRE(send || op.n_send() == 0, AMX_INVALID_ARGUMENT, "send is NULL");
        if (op.n_send() == 0) return;
        DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        if (!op.halo_send_ids.p) {
            amx::upload(op.halo_send_ids, op.send_ids, op.stream, op.device_bytes);
            AMX_HIP(hipStreamSynchronize(op.stream));
        }
        amx::launch_pack(send, u, op.halo_send_ids.as<uint32_t>(), op.n_send(), s);
    });
}

int amx_op_halo_import_dev(amx_op* h, double* u, const double* halo, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        AMX_REQUIRE(halo || op.n_ghost() == 0, AMX_INVALID_ARGUMENT, "halo is NULL");
        DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        if (op.n_pre) AMX_HIP(hipMemcpyAsync(u, halo, op.n_pre * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (op.n_post)
            AMX_HIP(hipMemcpyAsync(u + op.n_pre + op.n_owned, halo + op.n_pre, op.n_post * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
        if (op.bc_operator == AMX_BCOP_IDENTITY)
            amx::launch_set_entries(u, op.ghost_c_ids.as<uint32_t>(), op.n_ghost_c, 0.0, s);
    });
}

int amx_op_halo_export_dev(amx_op* h, const double* v, double* halo, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && halo, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        if (op.n_pre) AMX_HIP(hipMemcpyAsync(halo, v, op.n_pre * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (op.n_post)
            AMX_HIP(hipMemcpyAsync(halo + op.n_pre, v + op.n_pre + op.n_owned, op.n_post * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
    });
}

int amx_op_halo_accumulate_dev(amx_op* h, double* v, const double* recv, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        AMX_REQUIRE(recv || op.n_halo_unique == 0, AMX_INVALID_ARGUMENT, "recv is NULL");
        DeviceGuard dg(op.device);
        amx::launch_halo_accumulate(v, recv, op.halo_dst.as<uint32_t>(), op.halo_off.as<uint32_t>(),
                                    op.halo_src.as<uint32_t>(), op.n_halo_unique, op.pick(stream));
    });
}

int amx_op_bc_fixup_dev(amx_op* h, double* v, const double* u, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        DeviceGuard dg(op.device);
        amx::launch_bc_fixup(true, op.bc_operator, op.cdofs.as<uint32_t>(), op.n_cdofs, op.coef64.as<double>(), u, v,
                             op.pick(stream));
    });
}

int amx_op_matvec(amx_op* h, double* v, const double* u, int is_ghosted)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        amx::require_exchange(op);
        DeviceGuard dg(op.device);
        amx::ensure_scratch(op);
        hipStream_t s = op.stream;
        double* du = op.scratch_u.as<double>();
        double* dv = op.scratch_v.as<double>();
        amx::HostCallTimer tm(op);
        tm.mark(0, s);
        if (is_ghosted) {
            AMX_HIP(hipMemcpyAsync(du, u, op.n_total * sizeof(double), hipMemcpyHostToDevice, s));
        } else {  // local_to_ghost (:824-838): ghost slots are 0
            if (op.n_ghost()) amx::launch_fill(du, op.n_total, true, 0.0, s);
            AMX_HIP(hipMemcpyAsync(du + op.n_pre, u, op.n_owned * sizeof(double), hipMemcpyHostToDevice, s));
        }
        tm.mark(1, s);
        amx::apply_full(op, AMX_BCOP_NONE, du, dv, s);
        tm.mark(2, s);
        if (is_ghosted) AMX_HIP(hipMemcpyAsync(v, dv, op.n_total * sizeof(double), hipMemcpyDeviceToHost, s));
        else AMX_HIP(hipMemcpyAsync(v, dv + op.n_pre, op.n_owned * sizeof(double), hipMemcpyDeviceToHost, s));
        tm.mark(3, s);
        AMX_HIP(hipStreamSynchronize(s));
        tm.finish();
    });
}

int amx_op_matmult_mf(amx_op* h, double* vv, const double* uu)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && vv && uu, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        amx::require_exchange(op);
        DeviceGuard dg(op.device);
        amx::ensure_scratch(op);
        hipStream_t s = op.stream;
        double* du = op.scratch_u.as<double>();
        double* dv = op.scratch_v.as<double>();
        amx::HostCallTimer tm(op);
        tm.mark(0, s);
        // local_to_ghost (:3913): the ghost slots are filled by the exchange
        if (op.n_ghost()) amx::launch_fill(du, op.n_total, true, 0.0, s);
        AMX_HIP(hipMemcpyAsync(du + op.n_pre, uu, op.n_owned * sizeof(double), hipMemcpyHostToDevice, s));
        tm.mark(1, s);
        amx::apply_full(op, op.bc_operator, du, dv, s);
        tm.mark(2, s);
        AMX_HIP(hipMemcpyAsync(vv, dv + op.n_pre, op.n_owned * sizeof(double), hipMemcpyDeviceToHost, s));
        tm.mark(3, s);
        AMX_HIP(hipStreamSynchronize(s));
        tm.finish();
    });
}

int amx_op_get_timer(amx_op* h, double* scatter_s, double* gather_s, double* mv_s, double* total_s)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && scatter_s && gather_s && mv_s && total_s, AMX_INVALID_ARGUMENT, "NULL argument");
        const Op& op = h->op;
        amx::require_finalized(op);
        *scatter_s = op.t_scatter;
        *gather_s = op.t_gather;
        *mv_s = op.t_mv;
        *total_s = op.t_total;
    });
}

int amx_op_get_diagonal_dev(amx_op* h, double* diag, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && diag, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        amx::require_exchange(op);  // the owned rows hold the assembled diagonal
        DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        AMX_HIP(hipMemcpyAsync(diag, op.diag.p, op.n_total * sizeof(double), hipMemcpyDeviceToDevice, s));
        // apply_bc_diagonal (:4309-4337): 1 for BC_IMATRIX, the penalty for BC_PENALTY
        const double val = op.bc_meth == AMX_BC_IMATRIX ? 1.0 : op.penalty();
        amx::launch_set_entries(diag, op.diag_bc_ids.as<uint32_t>(), op.n_diag_bc, val, s);
    });
}

int amx_op_get_diagonal(amx_op* h, double* diag)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && diag, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        DeviceGuard dg(op.device);
        amx::ensure_scratch(op);
        int rc = amx_op_get_diagonal_dev(h, op.scratch_v.as<double>(), op.stream);
        if (rc != AMX_SUCCESS) fail(rc, "get_diagonal_dev failed");
        AMX_HIP(hipMemcpyAsync(diag, op.scratch_v.as<double>() + op.n_pre, op.n_owned * sizeof(double),
                               hipMemcpyDeviceToHost, op.stream));
        AMX_HIP(hipStreamSynchronize(op.stream));
    });
}

int amx_op_apply_bc_rhs(amx_op* h, double* rhs)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && rhs, AMX_INVALID_ARGUMENT, "NULL argument");
        Op& op = h->op;
        amx::require_finalized(op);
        AMX_REQUIRE(op.has_bdr, AMX_INVALID_STATE, "no boundary map");
        DeviceGuard dg(op.device);
        amx::ensure_scratch(op);
        if (op.bc_meth == AMX_BC_IMATRIX) {
            amx::require_exchange(op);
            // K_fc U_c: the operator on the vector holding the prescribed values
            // at constrained DoFs (aMatBased::apply_bc_rhs :364-400 form of
            // aMatFree::apply_bc_rhs :4556-4648)
            std::vector<double> uc(op.n_total, 0.0);
            for (uint64_t k = 0; k < op.local_map.size(); k++)
                if (op.bdr_map[k] == 1) uc[op.local_map[k]] = op.presval_map[k];
            std::vector<double> w(op.n_total);
            hipStream_t s = op.stream;
            AMX_HIP(hipMemcpyAsync(op.scratch_u.p, uc.data(), op.n_total * sizeof(double), hipMemcpyHostToDevice, s));
            // the product of this rank's elements; the ghost rows reach their
            // owners (the reference's ADD_VALUES into KfcUcVec, :4640-4646);
            // uc keeps this rank's prescribed values in its ghost slots
            amx::apply(op, KType::F64, true, AMX_BCOP_NONE, op.scratch_u.p, op.scratch_v.p, s);
            if (op.size > 1) amx::exchange_v(op, op.scratch_v.as<double>(), s);
            AMX_HIP(hipMemcpyAsync(w.data(), op.scratch_v.p, op.n_total * sizeof(double), hipMemcpyDeviceToHost, s));
            AMX_HIP(hipStreamSynchronize(s));
            for (uint32_t d : op.owned_free_local) rhs[d - op.n_pre] += -1.0 * w[d];
            for (size_t i = 0; i < op.owned_c_local.size(); i++) rhs[op.owned_c_local[i] - op.n_pre] = op.owned_c_vals[i];
        } else {
            const double p = op.penalty();
            for (size_t i = 0; i < op.owned_c_local.size(); i++)
                rhs[op.owned_c_local[i] - op.n_pre] = p * op.owned_c_vals[i];
        }
    });
}

int amx_op_time_kernels_dev(amx_op* h, int n_iter, double* v, const double* u, double* ms_element,
                            double* ms_gather, double* ms_total, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && v && u && n_iter > 0, AMX_INVALID_ARGUMENT, "bad argument");
        Op& op = h->op;
        amx::require_finalized(op);
        AMX_REQUIRE(op.size == 1, AMX_NOT_IMPLEMENTED, "timers run the one-rank operator");
        DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        struct Events {  // destroyed on every exit, errors included
            std::vector<hipEvent_t> v;
            ~Events()
            {
                for (hipEvent_t e : v) (void)hipEventDestroy(e);
            }
        } evs;
        evs.v.reserve(3 * (size_t)n_iter);
        for (int i = 0; i < 3 * n_iter; i++) {
            hipEvent_t e = nullptr;
            AMX_HIP(hipEventCreate(&e));
            evs.v.push_back(e);
        }
        std::vector<hipEvent_t>& ev = evs.v;
        const int bc = op.bc_operator;
        const bool masked = bc == AMX_BCOP_IDENTITY;
        for (int i = 0; i < n_iter; i++) {
            AMX_HIP(hipEventRecord(ev[3 * i], s));
            if (op.scatter == AMX_SCATTER_ATOMIC) amx::launch_fill(v, op.n_total, true, 0.0, s);
            amx::ebe_launch(op, KType::F64, true, masked, 2, u, v, s);
            AMX_HIP(hipEventRecord(ev[3 * i + 1], s));
            if (op.scatter == AMX_SCATTER_ATOMIC)
                amx::launch_bc_fixup(true, bc, op.cdofs.as<uint32_t>(), op.n_cdofs, op.coef64.as<double>(), u, v, s);
            else
                amx::gather_launch(op, true, bc, u, v, s);
            AMX_HIP(hipEventRecord(ev[3 * i + 2], s));
        }
        AMX_HIP(hipEventSynchronize(ev.back()));
        double te = 0, tg = 0, tt = 0;
        for (int i = 0; i < n_iter; i++) {
            float a = 0, b = 0;
            AMX_HIP(hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 1]));
            AMX_HIP(hipEventElapsedTime(&b, ev[3 * i + 1], ev[3 * i + 2]));
            te += a;
            tg += b;
            tt += a + b;
        }
        if (ms_element) *ms_element = te / n_iter;
        if (ms_gather) *ms_gather = tg / n_iter;
        if (ms_total) *ms_total = tt / n_iter;
    });
}

void* amx_op_stream(amx_op* h) { return h ? (void*)h->op.stream : nullptr; }

int amx_op_synchronize(amx_op* h)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h, AMX_INVALID_ARGUMENT, "op is NULL");
        DeviceGuard dg(h->op.device);
        AMX_HIP(hipStreamSynchronize(h->op.stream));
    });
}

}  // extern "C"
