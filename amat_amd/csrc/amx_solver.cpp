// amx_solver.cpp -- device Jacobi-PCG and FP64 iterative refinement around
// the EbE operator (the KSPCG + PCJACOBI solve of examples/include/solve.hpp:45-65,
// run entirely on the GPU).
//
// The iteration runs without host round trips: scalars stay on the device and
// the host only polls the converged flag every few iterations (launches after
// convergence are no-ops, the matvec kernels included).
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

// Jacobi diagonal of the operator being solved -> 1/diag.  Constrained rows
// take the value that BC form puts there: 1 for identity rows, the penalty for
// penalty rows (apply_bc_diagonal, include/aMatFree.hpp:4309-4337, does the
// same for BC_IMATRIX / BC_PENALTY).
void build_dinv(Op& op, double* dinv, double* d, hipStream_t s)
{
    AMX_HIP(hipMemcpyAsync(d, op.diag.p, op.n_total * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (op.bc_operator == AMX_BCOP_IDENTITY)
        launch_set_entries(d, op.cdofs.as<uint32_t>(), op.n_cdofs, 1.0, s);
    else if (op.bc_operator == AMX_BCOP_PENALTY_DIAG)
        launch_set_entries(d, op.cdofs.as<uint32_t>(), op.n_cdofs, op.penalty(), s);
    invert(d, dinv, op.n_total, s);
}

namespace {

void ensure(DevBuf& b, size_t bytes)
{
    if (b.bytes < bytes) b.alloc(bytes);
}

void ensure_ws(Op& op)
{
    const size_t n = std::max<uint32_t>(op.n_total, 1);
    for (DevBuf* b : {&op.ws_r, &op.ws_z, &op.ws_p, &op.ws_q, &op.ws_dinv}) ensure(*b, n * sizeof(double));
    for (DevBuf* b : {&op.ws_x32, &op.ws_b32, &op.ws_r32, &op.ws_z32, &op.ws_p32, &op.ws_q32}) ensure(*b, n * sizeof(float));
    if (op.ws_partials.bytes < kPartialsBytes) {
        op.ws_partials.alloc(kPartialsBytes);
        AMX_HIP(hipMemset(op.ws_partials.p, 0, kPartialsBytes));  // reduction counters start at zero
    }
    ensure(op.ws_sc, 2 * sizeof(CgScalars));
    if (!op.host_sc) {
        void* p = nullptr;
        AMX_HIP(hipHostMalloc(&p, 2 * sizeof(CgScalars), hipHostMallocDefault));
        op.host_sc = static_cast<CgScalars*>(p);
    }
}

// host copies of the device scalars go through `host` (the handle's pinned
// copy), or through a local when it is NULL
void set_scalars_via(CgScalars* host, CgScalars* dsc, double rtol, int max_it, hipStream_t s, double delta = 0.0,
                     int max_upd = 0)
{
    CgScalars h{};
    h.tol2 = rtol * rtol;
    h.max_it = (uint32_t)max_it;
    h.d2 = delta * delta;
    h.max_upd = (uint32_t)max_upd;
    h.counter = 0;
    CgScalars* src = host ? host : &h;
    *src = h;
    AMX_HIP(hipMemcpyAsync(dsc, src, sizeof(CgScalars), hipMemcpyHostToDevice, s));
    AMX_HIP(hipStreamSynchronize(s));
}

CgScalars read_scalars_via(CgScalars* host, const CgScalars* dsc, hipStream_t s)
{
    CgScalars h{};
    CgScalars* dst = host ? host : &h;
    AMX_HIP(hipMemcpyAsync(dst, dsc, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    AMX_HIP(hipStreamSynchronize(s));
    return *dst;
}

void set_scalars(Op& op, CgScalars* dsc, double rtol, int max_it, hipStream_t s, double delta = 0.0,
                 int max_upd = 0)
{
    set_scalars_via(op.host_sc, dsc, rtol, max_it, s, delta, max_upd);
}

CgScalars read_scalars(Op& op, const CgScalars* dsc, hipStream_t s) { return read_scalars_via(op.host_sc, dsc, s); }

void build_dinv_ws(Op& op, hipStream_t s) { amx::build_dinv(op, op.ws_dinv.as<double>(), op.ws_q.as<double>(), s); }

// PCG on A x = b (x starts at 0), vectors of type f64 ? double : float, the
// operator with storage kt and the handle's BC form.  Returns iterations.
// PCG iterations from the current state until the device stops them.
CgScalars cg_loop_ws(Op& op, KType kt, bool f64, void* x, void* r, void* z, void* p, void* q, int max_it,
                     CgScalars* dsc, double* partials, const double* dinv, CgScalars* host, hipStream_t s)
{
    const uint32_t n = op.n_total;
    // poll every `chunk` iterations: about 1/8 of the iterations of a small
    // solve, every other one of a large solve
    const uint32_t chunk = n > (1u << 21) ? 2 : (n > (1u << 17) ? 4 : 8);
    const uint32_t* skip = &dsc->done;
    for (int launched = 0;;) {
        for (uint32_t k = 0; k < chunk; k++) {
            apply(op, kt, f64, op.bc_operator, p, q, s, skip);
            cg_pq(f64, p, q, n, partials, dsc, s);
            cg_update(f64, x, r, z, p, q, dinv, n, partials, dsc, s);
            cg_p(f64, p, z, n, dsc, s);
        }
        launched += (int)chunk;
        const CgScalars sc = read_scalars_via(host, dsc, s);
        if (sc.done || launched >= max_it) return sc;
    }
}

CgScalars cg_loop(Op& op, KType kt, bool f64, void* x, void* r, void* z, void* p, void* q, int max_it,
                  CgScalars* dsc, hipStream_t s)
{
    return cg_loop_ws(op, kt, f64, x, r, z, p, q, max_it, dsc, op.ws_partials.as<double>(),
                      op.ws_dinv.as<double>(), op.host_sc, s);
}

CgScalars pcg(Op& op, KType kt, bool f64, const void* b, void* x, void* r, void* z, void* p, void* q, double rtol,
              int max_it, CgScalars* dsc, hipStream_t s, int max_upd = 0)
{
    set_scalars(op, dsc, rtol, max_it, s, 0.0, max_upd);
    cg_init(f64, b, op.ws_dinv.as<double>(), x, r, z, p, op.n_total, op.ws_partials.as<double>(), dsc, s);
    return cg_loop(op, kt, f64, x, r, z, p, q, max_it, dsc, s);
}

// Mixed-precision PCG with reliable updates (residual replacement: Sleijpen
// and van der Vorst, "Reliable updated residuals in hybrid Bi-CG methods",
// Computing 56 (1996); the form GPU mixed-precision solvers use).  The CG
// recurrence runs with the operator of storage kt and vectors of type f64v
// from one start, never restarted.  The device stops it (scalars' state 1)
// when its ||r|| has fallen below delta times its largest value since the last
// replacement, or it reports convergence or the iteration limit; the host then
// flushes the recurrence's increment into the FP64 solution, replaces the
// recurrence residual by the true b - A x (FP64 operator), continues the
// direction update p = z + beta p and lets it run again unless the true
// residual has converged or a limit is reached (state 2).  A replacement that
// finds the true residual norm more than 4x the recurrence's restarts the
// direction (beta = 0): then the low-precision operator's error, not the
// recurrence, limits progress, and the step is a restarted refinement step.
// A phase also ends after twice the previous phase's length (64 iterations at
// least), so that a stalled recurrence is replaced too.  The stop is decided on the device, so the
// iterates do not depend on how often the host polls.
CgScalars ru_pcg(Op& op, KType kt, bool f64v, const double* b, double* x, void* xs, void* rs, void* zs, void* ps,
                 void* qs, double rtol, double delta, int max_it, int max_upd, CgScalars* dsc, hipStream_t s,
                 int& n_upd, bool& true_rr)
{
    const uint32_t n = op.n_total;
    double* partials = op.ws_partials.as<double>();
    const double* dinv = op.ws_dinv.as<double>();
    double* q64 = op.ws_q.as<double>();
    set_scalars(op, dsc, rtol, max_it, s, delta, max_upd);
    launch_fill(x, n, true, 0.0, s);
    launch_fill(xs, n, f64v, 0.0, s);
    launch_fill(q64, n, true, 0.0, s);
    ru_replace(f64v, true, b, q64, rs, zs, ps, dinv, n, partials, dsc, s);
    CgScalars sc = read_scalars(op, dsc, s);
    n_upd = 0;
    true_rr = true;
    if (sc.done) return sc;
    const uint32_t chunk = n > (1u << 21) ? 2 : (n > (1u << 17) ? 4 : 8);
    const uint32_t* skip = &dsc->done;
    for (;;) {
        for (uint32_t k = 0; k < chunk; k++) {
            apply(op, kt, f64v, op.bc_operator, ps, qs, s, skip);
            cg_pq(f64v, ps, qs, n, partials, dsc, s);
            cg_update(f64v, xs, rs, zs, ps, qs, dinv, n, partials, dsc, s);
            cg_p(f64v, ps, zs, n, dsc, s);
        }
        sc = read_scalars(op, dsc, s);
        if (!sc.done) continue;
        if (sc.breakdown) {
            ru_flush(f64v, x, xs, n, s);
            true_rr = false;
            return sc;
        }
        // state 1: replace the residual, continue the direction
        ru_flush(f64v, x, xs, n, s);
        apply(op, KType::F64, true, op.bc_operator, x, q64, s);
        ru_replace(f64v, false, b, q64, rs, zs, ps, dinv, n, partials, dsc, s);
        cg_p(f64v, ps, zs, n, dsc, s);  // a no-op once final
        sc = read_scalars(op, dsc, s);
        n_upd = (int)sc.n_upd;
        if (sc.done) return sc;
    }
}

}  // namespace

// FP64 Jacobi-PCG with true-residual replacements (amx_cg_solve_dev) on the
// workspace w; returns the final scalars
void fp64_cg(Op& op, const double* b, double* x, double rtol, int max_it, const CgWork& w, amx_solve_stats* stats,
             hipStream_t s)
{
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = op.n_total;
    build_dinv(op, w.dinv, w.q, s);
    CgScalars* dsc = w.sc;
    set_scalars_via(w.host, dsc, rtol, max_it, s, 0.0, kCgReplacements);
    cg_init(true, b, w.dinv, x, w.r, w.z, w.p, n, w.partials, dsc, s);
    CgScalars sc = cg_loop_ws(op, KType::F64, true, x, w.r, w.z, w.p, w.q, max_it, dsc, w.partials, w.dinv, w.host, s);
    // the recurrence stopped: its residual is replaced by the true b - A x
    // (FP64); the solve ends there when that meets rtol (or a limit is
    // reached), else CG continues the direction from the true residual
    CgScalars fin;
    for (;;) {
        apply(op, KType::F64, true, op.bc_operator, x, w.q, s);
        ru_replace(true, false, b, w.q, w.r, w.z, w.p, w.dinv, n, w.partials, dsc, s);
        cg_p(true, w.p, w.z, n, dsc, s);  // a no-op once final
        fin = read_scalars_via(w.host, dsc, s);
        if (fin.done || fin.breakdown) break;
        sc = cg_loop_ws(op, KType::F64, true, x, w.r, w.z, w.p, w.q, max_it, dsc, w.partials, w.dinv, w.host, s);
        if (sc.breakdown) break;
    }
    const double rel = fin.bb > 0 ? std::sqrt(fin.rr / fin.bb) : 0.0;
    if (stats) {
        stats->iterations = (int32_t)fin.iter;
        stats->outer_iterations = (int32_t)fin.n_upd - 1;
        stats->converged = rel <= rtol * 1.0000001 || fin.bb == 0.0;
        stats->precision = AMX_PREC_FP64;
        stats->rel_residual = rel;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    AMX_REQUIRE(!fin.breakdown, AMX_NOT_CONVERGED, "CG breakdown: p^T A p <= 0 (operator not SPD)");
}

}  // namespace amx

extern "C" {

int amx_cg_solve_dev(amx_op* h, const double* b, double* x, double rtol, int max_it, amx_solve_stats* stats,
                     void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && b && x, AMX_INVALID_ARGUMENT, "NULL argument");
        AMX_REQUIRE(rtol > 0.0 && max_it > 0, AMX_INVALID_ARGUMENT, "rtol and max_it must be positive");
        Op& op = h->op;
        AMX_REQUIRE(op.finalized, AMX_INVALID_STATE, "finalize() has not been called");
        AMX_REQUIRE(op.size == 1, AMX_NOT_IMPLEMENTED, "device solver is one-rank");
        amx::DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        amx::ensure_ws(op);
        amx::CgWork w;
        w.r = op.ws_r.as<double>();
        w.z = op.ws_z.as<double>();
        w.p = op.ws_p.as<double>();
        w.q = op.ws_q.as<double>();
        w.dinv = op.ws_dinv.as<double>();
        w.partials = op.ws_partials.as<double>();
        w.sc = op.ws_sc.as<amx::CgScalars>();
        w.host = op.host_sc;
        amx::fp64_cg(op, b, x, rtol, max_it, w, stats, s);
    });
}

int amx_ir_cg_solve_dev(amx_op* h, int precision, const double* b, double* x, double rtol, double inner_rtol,
                        int max_outer, int max_inner, amx_solve_stats* stats, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && b && x, AMX_INVALID_ARGUMENT, "NULL argument");
        const bool v64 = (precision & AMX_IR_F64_VECTORS) != 0;
        const bool reliable = (precision & AMX_IR_RELIABLE) != 0;
        const int flags = precision & (AMX_IR_F64_VECTORS | AMX_IR_RELIABLE);
        precision &= ~(AMX_IR_F64_VECTORS | AMX_IR_RELIABLE);
        AMX_REQUIRE(precision == AMX_PREC_FP32 || precision == AMX_PREC_BF16, AMX_INVALID_ARGUMENT,
                    "inner precision must be FP32 or BF16 (optionally | AMX_IR_F64_VECTORS | AMX_IR_RELIABLE)");
        AMX_REQUIRE(rtol > 0.0 && inner_rtol > 0.0 && inner_rtol < 1.0 && max_outer > 0 && max_inner > 0,
                    AMX_INVALID_ARGUMENT, "bad tolerances or iteration limits");
        Op& op = h->op;
        AMX_REQUIRE(op.finalized, AMX_INVALID_STATE, "finalize() has not been called");
        AMX_REQUIRE(op.has(precision), AMX_INVALID_STATE, "inner precision was not built at finalize");
        AMX_REQUIRE(op.size == 1, AMX_NOT_IMPLEMENTED, "device solver is one-rank");
        amx::DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        const auto t0 = std::chrono::steady_clock::now();
        amx::ensure_ws(op);
        if (v64)
            for (DevBuf* v : {&op.ws_ib, &op.ws_ix, &op.ws_ir, &op.ws_iz, &op.ws_ip, &op.ws_iq})
                amx::ensure(*v, (size_t)std::max<uint32_t>(op.n_total, 1) * sizeof(double));
        amx::build_dinv_ws(op, s);
        const KType kt = precision == AMX_PREC_FP32 ? KType::F32 : KType::BF16;
        amx::CgScalars* outer = op.ws_sc.as<amx::CgScalars>();
        amx::CgScalars* inner = outer + 1;
        const uint32_t n = op.n_total;
        double* partials = op.ws_partials.as<double>();
        if (reliable) {
            int n_upd = 0;
            bool true_rr = true;
            amx::CgScalars sc =
                v64 ? amx::ru_pcg(op, kt, true, b, x, op.ws_ix.p, op.ws_ir.p, op.ws_iz.p, op.ws_ip.p, op.ws_iq.p, rtol,
                                  inner_rtol, max_inner, max_outer, outer, s, n_upd, true_rr)
                    : amx::ru_pcg(op, kt, false, b, x, op.ws_x32.p, op.ws_r32.p, op.ws_z32.p, op.ws_p32.p,
                                  op.ws_q32.p, rtol, inner_rtol, max_inner, max_outer, outer, s, n_upd, true_rr);
            const bool breakdown = sc.breakdown != 0;
            if (!true_rr) {  // stopped between replacements: measure the true residual
                amx::apply(op, KType::F64, true, op.bc_operator, x, op.ws_q.p, s);
                amx::residual(b, op.ws_q.as<double>(), op.ws_r.as<double>(), n, partials, outer, s);
                sc = amx::read_scalars(op, outer, s);
            }
            const double rel = sc.bb > 0 ? std::sqrt(sc.rr / sc.bb) : 0.0;
            if (stats) {
                stats->iterations = (int32_t)sc.iter;
                stats->outer_iterations = n_upd;
                stats->converged = rel <= rtol * 1.0000001 || sc.bb == 0.0;
                stats->precision = precision | flags;
                stats->rel_residual = rel;
                stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            }
            AMX_REQUIRE(!breakdown, AMX_NOT_CONVERGED, "CG breakdown (low-precision operator not SPD)");
            return;
        }
        // x = 0, r = b (the outer scalars start zeroed: their reduction counter included)
        amx::cg_reset(outer, rtol, (uint32_t)max_outer, s);
        amx::launch_fill(x, n, true, 0.0, s);
        amx::launch_fill(op.ws_q.p, n, true, 0.0, s);
        amx::residual(b, op.ws_q.as<double>(), op.ws_r.as<double>(), n, partials, outer, s);
        int total_inner = 0, outer_it = 0;
        double rel = 1.0;
        bool breakdown = false;
        for (;;) {
            const amx::CgScalars o = amx::read_scalars(op, outer, s);
            rel = o.bb > 0 ? std::sqrt(o.rr / o.bb) : 0.0;
            if (rel <= rtol || outer_it >= max_outer) break;
            // inner solve A_low d = r / ||r|| (FP32 or FP64 vectors, FP64 scalars)
            amx::CgScalars in;
            if (v64) {
                amx::scale_to_f64(op.ws_r.as<double>(), op.ws_ib.as<double>(), n, outer, s);
                in = amx::pcg(op, kt, true, op.ws_ib.p, op.ws_ix.p, op.ws_ir.p, op.ws_iz.p, op.ws_ip.p, op.ws_iq.p,
                              inner_rtol, max_inner, inner, s);
            } else {
                amx::scale_to_f32(op.ws_r.as<double>(), op.ws_b32.as<float>(), n, outer, s);
                in = amx::pcg(op, kt, false, op.ws_b32.p, op.ws_x32.p, op.ws_r32.p, op.ws_z32.p, op.ws_p32.p,
                              op.ws_q32.p, inner_rtol, max_inner, inner, s);
            }
            total_inner += (int)in.iter;
            breakdown = breakdown || in.breakdown;
            // x += ||r|| d ; r = b - A x with the FP64 operator
            if (v64) amx::correct64(x, op.ws_ix.as<double>(), n, outer, s);
            else amx::correct(x, op.ws_x32.as<float>(), n, outer, s);
            amx::apply(op, KType::F64, true, op.bc_operator, x, op.ws_q.p, s);
            amx::residual(b, op.ws_q.as<double>(), op.ws_r.as<double>(), n, partials, outer, s);
            outer_it++;
            if (in.iter == 0) break;  // no progress possible
        }
        if (stats) {
            stats->iterations = total_inner;
            stats->outer_iterations = outer_it;
            stats->converged = rel <= rtol;
            stats->precision = precision | flags;
            stats->rel_residual = rel;
            stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        AMX_REQUIRE(!breakdown, AMX_NOT_CONVERGED, "inner CG breakdown (low-precision operator not SPD)");
    });
}

int amx_dev_dot2(const double* x, const double* y, uint64_t n, uint32_t n_blocks, uint32_t repeat, double* out,
                 int device, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE((x && y) || n == 0, AMX_INVALID_ARGUMENT, "NULL argument");
        AMX_REQUIRE(out, AMX_INVALID_ARGUMENT, "NULL argument");
        AMX_REQUIRE(n <= 0xffffffffull, AMX_INVALID_ARGUMENT, "n exceeds 2^32 - 1");
        AMX_REQUIRE(n_blocks >= 1 && n_blocks <= amx::kMaxRedBlocks, AMX_INVALID_ARGUMENT,
                    "n_blocks must be in [1, 1024]");
        AMX_REQUIRE(repeat >= 1, AMX_INVALID_ARGUMENT, "repeat must be positive");
        amx::DeviceGuard dg(device);
        hipStream_t s = static_cast<hipStream_t>(stream);
        // partials and group counters, then the top counter, all zero
        DevBuf ws;
        ws.alloc(amx::kPartialsBytes + 256);
        AMX_HIP(hipMemsetAsync(ws.p, 0, ws.bytes, s));
        double* partials = ws.as<double>();
        uint32_t* counter = reinterpret_cast<uint32_t*>(ws.as<char>() + amx::kPartialsBytes);
        for (uint32_t k = 0; k < repeat; k++)
            amx::dot2(x, y, (uint32_t)n, n_blocks, partials, counter, out + 2 * (size_t)k, s);
        AMX_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
