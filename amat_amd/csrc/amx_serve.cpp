// amx_serve.cpp -- continuous-batching CG service on one GPU.
//
// A server runs many CG solves (jobs) at once.  Each job keeps its Krylov
// vectors, its element-row buffer and its scalars resident on the device; the
// element matrices stay in the job's (shared, read-only) operator.  One server
// step launches `n_iter` CG iterations for ALL active jobs, each phase of the
// iteration as one batched launch over the jobs (element kernels per element
// type and storage, then the transposed gather, p.q, the x/r/z update and the
// p update), so small solves fill the GPU together instead of one launch each.
// The host looks at the jobs' scalars once per step (one device-to-host copy);
// jobs leave when done and new ones are admitted between steps.
//
// Jobs on method-2 (FREE) operators (FP64 CG, or FP32 inner solves) launch their in-kernel element
// kernels one job at a time; the gather and the CG phases stay batched.
//
// Every job's arithmetic depends on its own data only: the scatter is the
// deterministic transposed gather and every reduction uses the job's fixed
// single-solve grid, so a job's result is bit-identical alone, in any mix,
// and to amx_cg_solve_dev / amx_ir_cg_solve_dev on the same operator.
//
// Methods: FP64 Jacobi-PCG (the KSPCG + PCJACOBI solve of
// examples/include/solve.hpp:45-65), or FP64 iterative refinement whose inner
// Jacobi-PCG iterates on the FP32 or BF16 element matrices with FP32 vectors
// (outer residual, scalars and penalty rows in FP64), restarted per outer
// step or (AMX_IR_RELIABLE) one recurrence with reliable updates.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "amx_op.hpp"

using amx::DevBuf;
using amx::fail;
using amx::guarded;
using amx::KType;
using amx::Op;

namespace {

using Clock = std::chrono::steady_clock;

enum Phase { kIter = 0, kOuterPending = 1, kFinalPending = 2, kDone = 3 };

struct Job {
    int64_t id = 0;
    amx_op* h = nullptr;
    int method = AMX_PREC_FP64;  // FP64 CG, or IR with FP32 / BF16 inner storage
    bool reliable = false;       // IR form: reliable updates instead of restarts
    double rtol = 1e-10, inner_rtol = 1e-2;
    int max_it = 10000, max_outer = 50;
    int slot = -1;
    Phase phase = kIter;
    bool f64iter = true;   // iteration vectors FP64 (CG) or FP32 (IR inner)
    KType kt = KType::F64; // element storage of the iteration
    uint32_t n = 0;
    // device state
    DevBuf b, x, dinv, ve, partials;  // FP64 rhs, solution, 1/diag; element rows; reduction partials
    DevBuf r, z, p, q;                // iteration vectors
    DevBuf xi_buf;                    // IR: inner solution (FP32); CG: unused (x iterates)
    DevBuf r64, q64, b32;             // IR: FP64 residual, A x, scaled residual (FP32; restarted form)
    void* xi = nullptr;               // iteration solution vector
    uint64_t bytes = 0;
    // progress
    int outer = 0, inner_total = 0, cg_iters = 0;
    bool breakdown = false, converged = false, stalled = false;
    double rel = 1.0;
    Clock::time_point t_submit, t_done;
};

struct Table {
    DevBuf dev_jobs, dev_prefix;
    uint32_t n = 0, total = 0;
    uint32_t max_nx = 0;    // element tables of records: the largest exception count
};

}  // namespace

struct amx_server {
    int device = 0;
    int max_jobs = 64;
    hipStream_t stream = nullptr;
    int64_t next_id = 1;
    std::map<int64_t, std::unique_ptr<Job>> jobs;  // active and finished, until fetched
    std::vector<int> free_slots;
    DevBuf sc_dev;                       // [max_jobs][2] CgScalars: iteration, outer
    amx::CgScalars* sc_host = nullptr;   // pinned mirror
    std::deque<int64_t> finished;        // not yet reported by amx_server_poll
    uint64_t workspace_bytes = 0;
    uint64_t steps = 0, iterations = 0, completed = 0;
    bool tables_dirty = true;
    // batched element launches, key: n | kt << 16 | f64 << 24 | records << 25 |
    // (records) identity BC form << 26
    std::map<uint64_t, Table> ebe;
    Table gather[2], cg[2];         // [f64]
    std::vector<Job*> free_jobs;    // jobs on method-2 (FREE) operators: element kernels launched per job
    // the per-job FREE element kernels run on kAux side streams (default
    // priority) forked from and joined back into `stream` every iteration, so
    // that the kernels of small jobs share the GPU
    static constexpr int kAux = 4;
    hipStream_t aux[kAux] = {};
    hipEvent_t fork = nullptr, join[kAux] = {};
    ~amx_server()
    {
        jobs.clear();
        if (sc_host) (void)hipHostFree(sc_host);
        for (int k = 0; k < kAux; k++) {
            if (join[k]) (void)hipEventDestroy(join[k]);
            if (aux[k]) (void)hipStreamDestroy(aux[k]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (stream) (void)hipStreamDestroy(stream);
    }
    amx::CgScalars* it_sc(const Job& j) { return sc_dev.as<amx::CgScalars>() + 2 * j.slot; }
    amx::CgScalars* out_sc(const Job& j) { return sc_dev.as<amx::CgScalars>() + 2 * j.slot + 1; }
    const amx::CgScalars& it_host(const Job& j) const { return sc_host[2 * j.slot]; }
    const amx::CgScalars& out_host(const Job& j) const { return sc_host[2 * j.slot + 1]; }
};

namespace {

uint64_t job_bytes(const Op& op, int method)
{
    const uint64_t n = std::max<uint32_t>(op.n_total, 1);
    const uint64_t ve = std::max<uint64_t>(op.ve_rows, 1) * 8;
    const uint64_t partials = amx::kPartialsBytes;
    if (method == AMX_PREC_FP64) return 7 * n * 8 + ve + partials;           // b x dinv r z p q
    if (method & AMX_IR_RELIABLE) return 5 * n * 8 + 5 * n * 4 + ve + partials;  // b x dinv r64 q64 + r z p q xs
    return 5 * n * 8 + 6 * n * 4 + ve + partials;                           // b x dinv r64 q64 + r z p q xi b32
}

bool valid_method(int method)
{
    const int base = method & ~AMX_IR_RELIABLE;
    if (method & AMX_IR_RELIABLE) return base == AMX_PREC_FP32 || base == AMX_PREC_BF16;
    return base == AMX_PREC_FP64 || base == AMX_PREC_FP32 || base == AMX_PREC_BF16;
}

template <typename T>
void upload_table(Table& t, const std::vector<T>& rows, const std::vector<uint32_t>& prefix)
{
    t.n = (uint32_t)rows.size();
    t.total = prefix.empty() ? 0 : prefix.back();
    if (rows.empty()) return;
    if (t.dev_jobs.bytes < rows.size() * sizeof(T)) t.dev_jobs.alloc(rows.size() * sizeof(T));
    if (t.dev_prefix.bytes < prefix.size() * 4) t.dev_prefix.alloc(prefix.size() * 4);
    AMX_HIP(hipMemcpy(t.dev_jobs.p, rows.data(), rows.size() * sizeof(T), hipMemcpyHostToDevice));
    AMX_HIP(hipMemcpy(t.dev_prefix.p, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice));
}

// launch tables over the unfinished jobs (a job whose iteration scalars say
// done turns its parts of the launches into no-ops)
void rebuild_tables(amx_server& S)
{
    std::map<uint64_t, std::pair<std::vector<amx::EbeJob>, std::vector<uint32_t>>> ebe;
    std::vector<amx::GatherJob> g[2];
    std::vector<uint32_t> gp[2] = {{0}, {0}};
    std::vector<amx::CgJob> c[2];
    std::vector<uint32_t> cp[2] = {{0}, {0}};
    S.free_jobs.clear();
    for (auto& kv : S.jobs) {
        Job& j = *kv.second;
        if (j.phase == kDone) continue;
        Op& op = j.h->op;
        const int f = j.f64iter ? 1 : 0;
        const uint32_t* skip = &S.it_sc(j)->done;
        if (op.is_free()) S.free_jobs.push_back(&j);
        for (auto& grp : op.groups) {
            if (op.is_free()) break;
            // 60-DoF records: one-wave-block launches, one per BC form
            const bool rec = grp.L[(int)j.kt].nsym, masked = op.bc_operator == AMX_BCOP_IDENTITY;
            const uint64_t key = (uint64_t)grp.n | ((uint64_t)j.kt << 16) | ((uint64_t)f << 24) |
                                 ((uint64_t)(rec ? 1 : 0) << 25) | ((uint64_t)(rec && masked ? 1 : 0) << 26);
            auto& e = ebe[key];
            if (e.second.empty()) e.second.push_back(0);
            amx::EbeJob a{};
            a.K = grp.K[(int)j.kt].p;
            a.map = grp.map.as<uint32_t>();
            a.u = j.p.p;
            a.out = j.ve.p;
            a.ve_base = grp.ve_base;
            a.ve_rs = grp.ve_rs;
            a.ve_ss = grp.ve_ss;
            a.tile_begin = 0;
            a.n_tiles = grp.n_tiles;
            a.skip = skip;
            a.kdiag = j.kt == KType::BF16 ? grp.kdiag16.as<float>() : nullptr;
            a.masked = op.bc_operator == AMX_BCOP_IDENTITY ? 1u : 0u;
            a.sym = grp.L[(int)j.kt].sym ? 1u : 0u;
            a.nsym_nx = grp.L[(int)j.kt].nx;
            a.nsym_xs = j.kt == KType::F64 ? grp.nsym_xs.as<uint32_t>() : grp.psx_tab[(int)j.kt].as<uint32_t>();
            a.nsym_colany = grp.nsym_colany;
            e.first.push_back(a);
            e.second.push_back(e.second.back() + grp.n_tiles);
        }
        if (!op.is_free()) {  // method-2 jobs gather on their side stream (launch_iteration)
            amx::GatherJob ga{};
            ga.off = op.off.as<uint32_t>();
            ga.pos = op.pos.as<uint32_t>();
            ga.ve = j.ve.p;
            ga.u = j.p.p;
            ga.flag = op.flag.as<uint8_t>();
            ga.coef = op.coef64.as<double>();
            ga.v = j.q.p;
            ga.n = j.n;
            ga.bc = (uint32_t)op.bc_operator;
            ga.skip = skip;
            g[f].push_back(ga);
            gp[f].push_back(gp[f].back() + (j.n + 255) / 256);
        }
        amx::CgJob cj{};
        cj.x = j.xi;
        cj.r = j.r.p;
        cj.z = j.z.p;
        cj.p = j.p.p;
        cj.q = j.q.p;
        cj.dinv = j.dinv.as<double>();
        cj.partials = j.partials.as<double>();
        cj.sc = S.it_sc(j);
        cj.n = j.n;
        cj.nblk = amx::reduction_grid(j.n);
        c[f].push_back(cj);
        cp[f].push_back(cp[f].back() + cj.nblk);
    }
    S.ebe.clear();
    for (auto& kv : ebe) {
        Table& t = S.ebe[kv.first];
        upload_table(t, kv.second.first, kv.second.second);
        t.max_nx = 0;
        for (const auto& a : kv.second.first) t.max_nx = std::max(t.max_nx, a.nsym_nx);
    }
    for (int f = 0; f < 2; f++) {
        upload_table(S.gather[f], g[f], gp[f]);
        upload_table(S.cg[f], c[f], cp[f]);
    }
    S.tables_dirty = false;
}

// one CG iteration of every unfinished job: the order of the single-solve
// loop (apply = element kernels + gather, then p.q, update, p)
void launch_iteration(amx_server& S)
{
    hipStream_t s = S.stream;
    for (auto& kv : S.ebe) {
        const uint32_t n = (uint32_t)(kv.first & 0xffff);
        const KType kt = (KType)((kv.first >> 16) & 0xff);
        const bool f64 = ((kv.first >> 24) & 1) != 0, rec = ((kv.first >> 25) & 1) != 0;
        const bool masked = ((kv.first >> 26) & 1) != 0;
        Table& t = kv.second;
        amx::launch_ebe_batched(n, kt, f64, t.dev_jobs.as<amx::EbeJob>(), t.dev_prefix.as<uint32_t>(), t.n, t.total,
                                s, rec, t.max_nx, masked);
    }
    if (!S.free_jobs.empty()) {  // element kernels from coordinates (FP64 or FP32) and the job's gather
        const int na = std::min<int>(amx_server::kAux, (int)S.free_jobs.size());
        AMX_HIP(hipEventRecord(S.fork, s));
        for (int k = 0; k < na; k++) AMX_HIP(hipStreamWaitEvent(S.aux[k], S.fork, 0));
        for (size_t i = 0; i < S.free_jobs.size(); i++) {
            Job* jp = S.free_jobs[i];
            Op& op = jp->h->op;
            const uint32_t* skip = &S.it_sc(*jp)->done;
            amx::ebe_launch(op, jp->kt, jp->f64iter, op.bc_operator == AMX_BCOP_IDENTITY, 2, jp->p.p, nullptr,
                            S.aux[i % na], skip, jp->ve.p);
            // the job's gather follows on the same stream (the arithmetic of the
            // batched gather), so it overlaps the other jobs' element kernels
            amx::gather_launch(op, jp->f64iter, op.bc_operator, jp->p.p, jp->q.p, S.aux[i % na], skip, jp->ve.p);
        }
        for (int k = 0; k < na; k++) {
            AMX_HIP(hipEventRecord(S.join[k], S.aux[k]));
            AMX_HIP(hipStreamWaitEvent(s, S.join[k], 0));
        }
    }
    for (int f = 0; f < 2; f++) {
        Table& t = S.gather[f];
        amx::launch_gather_batched(f == 1, t.dev_jobs.as<amx::GatherJob>(), t.dev_prefix.as<uint32_t>(), t.n,
                                   t.total, s);
    }
    for (int f = 0; f < 2; f++) {
        Table& t = S.cg[f];
        amx::cg_pq_batched(f == 1, t.dev_jobs.as<amx::CgJob>(), t.dev_prefix.as<uint32_t>(), t.n, t.total, s);
    }
    for (int f = 0; f < 2; f++) {
        Table& t = S.cg[f];
        amx::cg_update_batched(f == 1, t.dev_jobs.as<amx::CgJob>(), t.dev_prefix.as<uint32_t>(), t.n, t.total, s);
    }
    for (int f = 0; f < 2; f++) {
        Table& t = S.cg[f];
        amx::cg_p_batched(f == 1, t.dev_jobs.as<amx::CgJob>(), t.dev_prefix.as<uint32_t>(), t.n, t.total, s);
    }
}

// IR: inner solve on r64 / ||r64|| (the outer scalars hold ||r64||^2)
void start_inner(amx_server& S, Job& j)
{
    hipStream_t s = S.stream;
    amx::scale_to_f32(j.r64.as<double>(), j.b32.as<float>(), j.n, S.out_sc(j), s);
    amx::cg_reset(S.it_sc(j), j.inner_rtol, (uint32_t)j.max_it, s);
    amx::cg_init(false, j.b32.p, j.dinv.as<double>(), j.xi, j.r.p, j.z.p, j.p.p, j.n, j.partials.as<double>(),
                 S.it_sc(j), s);
}

// IR after an inner solve: x += ||r|| d, r64 = b - A x (FP64 operator)
void outer_update(amx_server& S, Job& j)
{
    hipStream_t s = S.stream;
    Op& op = j.h->op;
    amx::correct(j.x.as<double>(), static_cast<const float*>(j.xi), j.n, S.out_sc(j), s);
    amx::apply(op, KType::F64, true, op.bc_operator, j.x.p, j.q64.p, s, nullptr, j.ve.p);
    amx::residual(j.b.as<double>(), j.q64.as<double>(), j.r64.as<double>(), j.n, j.partials.as<double>(), S.out_sc(j),
                  s);
}

// reliable updates: x += xs, r = b - A x (FP64 operator) replaces the
// recurrence residual, p = z + beta p continues the direction (the sequence of
// the single-solve form)
void replace_residual(amx_server& S, Job& j)
{
    hipStream_t s = S.stream;
    Op& op = j.h->op;
    amx::ru_flush(false, j.x.as<double>(), j.xi, j.n, s);
    amx::apply(op, KType::F64, true, op.bc_operator, j.x.p, j.q64.p, s, nullptr, j.ve.p);
    amx::ru_replace(false, false, j.b.as<double>(), j.q64.as<double>(), j.r.p, j.z.p, j.p.p, j.dinv.as<double>(), j.n,
                    j.partials.as<double>(), S.it_sc(j), s);
    amx::cg_p(false, j.p.p, j.z.p, j.n, S.it_sc(j), s);
}

// FP64 CG: the stopped recurrence's residual replaced by the true b - A x
// (the job ends when that meets rtol or a limit, else it iterates on), as
// amx_cg_solve_dev does
void replace_residual64(amx_server& S, Job& j)
{
    hipStream_t s = S.stream;
    Op& op = j.h->op;
    amx::apply(op, KType::F64, true, op.bc_operator, j.x.p, j.q.p, s, nullptr, j.ve.p);
    amx::ru_replace(true, false, j.b.as<double>(), j.q.as<double>(), j.r.p, j.z.p, j.p.p, j.dinv.as<double>(), j.n,
                    j.partials.as<double>(), S.it_sc(j), s);
    amx::cg_p(true, j.p.p, j.z.p, j.n, S.it_sc(j), s);
}

// true residual ||b - A x|| of a reliable-update job stopped by a breakdown
void final_residual64(amx_server& S, Job& j)
{
    hipStream_t s = S.stream;
    Op& op = j.h->op;
    amx::ru_flush(false, j.x.as<double>(), j.xi, j.n, s);
    amx::apply(op, KType::F64, true, op.bc_operator, j.x.p, j.q64.p, s, nullptr, j.ve.p);
    amx::residual(j.b.as<double>(), j.q64.as<double>(), j.r64.as<double>(), j.n, j.partials.as<double>(),
                  S.out_sc(j), s);
}

// FP64 CG after convergence: true residual ||b - A x|| into the outer scalars
void final_residual(amx_server& S, Job& j)
{
    hipStream_t s = S.stream;
    Op& op = j.h->op;
    amx::apply(op, KType::F64, true, op.bc_operator, j.x.p, j.q.p, s, nullptr, j.ve.p);
    amx::residual(j.b.as<double>(), j.q.as<double>(), j.r.as<double>(), j.n, j.partials.as<double>(), S.out_sc(j), s);
}

void finish(amx_server& S, Job& j)
{
    j.phase = kDone;
    j.t_done = Clock::now();
    S.finished.push_back(j.id);
    S.completed++;
    S.tables_dirty = true;
}

// read the scalars of every job (one copy), advance each job's state machine
void advance(amx_server& S)
{
    AMX_HIP(hipMemcpyAsync(S.sc_host, S.sc_dev.p, (size_t)S.max_jobs * 2 * sizeof(amx::CgScalars),
                           hipMemcpyDeviceToHost, S.stream));
    AMX_HIP(hipStreamSynchronize(S.stream));
    for (auto& kv : S.jobs) {
        Job& j = *kv.second;
        const amx::CgScalars& it = S.it_host(j);
        const amx::CgScalars& ot = S.out_host(j);
        switch (j.phase) {
        case kIter:
            if (!it.done) break;
            j.breakdown = j.breakdown || it.breakdown;
            if (j.reliable) {
                j.inner_total = (int)it.iter;
                j.outer = (int)it.n_upd;
                if (it.breakdown) {
                    final_residual64(S, j);
                    j.phase = kFinalPending;
                } else if (it.state == 2) {  // decided on the true residual
                    j.rel = it.bb > 0 ? std::sqrt(it.rr / it.bb) : 0.0;
                    j.converged = j.rel <= j.rtol * 1.0000001 || it.bb == 0.0;
                    finish(S, j);
                } else {
                    replace_residual(S, j);  // the job runs on in the next step unless final
                }
            } else if (j.method == AMX_PREC_FP64) {
                j.cg_iters = (int)it.iter;
                if (it.breakdown) {
                    final_residual(S, j);
                    j.phase = kFinalPending;
                } else if (it.state == 2) {  // decided on the true residual
                    j.outer = (int)it.n_upd - 1;
                    j.rel = it.bb > 0 ? std::sqrt(it.rr / it.bb) : 0.0;
                    j.converged = j.rel <= j.rtol * 1.0000001 || it.bb == 0.0;
                    finish(S, j);
                } else {
                    replace_residual64(S, j);  // the job runs on in the next step unless final
                }
            } else {
                j.inner_total += (int)it.iter;
                j.stalled = it.iter == 0;
                outer_update(S, j);
                j.phase = kOuterPending;
            }
            break;
        case kOuterPending: {
            // outer scalars of the residual launched after the previous step
            j.outer++;
            j.rel = ot.bb > 0 ? std::sqrt(ot.rr / ot.bb) : 0.0;
            if (j.rel <= j.rtol || j.outer >= j.max_outer || j.stalled || j.breakdown) {
                j.converged = j.rel <= j.rtol;
                finish(S, j);
            } else {
                start_inner(S, j);
                j.phase = kIter;
            }
            break;
        }
        case kFinalPending:
            j.rel = ot.bb > 0 ? std::sqrt(ot.rr / ot.bb) : 0.0;
            j.converged = (j.rel <= j.rtol * 1.0000001 || ot.bb == 0.0) && !j.breakdown;
            finish(S, j);
            break;
        case kDone:
            break;
        }
    }
}

}  // namespace

extern "C" {

int amx_server_create(int device, int max_jobs, amx_server** out)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(out && max_jobs > 0 && max_jobs <= 4096, AMX_INVALID_ARGUMENT, "max_jobs must be 1..4096");
        int ndev = 0;
        AMX_HIP(hipGetDeviceCount(&ndev));
        AMX_REQUIRE(device >= 0 && device < ndev, AMX_INVALID_ARGUMENT, "device not present");
        amx::DeviceGuard dg(device);
        std::unique_ptr<amx_server> S(new amx_server());
        S->device = device;
        S->max_jobs = max_jobs;
        AMX_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        AMX_HIP(hipEventCreateWithFlags(&S->fork, hipEventDisableTiming));
        for (int k = 0; k < amx_server::kAux; k++) {
            AMX_HIP(hipStreamCreateWithFlags(&S->aux[k], hipStreamNonBlocking));
            AMX_HIP(hipEventCreateWithFlags(&S->join[k], hipEventDisableTiming));
        }
        S->sc_dev.alloc((size_t)max_jobs * 2 * sizeof(amx::CgScalars));
        AMX_HIP(hipMemset(S->sc_dev.p, 0, S->sc_dev.bytes));
        void* p = nullptr;
        AMX_HIP(hipHostMalloc(&p, (size_t)max_jobs * 2 * sizeof(amx::CgScalars), hipHostMallocDefault));
        S->sc_host = static_cast<amx::CgScalars*>(p);
        std::fill(S->sc_host, S->sc_host + 2 * max_jobs, amx::CgScalars{});
        for (int i = max_jobs - 1; i >= 0; i--) S->free_slots.push_back(i);
        *out = S.release();
    });
}

int amx_server_destroy(amx_server* S)
{
    if (S) {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(S->device);
        if (S->stream) (void)hipStreamSynchronize(S->stream);
        delete S;
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    return AMX_SUCCESS;
}

int amx_server_job_bytes(const amx_op* h, int method, uint64_t* bytes)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(h && bytes, AMX_INVALID_ARGUMENT, "NULL argument");
        AMX_REQUIRE(valid_method(method), AMX_INVALID_ARGUMENT,
                    "method: AMX_PREC_FP64 (CG) or AMX_PREC_FP32 / AMX_PREC_BF16 (IR), optionally | AMX_IR_RELIABLE");
        *bytes = job_bytes(h->op, method);
    });
}

int amx_server_submit(amx_server* S, amx_op* h, int method, const double* b_host, double rtol, double inner_rtol,
                      int max_it, int max_outer, int64_t* job_id)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(S && h && b_host && job_id, AMX_INVALID_ARGUMENT, "NULL argument");
        AMX_REQUIRE(valid_method(method), AMX_INVALID_ARGUMENT,
                    "method: AMX_PREC_FP64 (CG) or AMX_PREC_FP32 / AMX_PREC_BF16 (IR), optionally | AMX_IR_RELIABLE");
        const bool reliable = (method & AMX_IR_RELIABLE) != 0;
        const int base = method & ~AMX_IR_RELIABLE;
        AMX_REQUIRE(rtol > 0.0 && max_it > 0 && max_outer > 0, AMX_INVALID_ARGUMENT, "bad tolerance or limits");
        AMX_REQUIRE(method == AMX_PREC_FP64 || (inner_rtol > 0.0 && inner_rtol < 1.0), AMX_INVALID_ARGUMENT,
                    "inner_rtol must be in (0, 1)");
        Op& op = h->op;
        AMX_REQUIRE(op.finalized, AMX_INVALID_STATE, "operator not finalized");
        AMX_REQUIRE(!op.is_free() || base != AMX_PREC_BF16, AMX_NOT_IMPLEMENTED,
                    "method-2 (FREE) operators are served in FP64 or FP32");
        AMX_REQUIRE(op.size == 1, AMX_NOT_IMPLEMENTED, "served solves are one-rank");
        AMX_REQUIRE(op.device == S->device, AMX_INVALID_ARGUMENT, "operator on another device");
        AMX_REQUIRE(op.scatter == AMX_SCATTER_DETERMINISTIC, AMX_NOT_IMPLEMENTED,
                    "served solves use the deterministic scatter");
        AMX_REQUIRE(op.has(base), AMX_INVALID_STATE, "the method's storage was not built at finalize");
        for (auto& g : op.groups)
            AMX_REQUIRE(op.is_free() || amx::batched_n(g.n), AMX_NOT_IMPLEMENTED,
                        "batched launches need a tile kernel for n = " + std::to_string(g.n));
        AMX_REQUIRE(!S->free_slots.empty(), AMX_REJECTED, "all job slots are in use");
        amx::DeviceGuard dg(S->device);
        hipStream_t s = S->stream;
        std::unique_ptr<Job> J(new Job());
        Job& j = *J;
        j.id = S->next_id++;
        j.h = h;
        j.method = base;
        j.reliable = reliable;
        j.rtol = rtol;
        j.inner_rtol = inner_rtol;
        j.max_it = max_it;
        j.max_outer = max_outer;
        j.n = op.n_total;
        j.f64iter = base == AMX_PREC_FP64;
        j.kt = base == AMX_PREC_FP64 ? KType::F64 : (base == AMX_PREC_FP32 ? KType::F32 : KType::BF16);
        const size_t n = std::max<uint32_t>(j.n, 1);
        const size_t vs = j.f64iter ? 8 : 4;
        j.b.alloc(n * 8);
        j.x.alloc(n * 8);
        j.dinv.alloc(n * 8);
        j.ve.alloc(std::max<uint64_t>(op.ve_rows, 1) * 8);
        j.partials.alloc(amx::kPartialsBytes);
        AMX_HIP(hipMemsetAsync(j.partials.p, 0, amx::kPartialsBytes, s));  // reduction counters start at zero
        for (DevBuf* v : {&j.r, &j.z, &j.p, &j.q}) v->alloc(n * vs);
        if (!j.f64iter) {
            j.xi_buf.alloc(n * 4);
            if (!reliable) j.b32.alloc(n * 4);
            j.r64.alloc(n * 8);
            j.q64.alloc(n * 8);
            j.xi = j.xi_buf.p;
        } else {
            j.xi = j.x.p;
        }
        j.bytes = job_bytes(op, method);
        j.slot = S->free_slots.back();
        S->free_slots.pop_back();
        j.t_submit = Clock::now();
        AMX_HIP(hipMemcpyAsync(j.b.p, b_host, (size_t)j.n * 8, hipMemcpyHostToDevice, s));
        amx::cg_reset(S->out_sc(j), rtol, (uint32_t)max_outer, s);  // zeroed outer scalars (reduction counter)
        amx::build_dinv(op, j.dinv.as<double>(), j.x.as<double>(), s);  // x as scratch (zeroed below)
        if (reliable) {
            // the single-solve start: x = xs = 0, the recurrence residual from b
            amx::cg_reset(S->it_sc(j), rtol, (uint32_t)max_it, s, inner_rtol, (uint32_t)max_outer);
            amx::launch_fill(j.x.p, j.n, true, 0.0, s);
            amx::launch_fill(j.xi, j.n, false, 0.0, s);
            amx::launch_fill(j.q64.p, j.n, true, 0.0, s);
            amx::ru_replace(false, true, j.b.as<double>(), j.q64.as<double>(), j.r.p, j.z.p, j.p.p,
                            j.dinv.as<double>(), j.n, j.partials.as<double>(), S->it_sc(j), s);
            j.phase = kIter;
        } else if (j.f64iter) {
            amx::cg_reset(S->it_sc(j), rtol, (uint32_t)max_it, s, 0.0, amx::kCgReplacements);
            amx::cg_init(true, j.b.p, j.dinv.as<double>(), j.x.p, j.r.p, j.z.p, j.p.p, j.n, j.partials.as<double>(),
                         S->it_sc(j), s);
            j.phase = kIter;
        } else {
            amx::launch_fill(j.x.p, j.n, true, 0.0, s);
            amx::launch_fill(j.q64.p, j.n, true, 0.0, s);
            amx::residual(j.b.as<double>(), j.q64.as<double>(), j.r64.as<double>(), j.n, j.partials.as<double>(),
                          S->out_sc(j), s);
            start_inner(*S, j);
            j.phase = kIter;
        }
        // synchronize here: b_host may be released by the caller afterwards
        AMX_HIP(hipStreamSynchronize(s));
        S->workspace_bytes += j.bytes;
        *job_id = j.id;
        S->jobs[j.id] = std::move(J);
        S->tables_dirty = true;
    });
}

int amx_server_step(amx_server* S, int n_iter, int* n_finished)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(S && n_iter > 0, AMX_INVALID_ARGUMENT, "bad argument");
        amx::DeviceGuard dg(S->device);
        const uint64_t before = S->completed;
        if (S->tables_dirty) rebuild_tables(*S);
        bool any = false;
        for (auto& kv : S->jobs) any = any || kv.second->phase != kDone;
        if (any) {
            for (int k = 0; k < n_iter; k++) launch_iteration(*S);
            S->iterations += (uint64_t)n_iter;
            S->steps++;
            advance(*S);
        }
        if (n_finished) *n_finished = (int)(S->completed - before);
    });
}

int amx_server_poll(amx_server* S, int64_t* job_ids, int max_ids, int* n_ids)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(S && n_ids && (job_ids || max_ids == 0), AMX_INVALID_ARGUMENT, "bad argument");
        int k = 0;
        while (k < max_ids && !S->finished.empty()) {
            job_ids[k++] = S->finished.front();
            S->finished.pop_front();
        }
        *n_ids = k;
    });
}

int amx_server_result(amx_server* S, int64_t job_id, double* x_host, amx_solve_stats* stats)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(S, AMX_INVALID_ARGUMENT, "NULL server");
        auto it = S->jobs.find(job_id);
        AMX_REQUIRE(it != S->jobs.end(), AMX_INVALID_ARGUMENT, "unknown job id " + std::to_string(job_id));
        Job& j = *it->second;
        AMX_REQUIRE(j.phase == kDone, AMX_INVALID_STATE, "job not finished");
        amx::DeviceGuard dg(S->device);
        if (x_host) AMX_HIP(hipMemcpy(x_host, j.x.p, (size_t)j.n * 8, hipMemcpyDeviceToHost));
        if (stats) {
            stats->iterations = j.method == AMX_PREC_FP64 ? j.cg_iters : j.inner_total;
            stats->outer_iterations = j.outer;
            stats->converged = j.converged ? 1 : 0;
            stats->precision = j.method | (j.reliable ? AMX_IR_RELIABLE : 0);
            stats->rel_residual = j.rel;
            stats->seconds = std::chrono::duration<double>(j.t_done - j.t_submit).count();
        }
        S->free_slots.push_back(j.slot);
        S->workspace_bytes -= j.bytes;
        S->jobs.erase(it);  // frees the job's device memory
    });
}

int amx_server_get_info(const amx_server* S, amx_server_info* info)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(S && info, AMX_INVALID_ARGUMENT, "NULL argument");
        uint32_t active = 0;
        for (auto& kv : S->jobs) active += kv.second->phase != kDone;
        info->active_jobs = active;
        info->held_jobs = (uint32_t)S->jobs.size();
        info->free_slots = (uint32_t)S->free_slots.size();
        info->workspace_bytes = S->workspace_bytes;
        info->steps = S->steps;
        info->iterations = S->iterations;
        info->completed = S->completed;
    });
}

}  // extern "C"
