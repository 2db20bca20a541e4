// amx_solver_kernels.hip -- fused vector kernels of the device PCG.
//
// All inner products accumulate in FP64 whatever the vector type.  Reductions
// are deterministic: a fixed grid, a fixed element-to-thread assignment, an
// LDS tree per block, and the partials summed in a fixed tree order by the
// blocks that finish last (per group, then overall).  The scalars
// (alpha, beta, residual norms, the converged flag) never leave the device;
// every kernel becomes a no-op once `done` is set.
#include <hip/hip_runtime.h>

#include "amx_internal.hpp"
#include "amx_solver.hpp"

namespace amx {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kGroup = 32;  // blocks per first-level reduction group (kMaxRedBlocks / kGroup <= kBlock)
static_assert(kMaxRedBlocks / kGroup <= (uint32_t)kBlock && kGroup <= (uint32_t)kBlock, "reduction layout");
constexpr size_t kGroupOff = 3 * kMaxRedBlocks;        // doubles: group sums [NV][kGroup]
constexpr size_t kCntOff = kGroupOff + 3 * kGroup;     // doubles: group counters (uint32) from here
static_assert((kCntOff + kMaxRedBlocks / kGroup) * sizeof(double) <= kPartialsBytes, "partials buffer size");

__device__ __forceinline__ double block_sum(double x, double* sh)
{
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Deposit this block's NV partial sums; return true in the block that must
// finish the reduction (all partials visible to it), with the totals in tot.
// blk / nblk: this block's index among the reduction's blocks and their count.
//
// No device-scope fence: on this chip a release/acquire fence at agent scope
// writes back / invalidates the whole L2 of the XCD (buffer_wbl2 / buffer_inv),
// which, once per block, dominated the solver kernels.  Instead the partials
// are deposited with agent-scope atomic exchanges and the finishing block
// reads them with agent-scope atomic loads; both are performed at the point
// of coherence of all XCDs, not in an XCD's L2.  The ordering argument: the
// exchanges return the previous values (their results feed an asm statement,
// so the compiler must emit returning atomics and wait for their data), and a
// returning atomic's data comes back only after the operation has been
// performed at that point; the counter increment is issued after that wait
// (the asm statement clobbers memory, so nothing moves above it).  Whoever
// sees the counter at its final value therefore finds every deposit already
// performed.  This rests on the hardware's atomics, not on the HIP memory
// model's fences: relaxed atomics carry no ordering there.
//
// Two levels of counters: atomics on one address serialize at the point of
// coherence, so the blocks count per group of kGroup (the group's last block
// sums its group and deposits the group sum), and only the group finishers
// count on the top counter.  The summation order is fixed (a tree over the
// blocks of a group, then a tree over the groups).  Group counters and group
// sums live after the block partials (kGroupOff, kCntOff); the counters start
// at zero (the buffer is cleared when allocated) and each finisher resets its
// own.
template <int NV>
__device__ bool reduce_partials_n(const double (&v)[NV], double* partials, uint32_t* top, double (&tot)[NV],
                                  double* sh, uint32_t blk, uint32_t nblk)
{
    __shared__ uint32_t stage;  // 0: done here, 1: finishes its group, 2: finishes the reduction
    double b[NV];
    for (int k = 0; k < NV; k++) b[k] = block_sum(v[k], sh);
    const uint32_t n_groups = (nblk + kGroup - 1) / kGroup, g = blk / kGroup;
    const uint32_t g0 = g * kGroup, gsize = min(kGroup, nblk - g0);
    double* gpart = partials + kGroupOff;
    uint32_t* gcount = reinterpret_cast<uint32_t*>(partials + kCntOff);
    if (threadIdx.x == 0) {
        double old[NV];
        for (int k = 0; k < NV; k++)
            old[k] = __hip_atomic_exchange(&partials[(size_t)k * nblk + blk], b[k], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        // every exchange has returned before the counter moves
        for (int k = 0; k < NV; k++) __asm__ volatile("" ::"v"(old[k]) : "memory");
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const uint32_t prev = __hip_atomic_fetch_add(&gcount[g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        stage = (prev == gsize - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (stage == 0) return false;
    // the group's finisher: its blocks' partials -> the group sum
    double gs[NV];
    for (int k = 0; k < NV; k++) {
        const double x = threadIdx.x < gsize
                             ? __hip_atomic_load(&partials[(size_t)k * nblk + g0 + threadIdx.x], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)
                             : 0.0;
        gs[k] = block_sum(x, sh);
    }
    if (threadIdx.x == 0) {
        double old[NV];
        for (int k = 0; k < NV; k++)
            old[k] = __hip_atomic_exchange(&gpart[(size_t)k * kGroup + g], gs[k], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&gcount[g], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int k = 0; k < NV; k++) __asm__ volatile("" ::"v"(old[k]) : "memory");
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const uint32_t prev = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        stage = (prev == n_groups - 1) ? 2u : 0u;
    }
    __syncthreads();
    if (stage != 2) return false;
    for (int k = 0; k < NV; k++) {
        const double x = threadIdx.x < n_groups
                             ? __hip_atomic_load(&gpart[(size_t)k * kGroup + threadIdx.x], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)
                             : 0.0;
        tot[k] = block_sum(x, sh);
    }
    if (threadIdx.x == 0) __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

template <int NV>
__device__ bool reduce_partials(const double (&v)[NV], double* partials, uint32_t* counter, double (&tot)[NV], double* sh)
{
    return reduce_partials_n<NV>(v, partials, counter, tot, sh, blockIdx.x, gridDim.x);
}

// end of a CG update (one thread): plain runs stop on convergence or the
// iteration limit; reliable-update runs (d2 > 0) stop there too, or when ||r||
// has fallen below sqrt(d2) of its largest value since the last replacement,
// or when the phase has run twice as long as the previous one (64 iterations
// at least: a recurrence stalled by the operator's error), and leave the
// decision to the replacement (state 1).  A stopped run keeps
// rz and beta of the previous iteration.
__device__ __forceinline__ void cg_update_finish(CgScalars* sc, double rz_new, double rr)
{
    sc->iter += 1;
    sc->rr = rr;
    const bool reliable = sc->d2 > 0.0;
    if (rr <= sc->tol2 * sc->bb || sc->iter >= sc->max_it) {
        sc->done = 1u;
        if (reliable) sc->state = 1u;
        return;
    }
    if (reliable) {
        const double m = rr > sc->rmax2 ? rr : sc->rmax2;
        sc->rmax2 = m;
        const uint32_t len = sc->iter - sc->upd_iter, cap = 2u * sc->phase_len;
        if (rr < sc->d2 * m || len >= (cap > 64u ? cap : 64u)) {
            sc->done = 1u;
            sc->state = 1u;
            return;
        }
    }
    sc->beta = rz_new / sc->rz;
    sc->rz = rz_new;
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_init_kernel(const VT* __restrict__ b, const double* __restrict__ dinv,
                                                        VT* __restrict__ x, VT* __restrict__ r, VT* __restrict__ z,
                                                        VT* __restrict__ p, uint32_t n, double* partials,
                                                        CgScalars* sc)
{
    __shared__ double sh[kBlock];
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const VT bi = b[i];
        const VT zi = (VT)(dinv[i] * (double)bi);
        x[i] = VT(0);
        r[i] = bi;
        z[i] = zi;
        p[i] = zi;
        acc[0] += (double)bi * (double)zi;
        acc[1] += (double)bi * (double)bi;
    }
    double tot[3];
    if (!reduce_partials<3>(acc, partials, &sc->counter, tot, sh)) return;
    if (threadIdx.x == 0) {
        sc->rz = tot[0];
        sc->rr = tot[1];
        sc->bb = tot[1];
        sc->iter = 0;
        sc->done = (tot[1] <= sc->tol2 * tot[1]) || tot[1] == 0.0 ? 1u : 0u;
        sc->breakdown = 0u;
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_pq_kernel(const VT* __restrict__ p, const VT* __restrict__ q, uint32_t n,
                                                      double* partials, CgScalars* sc)
{
    if (sc->done) return;
    __shared__ double sh[kBlock];
    double acc[1] = {0.0};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        acc[0] += (double)p[i] * (double)q[i];
    double tot[1];
    if (!reduce_partials<1>(acc, partials, &sc->counter, tot, sh)) return;
    if (threadIdx.x == 0) {
        sc->pq = tot[0];
        if (!(tot[0] > 0.0)) {  // operator not SPD on this direction
            sc->breakdown = 1u;
            sc->done = 1u;
        } else {
            sc->alpha = sc->rz / tot[0];
        }
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(VT* __restrict__ x, VT* __restrict__ r, VT* __restrict__ z,
                                                          const VT* __restrict__ p, const VT* __restrict__ q,
                                                          const double* __restrict__ dinv, uint32_t n,
                                                          double* partials, CgScalars* sc)
{
    if (sc->done) return;
    __shared__ double sh[kBlock];
    const double alpha = sc->alpha;
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] = (VT)((double)x[i] + alpha * (double)p[i]);
        const double ri = (double)r[i] - alpha * (double)q[i];
        const VT rv = (VT)ri;
        const VT zv = (VT)(dinv[i] * (double)rv);
        r[i] = rv;
        z[i] = zv;
        acc[0] += (double)rv * (double)zv;
        acc[1] += (double)rv * (double)rv;
    }
    double tot[2];
    if (!reduce_partials<2>(acc, partials, &sc->counter, tot, sh)) return;
    if (threadIdx.x == 0) cg_update_finish(sc, tot[0], tot[1]);
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_p_kernel(VT* __restrict__ p, const VT* __restrict__ z, uint32_t n,
                                                     const CgScalars* sc)
{
    if (sc->done) return;
    const double beta = sc->beta;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        p[i] = (VT)((double)z[i] + beta * (double)p[i]);
}

// r = b - Ax (FP64), partial ||r||^2 -> sc->rr
__global__ __launch_bounds__(kBlock) void residual_kernel(const double* __restrict__ b, const double* __restrict__ ax,
                                                         double* __restrict__ r, uint32_t n, double* partials,
                                                         CgScalars* sc)
{
    __shared__ double sh[kBlock];
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double ri = b[i] - ax[i];
        r[i] = ri;
        acc[0] += ri * ri;
        acc[1] += b[i] * b[i];
    }
    double tot[2];
    if (!reduce_partials<2>(acc, partials, &sc->counter, tot, sh)) return;
    if (threadIdx.x == 0) {
        sc->rr = tot[0];
        sc->bb = tot[1];
    }
}

// rs = r / s (s = ||r||) in the inner solve's vector type
template <typename VT>
__global__ void scale_kernel(const double* __restrict__ r, VT* __restrict__ rs, uint32_t n, const CgScalars* sc)
{
    const double inv = sc->rr > 0.0 ? 1.0 / sqrt(sc->rr) : 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) rs[i] = (VT)(r[i] * inv);
}

// x += s * d
template <typename VT>
__global__ void correct_kernel(double* __restrict__ x, const VT* __restrict__ d, uint32_t n, const CgScalars* sc)
{
    const double s = sqrt(sc->rr);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] += s * (double)d[i];
}

// x.y and y.y over the solvers' reduction (the element-to-thread assignment
// and summation order of cg_pq_kernel), result to out[0..1]
__global__ __launch_bounds__(kBlock) void dot2_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                     uint32_t n, double* partials, uint32_t* counter,
                                                     double* __restrict__ out)
{
    __shared__ double sh[kBlock];
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        acc[0] += x[i] * y[i];
        acc[1] += y[i] * y[i];
    }
    double tot[2];
    if (!reduce_partials<2>(acc, partials, counter, tot, sh)) return;
    if (threadIdx.x == 0) {
        out[0] = tot[0];
        out[1] = tot[1];
    }
}

__global__ void invert_kernel(const double* __restrict__ d, double* __restrict__ dinv, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dinv[i] = 1.0 / d[i];
}

// ---- batched: many jobs per launch ---------------------------------------
// Block -> (job j, local block lb) through the block prefix; every job keeps
// the grid of its single-job launch (reduction_grid(n) blocks, same stride),
// so its sums are bit-identical to a solve of that job alone.
__device__ __forceinline__ uint32_t job_of(const uint32_t* __restrict__ prefix, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_pq_batched_kernel(const CgJob* __restrict__ jobs,
                                                              const uint32_t* __restrict__ prefix, uint32_t n_jobs)
{
    const uint32_t gb = __builtin_amdgcn_readfirstlane(blockIdx.x);
    const uint32_t j = job_of(prefix, n_jobs, gb);
    const CgJob& jb = jobs[j];
    CgScalars* sc = jb.sc;
    if (sc->done) return;
    __shared__ double sh[kBlock];
    const uint32_t lb = gb - prefix[j], nblk = jb.nblk, n = jb.n;
    const VT* p = static_cast<const VT*>(jb.p);
    const VT* q = static_cast<const VT*>(jb.q);
    double acc[1] = {0.0};
    for (uint32_t i = lb * kBlock + threadIdx.x; i < n; i += nblk * kBlock) acc[0] += (double)p[i] * (double)q[i];
    double tot[1];
    if (!reduce_partials_n<1>(acc, jb.partials, &sc->counter, tot, sh, lb, nblk)) return;
    if (threadIdx.x == 0) {
        sc->pq = tot[0];
        if (!(tot[0] > 0.0)) {
            sc->breakdown = 1u;
            sc->done = 1u;
        } else {
            sc->alpha = sc->rz / tot[0];
        }
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_update_batched_kernel(const CgJob* __restrict__ jobs,
                                                                  const uint32_t* __restrict__ prefix, uint32_t n_jobs)
{
    const uint32_t gb = __builtin_amdgcn_readfirstlane(blockIdx.x);
    const uint32_t j = job_of(prefix, n_jobs, gb);
    const CgJob& jb = jobs[j];
    CgScalars* sc = jb.sc;
    if (sc->done) return;
    __shared__ double sh[kBlock];
    const uint32_t lb = gb - prefix[j], nblk = jb.nblk, n = jb.n;
    VT* x = static_cast<VT*>(jb.x);
    VT* r = static_cast<VT*>(jb.r);
    VT* z = static_cast<VT*>(jb.z);
    const VT* p = static_cast<const VT*>(jb.p);
    const VT* q = static_cast<const VT*>(jb.q);
    const double* dinv = jb.dinv;
    const double alpha = sc->alpha;
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = lb * kBlock + threadIdx.x; i < n; i += nblk * kBlock) {
        x[i] = (VT)((double)x[i] + alpha * (double)p[i]);
        const double ri = (double)r[i] - alpha * (double)q[i];
        const VT rv = (VT)ri;
        const VT zv = (VT)(dinv[i] * (double)rv);
        r[i] = rv;
        z[i] = zv;
        acc[0] += (double)rv * (double)zv;
        acc[1] += (double)rv * (double)rv;
    }
    double tot[2];
    if (!reduce_partials_n<2>(acc, jb.partials, &sc->counter, tot, sh, lb, nblk)) return;
    if (threadIdx.x == 0) cg_update_finish(sc, tot[0], tot[1]);
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_p_batched_kernel(const CgJob* __restrict__ jobs,
                                                             const uint32_t* __restrict__ prefix, uint32_t n_jobs)
{
    const uint32_t gb = __builtin_amdgcn_readfirstlane(blockIdx.x);
    const uint32_t j = job_of(prefix, n_jobs, gb);
    const CgJob& jb = jobs[j];
    const CgScalars* sc = jb.sc;
    if (sc->done) return;
    const uint32_t lb = gb - prefix[j], nblk = jb.nblk, n = jb.n;
    VT* p = static_cast<VT*>(jb.p);
    const VT* z = static_cast<const VT*>(jb.z);
    const double beta = sc->beta;
    for (uint32_t i = lb * kBlock + threadIdx.x; i < n; i += nblk * kBlock) p[i] = (VT)((double)z[i] + beta * (double)p[i]);
}

// ---- reliable updates (mixed-precision CG with residual replacement) ----
// x += xs (FP64), xs = 0
template <typename VT>
__global__ __launch_bounds__(kBlock) void ru_flush_kernel(double* __restrict__ x, VT* __restrict__ xs, uint32_t n)
{
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] += (double)xs[i];
        xs[i] = VT(0);
    }
}

// Replace the recurrence's residual by the true one: r = b - Ax in FP64,
// rs = (VT) r, zs = dinv rs; ||r||^2 and rs.zs in FP64.  The finishing block
// sets rr (true), rz, beta for the direction update that follows (rz_new /
// rz_old, or 0 to restart the direction) and done (converged, or a limit
// reached).  init: first call (Ax = 0): also bb = ||b||^2, p = zs, counters
// cleared.
template <typename VT, bool INIT>
__global__ __launch_bounds__(kBlock) void ru_replace_kernel(const double* __restrict__ b, const double* __restrict__ ax,
                                                           VT* __restrict__ rs, VT* __restrict__ zs,
                                                           VT* __restrict__ ps, const double* __restrict__ dinv,
                                                           uint32_t n, double* partials, CgScalars* sc)
{
    __shared__ double sh[kBlock];
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double ri = b[i] - ax[i];
        const VT rv = (VT)ri;
        const VT zv = (VT)(dinv[i] * (double)rv);
        rs[i] = rv;
        zs[i] = zv;
        if constexpr (INIT) ps[i] = zv;
        acc[0] += ri * ri;
        acc[1] += (double)rv * (double)zv;
    }
    double tot[2];
    if (!reduce_partials<2>(acc, partials, &sc->counter, tot, sh)) return;
    if (threadIdx.x == 0) {
        if constexpr (INIT) {
            sc->bb = tot[0];
            sc->iter = 0;
            sc->breakdown = 0u;
            sc->beta = 0.0;
            sc->n_upd = 0u;
            sc->upd_iter = 0u;
            sc->phase_len = 0u;
        } else {
            // keep the direction while the recurrence tracked the true residual
            // (true norm at most 4x the recurrence's); otherwise the
            // low-precision operator's error dominates and the direction
            // restarts (beta = 0), as in a restarted refinement step
            const bool tracked = tot[0] <= 16.0 * sc->rr;
            sc->beta = (tracked && sc->rz != 0.0) ? tot[1] / sc->rz : 0.0;
            sc->n_upd += 1u;
            sc->phase_len = sc->iter - sc->upd_iter;
            sc->upd_iter = sc->iter;
        }
        sc->rr_prev = tot[0];
        sc->rr = tot[0];
        sc->rz = tot[1];
        sc->rmax2 = tot[0];
        sc->done = (tot[0] <= sc->tol2 * sc->bb || tot[0] == 0.0 || sc->iter >= sc->max_it ||
                    sc->n_upd >= sc->max_upd)
                       ? 1u
                       : 0u;
        sc->state = sc->done ? 2u : 0u;
    }
}

// restart a CG: scalars zeroed, tolerance and iteration limit set (a kernel,
// so it can sit in a stream or a graph without host staging)
__global__ void cg_reset_kernel(CgScalars* sc, double tol2, uint32_t max_it, double d2, uint32_t max_upd)
{
    if (threadIdx.x != 0) return;
    CgScalars h{};
    h.tol2 = tol2;
    h.max_it = max_it;
    h.d2 = d2;
    h.max_upd = max_upd;
    *sc = h;
}

}  // namespace

uint32_t reduction_grid(uint32_t n)
{
    const uint32_t want = (n + kBlock * 4 - 1) / (kBlock * 4);
    return want < 1 ? 1 : (want > kMaxRedBlocks ? kMaxRedBlocks : want);
}

void cg_init(bool f64, const void* b, const double* dinv, void* x, void* r, void* z, void* p, uint32_t n,
             double* partials, CgScalars* sc, hipStream_t s)
{
    const uint32_t g = reduction_grid(n);
    if (f64) cg_init_kernel<double><<<g, kBlock, 0, s>>>((const double*)b, dinv, (double*)x, (double*)r, (double*)z, (double*)p, n, partials, sc);
    else cg_init_kernel<float><<<g, kBlock, 0, s>>>((const float*)b, dinv, (float*)x, (float*)r, (float*)z, (float*)p, n, partials, sc);
    AMX_HIP(hipGetLastError());
}

void cg_pq(bool f64, const void* p, const void* q, uint32_t n, double* partials, CgScalars* sc, hipStream_t s)
{
    const uint32_t g = reduction_grid(n);
    if (f64) cg_pq_kernel<double><<<g, kBlock, 0, s>>>((const double*)p, (const double*)q, n, partials, sc);
    else cg_pq_kernel<float><<<g, kBlock, 0, s>>>((const float*)p, (const float*)q, n, partials, sc);
    AMX_HIP(hipGetLastError());
}

void cg_update(bool f64, void* x, void* r, void* z, const void* p, const void* q, const double* dinv, uint32_t n,
               double* partials, CgScalars* sc, hipStream_t s)
{
    const uint32_t g = reduction_grid(n);
    if (f64) cg_update_kernel<double><<<g, kBlock, 0, s>>>((double*)x, (double*)r, (double*)z, (const double*)p, (const double*)q, dinv, n, partials, sc);
    else cg_update_kernel<float><<<g, kBlock, 0, s>>>((float*)x, (float*)r, (float*)z, (const float*)p, (const float*)q, dinv, n, partials, sc);
    AMX_HIP(hipGetLastError());
}

void cg_p(bool f64, void* p, const void* z, uint32_t n, const CgScalars* sc, hipStream_t s)
{
    const uint32_t g = reduction_grid(n);
    if (f64) cg_p_kernel<double><<<g, kBlock, 0, s>>>((double*)p, (const double*)z, n, sc);
    else cg_p_kernel<float><<<g, kBlock, 0, s>>>((float*)p, (const float*)z, n, sc);
    AMX_HIP(hipGetLastError());
}

void residual(const double* b, const double* ax, double* r, uint32_t n, double* partials, CgScalars* sc, hipStream_t s)
{
    residual_kernel<<<reduction_grid(n), kBlock, 0, s>>>(b, ax, r, n, partials, sc);
    AMX_HIP(hipGetLastError());
}

void scale_to_f32(const double* r, float* r32, uint32_t n, const CgScalars* sc, hipStream_t s)
{
    scale_kernel<float><<<reduction_grid(n), kBlock, 0, s>>>(r, r32, n, sc);
    AMX_HIP(hipGetLastError());
}

void scale_to_f64(const double* r, double* rs, uint32_t n, const CgScalars* sc, hipStream_t s)
{
    scale_kernel<double><<<reduction_grid(n), kBlock, 0, s>>>(r, rs, n, sc);
    AMX_HIP(hipGetLastError());
}

void correct(double* x, const float* d32, uint32_t n, const CgScalars* sc, hipStream_t s)
{
    correct_kernel<float><<<reduction_grid(n), kBlock, 0, s>>>(x, d32, n, sc);
    AMX_HIP(hipGetLastError());
}

void correct64(double* x, const double* d, uint32_t n, const CgScalars* sc, hipStream_t s)
{
    correct_kernel<double><<<reduction_grid(n), kBlock, 0, s>>>(x, d, n, sc);
    AMX_HIP(hipGetLastError());
}

void cg_pq_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                   hipStream_t s)
{
    if (!total_blocks) return;
    if (f64) cg_pq_batched_kernel<double><<<total_blocks, kBlock, 0, s>>>(jobs, prefix, n_jobs);
    else cg_pq_batched_kernel<float><<<total_blocks, kBlock, 0, s>>>(jobs, prefix, n_jobs);
    AMX_HIP(hipGetLastError());
}

void cg_update_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                       hipStream_t s)
{
    if (!total_blocks) return;
    if (f64) cg_update_batched_kernel<double><<<total_blocks, kBlock, 0, s>>>(jobs, prefix, n_jobs);
    else cg_update_batched_kernel<float><<<total_blocks, kBlock, 0, s>>>(jobs, prefix, n_jobs);
    AMX_HIP(hipGetLastError());
}

void cg_p_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                  hipStream_t s)
{
    if (!total_blocks) return;
    if (f64) cg_p_batched_kernel<double><<<total_blocks, kBlock, 0, s>>>(jobs, prefix, n_jobs);
    else cg_p_batched_kernel<float><<<total_blocks, kBlock, 0, s>>>(jobs, prefix, n_jobs);
    AMX_HIP(hipGetLastError());
}

void ru_flush(bool f64, double* x, void* xs, uint32_t n, hipStream_t s)
{
    const uint32_t g = reduction_grid(n);
    if (f64) ru_flush_kernel<double><<<g, kBlock, 0, s>>>(x, (double*)xs, n);
    else ru_flush_kernel<float><<<g, kBlock, 0, s>>>(x, (float*)xs, n);
    AMX_HIP(hipGetLastError());
}

void ru_replace(bool f64, bool init, const double* b, const double* ax, void* rs, void* zs, void* ps,
                const double* dinv, uint32_t n, double* partials, CgScalars* sc, hipStream_t s)
{
    const uint32_t g = reduction_grid(n);
#define AMX_RU(VT, I)                                                                                              \
    ru_replace_kernel<VT, I><<<g, kBlock, 0, s>>>(b, ax, (VT*)rs, (VT*)zs, (VT*)ps, dinv, n, partials, sc)
    if (f64) { if (init) AMX_RU(double, true); else AMX_RU(double, false); }
    else { if (init) AMX_RU(float, true); else AMX_RU(float, false); }
#undef AMX_RU
    AMX_HIP(hipGetLastError());
}

void cg_reset(CgScalars* sc, double rtol, uint32_t max_it, hipStream_t s, double delta, uint32_t max_upd)
{
    cg_reset_kernel<<<1, 64, 0, s>>>(sc, rtol * rtol, max_it, delta * delta, max_upd);
    AMX_HIP(hipGetLastError());
}

void dot2(const double* x, const double* y, uint32_t n, uint32_t n_blocks, double* partials, uint32_t* counter,
          double* out, hipStream_t s)
{
    dot2_kernel<<<n_blocks, kBlock, 0, s>>>(x, y, n, partials, counter, out);
    AMX_HIP(hipGetLastError());
}

void invert(const double* d, double* dinv, uint32_t n, hipStream_t s)
{
    if (n == 0) return;
    invert_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(d, dinv, n);
    AMX_HIP(hipGetLastError());
}

}  // namespace amx
