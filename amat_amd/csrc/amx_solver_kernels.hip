// amx_solver_kernels.hip -- fused vector kernels of the device PCG.
//
// All inner products accumulate in FP64 whatever the vector type.  Reductions
// are deterministic: a fixed grid, a fixed element-to-thread assignment, an
// LDS tree per block, and the partials summed in block order by whichever
// block finishes last (device-scope fences around a counter).  The scalars
// (alpha, beta, residual norms, the converged flag) never leave the device;
// every kernel becomes a no-op once `done` is set.
#include <hip/hip_runtime.h>

#include "amx_internal.hpp"
#include "amx_solver.hpp"

namespace amx {

namespace {

constexpr int kBlock = 256;

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
// are deposited with agent-scope atomic exchanges (performed at the point of
// coherence of all XCDs) whose completion is awaited before the counter is
// incremented, and the finishing block reads them with agent-scope atomic
// loads.  Summation order is unchanged (fixed, by block index).
template <int NV>
__device__ bool reduce_partials_n(const double (&v)[NV], double* partials, uint32_t* counter, double (&tot)[NV],
                                  double* sh, uint32_t blk, uint32_t nblk)
{
    __shared__ bool last;
    double b[NV];
    for (int k = 0; k < NV; k++) b[k] = block_sum(v[k], sh);
    if (threadIdx.x == 0) {
        double old = 0.0;
        for (int k = 0; k < NV; k++)
            old += __hip_atomic_exchange(&partials[(size_t)k * nblk + blk], b[k], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        // every exchange has returned (its value is used) before the counter moves
        const uint32_t one = old == old ? 1u : 1u;
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const uint32_t prev = __hip_atomic_fetch_add(counter, one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = (prev == nblk - 1);
    }
    __syncthreads();
    if (!last) return false;
    for (int k = 0; k < NV; k++) {
        double s = 0.0;
        for (uint32_t i = threadIdx.x; i < nblk; i += kBlock)
            s += __hip_atomic_load(&partials[(size_t)k * nblk + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        tot[k] = block_sum(s, sh);
    }
    if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

template <int NV>
__device__ bool reduce_partials(const double (&v)[NV], double* partials, uint32_t* counter, double (&tot)[NV], double* sh)
{
    return reduce_partials_n<NV>(v, partials, counter, tot, sh, blockIdx.x, gridDim.x);
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
    if (threadIdx.x == 0) {
        sc->iter += 1;
        sc->rr = tot[1];
        if (tot[1] <= sc->tol2 * sc->bb || sc->iter >= sc->max_it) {
            sc->done = 1u;
        } else {
            sc->beta = tot[0] / sc->rz;
            sc->rz = tot[0];
        }
    }
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
    if (threadIdx.x == 0) {
        sc->iter += 1;
        sc->rr = tot[1];
        if (tot[1] <= sc->tol2 * sc->bb || sc->iter >= sc->max_it) {
            sc->done = 1u;
        } else {
            sc->beta = tot[0] / sc->rz;
            sc->rz = tot[0];
        }
    }
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

// restart a CG: scalars zeroed, tolerance and iteration limit set (a kernel,
// so it can sit in a stream or a graph without host staging)
__global__ void cg_reset_kernel(CgScalars* sc, double tol2, uint32_t max_it)
{
    if (threadIdx.x != 0) return;
    CgScalars h{};
    h.tol2 = tol2;
    h.max_it = max_it;
    *sc = h;
}

}  // namespace

uint32_t reduction_grid(uint32_t n)
{
    const uint32_t want = (n + kBlock * 4 - 1) / (kBlock * 4);
    return want < 1 ? 1 : (want > 1024 ? 1024 : want);
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

void cg_reset(CgScalars* sc, double rtol, uint32_t max_it, hipStream_t s)
{
    cg_reset_kernel<<<1, 64, 0, s>>>(sc, rtol * rtol, max_it);
    AMX_HIP(hipGetLastError());
}

void invert(const double* d, double* dinv, uint32_t n, hipStream_t s)
{
    if (n == 0) return;
    invert_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(d, dinv, n);
    AMX_HIP(hipGetLastError());
}

}  // namespace amx
