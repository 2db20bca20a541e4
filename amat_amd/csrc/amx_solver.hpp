// amx_solver.hpp -- device PCG pieces (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace amx {

// Device-resident state of one PCG run.  Reliable-update runs (d2 > 0) also
// keep rmax2, the largest recurrence ||r||^2 since the last replacement, and
// `state`: 1 = the recurrence stopped for a replacement (its ||r|| fell below
// sqrt(d2) of rmax, it ran twice as long as the previous phase (at least 64
// iterations) without that, or it reports convergence or the iteration
// limit), 2 =
// final (decided on the true residual by the replacement, or the replacement
// limit reached).
struct CgScalars {
    double rz, pq, rr, bb, alpha, beta, tol2, rmax2, d2, rr_prev;  // rr_prev: true ||r||^2 at the last replacement
    uint32_t done, iter, max_it, breakdown, counter, state;
    uint32_t n_upd, max_upd;      // replacements done / allowed (reliable updates)
    uint32_t upd_iter, phase_len; // iteration of the last replacement; iterations of the phase before it
};

// one job of a batched CG launch (vectors of the launch's type)
struct CgJob {
    void* x;
    void* r;
    void* z;
    void* p;
    const void* q;
    const double* dinv;
    double* partials;
    CgScalars* sc;
    uint32_t n, nblk;  // nblk = reduction_grid(n)
};

// Reductions: at most kMaxRedBlocks blocks; their partial sums, group sums and
// counters live in a per-solve buffer of kPartialsBytes, zeroed when allocated.
constexpr uint32_t kMaxRedBlocks = 1024;
constexpr size_t kPartialsBytes = 4096 * sizeof(double);
// FP64 CG: residual replacements allowed after the recurrence stops (the
// first one measures the true residual; more run only when it misses rtol)
constexpr uint32_t kCgReplacements = 4;
uint32_t reduction_grid(uint32_t n);
void cg_init(bool f64, const void* b, const double* dinv, void* x, void* r, void* z, void* p, uint32_t n,
             double* partials, CgScalars* sc, hipStream_t s);
void cg_pq(bool f64, const void* p, const void* q, uint32_t n, double* partials, CgScalars* sc, hipStream_t s);
void cg_update(bool f64, void* x, void* r, void* z, const void* p, const void* q, const double* dinv, uint32_t n,
               double* partials, CgScalars* sc, hipStream_t s);
void cg_p(bool f64, void* p, const void* z, uint32_t n, const CgScalars* sc, hipStream_t s);
void residual(const double* b, const double* ax, double* r, uint32_t n, double* partials, CgScalars* sc, hipStream_t s);
void scale_to_f32(const double* r, float* r32, uint32_t n, const CgScalars* sc, hipStream_t s);
void correct(double* x, const float* d32, uint32_t n, const CgScalars* sc, hipStream_t s);
void scale_to_f64(const double* r, double* rs, uint32_t n, const CgScalars* sc, hipStream_t s);
void correct64(double* x, const double* d, uint32_t n, const CgScalars* sc, hipStream_t s);
void invert(const double* d, double* dinv, uint32_t n, hipStream_t s);
// out[0..1] = x.y, y.y (FP64) by the solvers' reduction over n_blocks blocks
// (partials: kPartialsBytes with zero counters; counter: zero, left at zero)
void dot2(const double* x, const double* y, uint32_t n, uint32_t n_blocks, double* partials, uint32_t* counter,
          double* out, hipStream_t s);
// batched iteration pieces: prefix[j] = first block of job j (prefix[n_jobs] = total_blocks)
void cg_pq_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                   hipStream_t s);
void cg_update_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                       hipStream_t s);
void cg_p_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                  hipStream_t s);
void cg_reset(CgScalars* sc, double rtol, uint32_t max_it, hipStream_t s, double delta = 0.0,
              uint32_t max_upd = 0);
// reliable updates: x += xs, xs = 0; and the residual replacement from Ax
// (init: Ax = 0 at the start, also sets bb and p = z)
void ru_flush(bool f64, double* x, void* xs, uint32_t n, hipStream_t s);
void ru_replace(bool f64, bool init, const double* b, const double* ax, void* rs, void* zs, void* ps,
                const double* dinv, uint32_t n, double* partials, CgScalars* sc, hipStream_t s);

}  // namespace amx
