// amx_solver.hpp -- device PCG pieces (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace amx {

// Device-resident state of one PCG run.
struct CgScalars {
    double rz, pq, rr, bb, alpha, beta, tol2;
    uint32_t done, iter, max_it, breakdown, counter, pad;
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
// batched iteration pieces: prefix[j] = first block of job j (prefix[n_jobs] = total_blocks)
void cg_pq_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                   hipStream_t s);
void cg_update_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                       hipStream_t s);
void cg_p_batched(bool f64, const CgJob* jobs, const uint32_t* prefix, uint32_t n_jobs, uint32_t total_blocks,
                  hipStream_t s);
void cg_reset(CgScalars* sc, double rtol, uint32_t max_it, hipStream_t s);

}  // namespace amx
