// amx_nsym.hpp -- near-symmetric packed records of one-element FP64 tiles.
//
// Element matrices built by quadrature (ke_hex20_iso and its like) are
// symmetric in value but not in every bit: K[r][c] and K[c][r] come out of
// different operation orders and differ in their last bits, and entries that
// are zero in exact arithmetic come out as +-1e-17 noise of either sign.  A
// record stores such a matrix losslessly in about 0.6 of the whole n*n FP64
// values, so that the element kernel reads K[r][c] bit for bit:
//
//   U  the upper triangle with the diagonal, row-major: (r <= c) at
//      r*n - r(r-1)/2 + (c - r); NSP = n(n+1)/2 rounded up to even values.
//   X  the exception values of the group (below), nx per element, in the
//      row-major order of their (r, c), each stored with its low 32-bit word
//      plus 128 (mod 2^32); NXP = nx rounded up to even.
//   D  one signed byte per lower entry (r > c): the low 32-bit word of K[r][c]
//      minus that of K[c][r], when both have the same high word and the
//      difference is in [-127, 127] (no carry reaches the high word); -128
//      marks an exception.  Bytes of row r in 16-byte chunks: chunk q holds
//      columns 16q .. 16q+15 (zero where c >= r) and exists for rows
//      r = 16q+1 .. n-1; chunk-major, D[q] then D[q+1]: the chunk of row r is
//      at dbase(q) + r - 16q - 1 (in chunks).
//
// So K[r][c] for r > c is U[c][r] with delta[r][c] added to its low word, and
// where delta is -128 it is the row's next X value with -128 added to its low
// word.  Exception positions are those where any matrix of the group fails
// the byte code; they are the same for every element of the group.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace amx {

__host__ __device__ constexpr uint32_t nsym_ns(uint32_t n) { return n * (n + 1) / 2; }
__host__ __device__ constexpr uint32_t nsym_nsp(uint32_t n) { return (nsym_ns(n) + 1u) & ~1u; }
__host__ __device__ constexpr uint32_t nsym_nxp(uint32_t nx) { return (nx + 1u) & ~1u; }
// delta chunks per row of the last row, and chunks before chunk column q
__host__ __device__ constexpr uint32_t nsym_nq(uint32_t n) { return (n + 14u) / 16u; }
__host__ __device__ constexpr uint32_t nsym_dbase(uint32_t n, uint32_t q)
{
    return q == 0 ? 0u : nsym_dbase(n, q - 1) + (n - 1u - 16u * (q - 1));
}
__host__ __device__ constexpr uint32_t nsym_dchunks(uint32_t n) { return nsym_dbase(n, nsym_nq(n)); }
__host__ __device__ constexpr uint64_t nsym_record_bytes_of(uint32_t n, uint32_t nx)
{
    return 8ull * (nsym_nsp(n) + nsym_nxp(nx)) + 16ull * nsym_dchunks(n);
}
__host__ __device__ constexpr uint32_t nsym_uidx(uint32_t n, uint32_t a, uint32_t b)  // a <= b
{
    return a * n - (a * (a - 1u)) / 2u + (b - a);
}

// Packed symmetric records with exceptions (FP32 / BF16 storage of the same
// groups): rounded to FP32 or BF16, such matrices are bitwise symmetric except
// at their exception positions, so a record holds U (the BF16 diagonal slot
// is 0: that storage keeps its diagonal in FP32 beside the matrices) and then
// the group's nx exception values of the storage type, in row-major order of
// their (r, c), padded to 16 bytes.
__host__ __device__ constexpr uint32_t psx_record_values(uint32_t n, uint32_t nx, uint32_t ksize)
{
    return ((nsym_ns(n) + nx) * ksize + 15u) / 16u * 16u / ksize;
}

// one value K[r][c] of a record (set-up kernels; the element kernel reads the
// record through LDS).  xidx[r*n + c]: index into X of an exception, else any.
__device__ __forceinline__ double nsym_value(const double* __restrict__ rec, uint32_t n, uint32_t nx,
                                             const uint16_t* __restrict__ xidx, uint32_t r, uint32_t c)
{
    if (c >= r) return rec[nsym_uidx(n, r, c)];
    const int8_t* D = reinterpret_cast<const int8_t*>(rec + nsym_nsp(n) + nsym_nxp(nx));
    const uint32_t q = c / 16u;
    const int32_t d = D[(uint64_t)(nsym_dbase(n, q) + r - 16u * q - 1u) * 16u + (c - 16u * q)];
    uint64_t bits = d == -128 ? __double_as_longlong(rec[nsym_nsp(n) + xidx[r * n + c]])
                              : __double_as_longlong(rec[nsym_uidx(n, c, r)]);
    const uint32_t lo = (uint32_t)bits + (uint32_t)d;
    bits = (bits & 0xffffffff00000000ull) | lo;
    return __longlong_as_double((long long)bits);
}

}  // namespace amx
