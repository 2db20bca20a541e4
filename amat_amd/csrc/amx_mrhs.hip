// amx_mrhs.hip -- multi-RHS element products on MFMA: V = sum_e P_e^T K_e P_e U
// for U, V of nrhs columns (DoF-major, columns contiguous), one persistent
// launch over several operators (meshes of different element types).
//
// A wave owns one element tile x one RHS tile (32 * NT columns).
//   FP32: v_mfma_f32_32x32x2_f32; A fragment lane l = A[l%32][k0 + l/32];
//   BF16: v_mfma_f32_32x32x16_bf16; A fragment lane l = A[l%32][k0 + 8(l/32) .. +7];
//         U is rounded to BF16 (nearest even) on load, accumulation is FP32;
//   C fragment (both): lane l, register i -> row 8(i/4) + 4(l/32) + i%4, column l%32.
// 60-DoF elements (hex20 elasticity) are one tile of two 32-row M-tiles and
// their matrices are stored in MFMA fragment order ("compact": rows 60..63 and
// the K padding are not stored), so every A load is a contiguous 16-byte load
// per lane.  Smaller elements are packed block-diagonally, G = 32/n elements per
// 32-row tile (4 hex8), from per-element matrices (FP32 column-major, BF16
// row-major with 8-aligned rows); the off-diagonal zeros are never stored.
// The element rows reach V either through the deterministic two-pass scatter
// (rows stored per element, then summed per DoF in ascending element order:
// bit-reproducible, independent of the launch shape) or FP32 atomics.
// Work split: every segment (operator) is cut into 8 contiguous slices, one per
// XCD (workgroups are dispatched round-robin over the 8 XCDs), so elements that
// share nodes run on the same L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "amx_nsym.hpp"
#include "amx_op.hpp"

using amx::DevBuf;
using amx::fail;
using amx::guarded;

namespace amx {
namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kMaxSeg = 8;
constexpr uint32_t kXcd = 8;

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

struct Seg {
    const void* K;          // see the layouts above
    const uint32_t* dofs;   // [tile][dstride] local DoF ids (kNone: padding)
    const float* U;
    float* V;
    float* ve;              // deterministic scatter: element rows [e*n + r][nrhs]; NULL: atomics into V
    uint32_t n, n_elem, n_tiles, ngroups, nrhs, dstride;
};

struct SegTable {
    Seg s[kMaxSeg];
    uint32_t n_seg;
    uint32_t chunked;  // 0: waves interleave over their XCD's slice; 1: each wave takes a contiguous chunk
};

__host__ __device__ constexpr int tile_elems(int n) { return n > 32 ? 1 : 32 / n; }
__host__ __device__ constexpr int tile_rows(int n) { return n > 32 ? n : (32 / n) * n; }
__host__ __device__ constexpr int mtiles(int n) { return (tile_rows(n) + 31) / 32; }
// K dimension of a small-element tile padded to the MFMA step (2 for FP32, 16 for BF16)
__host__ __device__ constexpr int kdim(int n, bool bf16) { return bf16 ? (tile_rows(n) + 15) / 16 * 16 : (tile_rows(n) + 1) / 2 * 2; }
__host__ __device__ constexpr int bf16_ld(int n) { return (n + 7) / 8 * 8; }
__host__ __device__ constexpr int dof_stride(int n, bool bf16) { return n == 60 ? 64 : kdim(n, bf16); }

__device__ __forceinline__ uint32_t rl(uint32_t x, int lane) { return (uint32_t)__builtin_amdgcn_readlane((int)x, lane); }

__host__ __device__ __forceinline__ unsigned short f2bf(float f)
{
    uint32_t b;
    memcpy(&b, &f, 4);
    return (unsigned short)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

// ---- 60-DoF elements: padded positions and compact MFMA-order storage ----
// A 60-DoF element is laid out on 64 "positions": p = 32 h + s holds element
// DoF 30 h + s for s < 30; s = 30, 31 are padding.  Rows (M-tile m = h) and
// columns (the K dimension) use the same positions, as does the DoF list
// (padding entries repeat DoF 0, a valid row).  Padding rows are never
// stored; padding columns multiply zeros of the storage.
//   FP32, k-step ks (0..29) = 4 g + j, lane (h, s): A = K[30 m + s][30 h + ks]
//     g < 7 : g*480 + m*240 + (30 h + s)*4 + j      (16-byte load per lane)
//     g = 7 : 3360 + m*120 + (30 h + s)*2 + j       (8-byte load)
//   BF16, k-step ks (0..3), lane (h, s), v (0..7): A = K[30 m + s][col(16 ks + 8 h + v)]
//     per M-tile m: m*1800 +
//       ks even        : (ks/2)*480 + (30 h + s)*8 + v   (16 bytes)
//       ks odd, h = 0  : 960 + (ks/2)*240 + s*8 + v      (16 bytes)
//       ks odd, h = 1  : 1440 + (ks/2)*180 + s*6 + v     (12 bytes: positions 24..29)
__host__ __device__ inline uint32_t pos60(bool bf16, uint32_t r, uint32_t c)
{
    const uint32_t m = r / 30, s = r % 30;
    if (!bf16) {
        const uint32_t h = c / 30, ks = c % 30, g = ks / 4, j = ks % 4;
        if (g < 7) return g * 480 + m * 240 + (30 * h + s) * 4 + j;
        return 3360 + m * 120 + (30 * h + s) * 2 + j;
    }
    const uint32_t p = 32 * (c / 30) + c % 30, ks = p / 16, h = (p % 16) / 8, v = p % 8;
    uint32_t sub;
    if (ks % 2 == 0) sub = (ks / 2) * 480 + (30 * h + s) * 8 + v;
    else if (h == 0) sub = 960 + (ks / 2) * 240 + s * 8 + v;
    else sub = 1440 + (ks / 2) * 180 + s * 6 + v;
    return m * 1800 + sub;
}

// atomically add the C fragments acc[m][q] (rows of the tile) into V
// (deterministic form: the rows go to their element rows of ve, row t*ROWS + tile row)
template <int MT, int NT, int ROWS>
__device__ __forceinline__ void scatter_rows(const Seg& sg, uint32_t t, uint32_t mydof, uint32_t col0,
                                             const f32x16 (&acc)[MT][NT])
{
    const uint32_t lane = threadIdx.x & 63u, half = lane >> 5, col_l = lane & 31u;
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r0 = m * 32 + 8 * (i / 4) + (i % 4);  // row of half 0; half 1 has r0 + 4
            if (r0 >= ROWS) continue;
            const uint32_t d0 = rl(mydof, r0), d1 = (r0 + 4 < ROWS) ? rl(mydof, r0 + 4) : kNone;
            const uint32_t d = half ? d1 : d0;
#pragma unroll
            for (int q = 0; q < NT; q++) {
                const uint32_t col = col0 + q * 32 + col_l;
                if (col < sg.nrhs && d != kNone) {
                    if (sg.ve) {
                        const uint64_t row = (uint64_t)t * ROWS + (uint32_t)r0 + 4 * half;
                        sg.ve[row * sg.nrhs + col] = acc[m][q][i];
                    } else {
                        unsafeAtomicAdd(&sg.V[(uint64_t)d * sg.nrhs + col], acc[m][q][i]);
                    }
                }
            }
        }
}

// U columns this lane reads for each RHS sub-tile q: clamped to the last
// column, so every load is in bounds and branch-free; the C columns of
// clamped lanes are never stored.
template <int NT>
__device__ __forceinline__ void rhs_columns(const Seg& sg, uint32_t col0, uint32_t (&uc)[NT])
{
    const uint32_t col_l = threadIdx.x & 31u;
#pragma unroll
    for (int q = 0; q < NT; q++) uc[q] = min(col0 + q * 32 + col_l, sg.nrhs - 1);
}

__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi)
{
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, bf2{(__bf16)lo, (__bf16)hi});
}

// A fragments of one 60-DoF element (compact MFMA order).  Padding lanes
// (s = 30, 31) load lane s = 29's data: their rows are never stored.
template <bool BF16> struct Frag60;
template <> struct Frag60<false> {
    f32x4 a[2][7];
    f32x2 t[2];
    __device__ __forceinline__ void load(const void* K, uint32_t e)
    {
        const uint32_t lane = threadIdx.x & 63u, h = lane >> 5, sl = min(lane & 31u, 29u);
        const float* Kb = static_cast<const float*>(K) + (uint64_t)e * 3600 + (30 * h + sl) * 4;
        const float* Kt = static_cast<const float*>(K) + (uint64_t)e * 3600 + 3360 + (30 * h + sl) * 2;
#pragma unroll
        for (int m = 0; m < 2; m++) {
#pragma unroll
            for (int g = 0; g < 7; g++)
                a[m][g] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(Kb + g * 480 + m * 240));
            t[m] = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(Kt + m * 120));
        }
    }
    __device__ __forceinline__ float get(int m, int ks) const
    {
        const int g = ks / 4, j = ks % 4;
        return g < 7 ? a[m][g < 7 ? g : 0][j] : t[m][j & 1];
    }
};
template <> struct Frag60<true> {
    u16x8 a[2][4];
    __device__ __forceinline__ void load(const void* K, uint32_t e)
    {
        typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const uint32_t lane = threadIdx.x & 63u, h = lane >> 5, sl = min(lane & 31u, 29u);
        const unsigned short* Kb = static_cast<const unsigned short*>(K) + (uint64_t)e * 3600;
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const unsigned short* Km = Kb + m * 1800;
#pragma unroll
            for (int k2 = 0; k2 < 2; k2++) {
                a[m][2 * k2] = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(Km + k2 * 480 + (30 * h + sl) * 8));
                // odd k-steps: half 0 reads 8 values, half 1 reads 6 (then zeros); both loads in every lane
                const u32x4 w0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(Km + 960 + k2 * 240 + sl * 8));
                const u32x3 w1 = __builtin_nontemporal_load(reinterpret_cast<const u32x3*>(Km + 1440 + k2 * 180 + sl * 6));
                const u32x4 w = h ? u32x4{w1[0], w1[1], w1[2], 0u} : w0;
                a[m][2 * k2 + 1] = __builtin_bit_cast(u16x8, w);
            }
        }
    }
};

// B fragments of one 60-DoF element: U rows at the element's DoFs (read from
// the wave's LDS copy of the position-ordered DoF list), RHS columns uc[q]
template <bool BF16, int NT> struct Rhs60;
template <int NT> struct Rhs60<false, NT> {
    float b[30][NT];  // lane half h, k-step ks: position 32 h + ks
    __device__ __forceinline__ void load(const float* __restrict__ U, uint32_t nrhs, const uint32_t* sd,
                                         const uint32_t (&uc)[NT])
    {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const uint32_t h = (threadIdx.x >> 5) & 1u;
        const u32x4* sp = reinterpret_cast<const u32x4*>(sd + 32 * h);
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const u32x4 d = sp[c];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int ks = 4 * c + r;
                if (ks >= 30) break;
#pragma unroll
                for (int q = 0; q < NT; q++) b[ks][q] = U[__umul24(d[r], nrhs) + uc[q]];
            }
        }
    }
};
template <int NT> struct Rhs60<true, NT> {
    float x[4][8][NT];  // lane half h, k-step ks, v: position 16 ks + 8 h + v
    __device__ __forceinline__ void load(const float* __restrict__ U, uint32_t nrhs, const uint32_t* sd,
                                         const uint32_t (&uc)[NT])
    {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const uint32_t h = (threadIdx.x >> 5) & 1u;
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const u32x4 d = *reinterpret_cast<const u32x4*>(sd + 16 * ks + 8 * h + 4 * c);
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int q = 0; q < NT; q++) x[ks][4 * c + r][q] = U[__umul24(d[r], nrhs) + uc[q]];
            }
    }
};

template <int NT>
__device__ __forceinline__ void mfma60(const Frag60<false>& A, const Rhs60<false, NT>& B, f32x16 (&acc)[2][NT])
{
#pragma unroll
    for (int ks = 0; ks < 30; ks++)
#pragma unroll
        for (int q = 0; q < NT; q++)
#pragma unroll
            for (int m = 0; m < 2; m++)
                acc[m][q] = __builtin_amdgcn_mfma_f32_32x32x2f32(A.get(m, ks), B.b[ks][q], acc[m][q], 0, 0, 0);
}

template <int NT>
__device__ __forceinline__ void mfma60(const Frag60<true>& A, const Rhs60<true, NT>& B, f32x16 (&acc)[2][NT])
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int ks = 0; ks < 4; ks++)
#pragma unroll
        for (int q = 0; q < NT; q++) {
            const u32x4 bw = {bf16_pair(B.x[ks][0][q], B.x[ks][1][q]), bf16_pair(B.x[ks][2][q], B.x[ks][3][q]),
                              bf16_pair(B.x[ks][4][q], B.x[ks][5][q]), bf16_pair(B.x[ks][6][q], B.x[ks][7][q])};
#pragma unroll
            for (int m = 0; m < 2; m++)
                acc[m][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, A.a[m][ks]),
                                                                     __builtin_bit_cast(bf16x8, bw), acc[m][q], 0, 0, 0);
        }
}

// add the rows of a 60-DoF element into V: lane half h, C register i of M-tile
// m is position 32 m + 8 (i/4) + 4 h + i%4 (padding positions 30, 31 skipped)
// (deterministic form: element row 30 m + 8 (i/4) + 4 h + i%4 of element e in ve)
template <int NT>
__device__ __forceinline__ void scatter60(const Seg& sg, uint32_t e, const uint32_t* sd, uint32_t col0,
                                          const f32x16 (&acc)[2][NT])
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const uint32_t lane = threadIdx.x & 63u, h = lane >> 5, col_l = lane & 31u;
    const uint32_t nrhs = sg.nrhs;
    float* __restrict__ V = sg.V;
    u32x4 d[2][4];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) d[m][j] = *reinterpret_cast<const u32x4*>(sd + 32 * m + 8 * j + 4 * h);
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool pad = j == 3 && r >= 2 && h;  // positions 30, 31 of the M-tile
#pragma unroll
                for (int q = 0; q < NT; q++) {
                    const uint32_t col = col0 + q * 32 + col_l;
                    if (col < nrhs && !pad) {
                        if (sg.ve) {
                            const uint64_t row = (uint64_t)e * 60 + 30 * m + 8 * j + 4 * h + r;
                            sg.ve[row * nrhs + col] = acc[m][q][4 * j + r];
                        } else {
                            unsafeAtomicAdd(V + __umul24(d[m][j][r], nrhs) + col, acc[m][q][4 * j + r]);
                        }
                    }
                }
            }
}

// one block-diagonal tile of G = 32/N small elements x one RHS tile
template <int N, bool BF16, int NT>
__device__ __forceinline__ void tile_small(const Seg& sg, uint32_t t, uint32_t ng)
{
    constexpr int G = tile_elems(N), ROWS = tile_rows(N), KD = kdim(N, BF16);
    static_assert(N <= 32 && KD <= 64, "small-element tile");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t half = lane >> 5, col_l = lane & 31u;
    const uint32_t mydof = lane < (uint32_t)KD ? sg.dofs[(uint64_t)t * KD + lane] : kNone;
    const uint32_t col0 = ng * 32u * NT;
    const uint32_t nrhs = sg.nrhs;
    const float* __restrict__ U = sg.U;
    uint32_t uc[NT];
    rhs_columns<NT>(sg, col0, uc);
    f32x16 acc[1][NT];
#pragma unroll
    for (int q = 0; q < NT; q++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[0][q][i] = 0.0f;
    const uint64_t e0 = (uint64_t)t * G;
    const int row = (int)col_l;
    const int g = row < ROWS ? row / N : 0, r = row < ROWS ? row % N : 0;
    const bool rv = row < ROWS && e0 + g < sg.n_elem;
    const uint64_t eg = rv ? e0 + g : e0;  // an existing element: loads stay in bounds
    const uint32_t dsafe = rl(mydof, 0);

    if constexpr (!BF16) {
        const float* K = static_cast<const float*>(sg.K) + eg * N * N;
        float a[KD / 2], b[KD / 2][NT];
#pragma unroll
        for (int ks = 0; ks < KD / 2; ks++) {
            const int k = 2 * ks + (int)half;
            const bool kv = rv && k < ROWS && k / N == g;
            const float x = K[kv ? (k % N) * N + r : 0];
            a[ks] = kv ? x : 0.0f;
            const uint32_t d0 = rl(mydof, 2 * ks), d1 = rl(mydof, 2 * ks + 1);
            const uint32_t d = half ? d1 : d0;
            const uint32_t off = (d != kNone ? d : dsafe) * nrhs;
#pragma unroll
            for (int q = 0; q < NT; q++) {
                const float y = U[off + uc[q]];
                b[ks][q] = d != kNone ? y : 0.0f;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int ks = 0; ks < KD / 2; ks++)
#pragma unroll
            for (int q = 0; q < NT; q++) acc[0][q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ks], b[ks][q], acc[0][q], 0, 0, 0);
    } else {
        constexpr int LD = bf16_ld(N);
        const unsigned short* K = static_cast<const unsigned short*>(sg.K) + eg * N * LD + (uint64_t)r * LD;
        u16x8 a[KD / 16];
        float y[KD / 16][8][NT];
#pragma unroll
        for (int ks = 0; ks < KD / 16; ks++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = 16 * ks + 8 * (int)half + j;
                const bool kv = rv && k < ROWS && k / N == g;
                const unsigned short x = K[kv ? k % N : 0];
                a[ks][j] = kv ? x : (unsigned short)0;
                const int k0 = 16 * ks + j;
                const uint32_t d0 = rl(mydof, k0), d1 = rl(mydof, k0 + 8);
                const uint32_t d = half ? d1 : d0;
                const uint32_t off = (d != kNone ? d : dsafe) * nrhs;
#pragma unroll
                for (int q = 0; q < NT; q++) {
                    const float v = U[off + uc[q]];
                    y[ks][j][q] = d != kNone ? v : 0.0f;
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int ks = 0; ks < KD / 16; ks++)
#pragma unroll
            for (int q = 0; q < NT; q++) {
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                const u32x4 bw = {bf16_pair(y[ks][0][q], y[ks][1][q]), bf16_pair(y[ks][2][q], y[ks][3][q]),
                                  bf16_pair(y[ks][4][q], y[ks][5][q]), bf16_pair(y[ks][6][q], y[ks][7][q])};
                acc[0][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a[ks]),
                                                                     __builtin_bit_cast(bf16x8, bw), acc[0][q], 0, 0, 0);
            }
    }
    scatter_rows<1, NT, ROWS>(sg, t, mydof, col0, acc);
}

template <int N, bool BF16, int NT>
__device__ __forceinline__ void slice_small(const Seg& sg, uint32_t lo, uint32_t hi, uint32_t stride)
{
    for (uint32_t w = lo; w < hi; w += stride) {
        const uint32_t t = w / sg.ngroups;
        tile_small<N, BF16, NT>(sg, t, w - t * sg.ngroups);
    }
}

// One item of a 60-DoF slice, software-pipelined: the DoF list and K
// fragments of the next item (tn) are loaded while the current item's U rows
// arrive and its MFMAs run.  The DoF list of the current item sits in the
// wave's LDS slot sd (64 entries); the next list is written there after the
// current item's rows have been added.  Load order: next DoFs, current U rows,
// next K, so that waiting for the U rows leaves the prefetch in flight.
template <bool BF16, int NT>
__device__ __forceinline__ void item60(const Seg& sg, uint32_t t, uint32_t ng, uint32_t tn, const Frag60<BF16>& A,
                                       Frag60<BF16>& An, uint32_t* sd)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t col0 = ng * 32u * NT;
    const uint32_t dn = sg.dofs[(uint64_t)tn * 64 + lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t uc[NT];
    rhs_columns<NT>(sg, col0, uc);
    Rhs60<BF16, NT> B;
    B.load(sg.U, sg.nrhs, sd, uc);
    An.load(sg.K, tn);
    __builtin_amdgcn_sched_barrier(0);
    f32x16 acc[2][NT];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int q = 0; q < NT; q++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[m][q][i] = 0.0f;
    mfma60<NT>(A, B, acc);
    scatter60<NT>(sg, t, sd, col0, acc);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    sd[lane] = dn;
}

// 60-DoF slice: items lo, lo + stride, ... < hi (wave-uniform).  Two register
// sets of K fragments alternate, so the prefetch needs no copy; the last item
// prefetches itself (a cache hit).
template <bool BF16, int NT>
__device__ __forceinline__ void slice60(const Seg& sg, uint32_t lo, uint32_t hi, uint32_t stride, uint32_t* sd)
{
    if (lo >= hi) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t ngs = sg.ngroups;
    Frag60<BF16> A0, A1;
    uint32_t t = lo / ngs;
    A0.load(sg.K, t);
    sd[lane] = sg.dofs[(uint64_t)t * 64 + lane];
    for (uint32_t w = lo; w < hi;) {
        t = w / ngs;
        uint32_t wn = w + stride < hi ? w + stride : w;
        item60<BF16, NT>(sg, t, w - t * ngs, wn / ngs, A0, A1, sd);
        w += stride;
        if (w >= hi) break;
        t = w / ngs;
        wn = w + stride < hi ? w + stride : w;
        item60<BF16, NT>(sg, t, w - t * ngs, wn / ngs, A1, A0, sd);
        w += stride;
    }
}

// Persistent: the grid is a multiple of 8 blocks; block b runs on XCD b % 8 and
// works through slice b % 8 of every segment, its waves striding over it.
template <bool BF16, int NT>
__global__ __launch_bounds__(256) void mrhs_persistent(SegTable tab)
{
    __shared__ __align__(16) uint32_t sdofs[4][64];  // per wave: DoF list of its current 60-DoF element
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t xcd = blockIdx.x % kXcd;
    const uint32_t nbx = gridDim.x / kXcd;
    const uint32_t stride = nbx * wpb;
    const uint32_t w0 = __builtin_amdgcn_readfirstlane((blockIdx.x / kXcd) * wpb + (threadIdx.x >> 6));
    for (uint32_t si = 0; si < tab.n_seg; si++) {
        const Seg& sg = tab.s[si];
        const uint64_t total = (uint64_t)sg.n_tiles * sg.ngroups;
        const uint64_t xlo = total * xcd / kXcd, xhi = total * (xcd + 1) / kXcd;
        uint32_t lo, hi, step;
        if (tab.chunked) {
            lo = (uint32_t)(xlo + (xhi - xlo) * w0 / stride);
            hi = (uint32_t)(xlo + (xhi - xlo) * (w0 + 1) / stride);
            step = 1;
        } else {
            lo = (uint32_t)xlo + w0;
            hi = (uint32_t)xhi;
            step = stride;
        }
        switch (sg.n) {
        case 8: slice_small<8, BF16, NT>(sg, lo, hi, step); break;
        case 10: slice_small<10, BF16, NT>(sg, lo, hi, step); break;
        case 20: slice_small<20, BF16, NT>(sg, lo, hi, step); break;
        case 60: slice60<BF16, NT>(sg, lo, hi, step, sdofs[threadIdx.x >> 6]); break;
        default: break;
        }
    }
}

__global__ void zero_f32(float* p, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// deterministic scatter, second pass: V[d][c] = sum over d's element rows (fixed
// order: ascending element, then row) of ve[row][c]
__global__ __launch_bounds__(256) void mrhs_gather_kernel(const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ pos, const float* __restrict__ ve,
                                                         float* __restrict__ V, uint32_t n_total, uint32_t nrhs)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n_total * nrhs) return;
    const uint32_t d = (uint32_t)(i / nrhs), c = (uint32_t)(i - (uint64_t)d * nrhs);
    const uint32_t b = off[d], cnt = off[d + 1] - b;
    float s = 0.0f;
    // chunks of 8 entries: index loads, then row loads, in flight together;
    // the sum keeps the list order (entries past the end are not added)
    for (uint32_t q0 = 0; q0 < cnt; q0 += 8) {
        uint32_t p[8];
        float x[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) p[j] = pos[b + min(q0 + j, cnt - 1)];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) x[j] = ve[(uint64_t)p[j] * nrhs + c];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++)
            if (q0 + j < cnt) s += x[j];
    }
    V[i] = s;
}

// the same sums, four columns per thread (nrhs a multiple of 4, 16-byte
// aligned rows): one 16-byte load per entry; each column's sum keeps the order
// of mrhs_gather_kernel, so the results are the same bits
__global__ __launch_bounds__(256) void mrhs_gather4_kernel(const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ pos,
                                                          const float4* __restrict__ ve, float4* __restrict__ V,
                                                          uint32_t n_total, uint32_t nq)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n_total * nq) return;
    const uint32_t d = (uint32_t)(i / nq), c = (uint32_t)(i - (uint64_t)d * nq);
    const uint32_t b = off[d], cnt = off[d + 1] - b;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t q0 = 0; q0 < cnt; q0 += 8) {
        uint32_t p[8];
        float4 x[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) p[j] = pos[b + min(q0 + j, cnt - 1)];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) x[j] = ve[(uint64_t)p[j] * nq + c];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++)
            if (q0 + j < cnt) {
                s.x += x[j].x;
                s.y += x[j].y;
                s.z += x[j].z;
                s.w += x[j].w;
            }
    }
    V[i] = s;
}

// element matrices for the MFMA kernels from the tiled FP64 storage:
// 60 DoFs -> compact MFMA order (pos60); smaller -> FP32 column-major n x n or
// BF16 row-major n x ld.  One thread per stored value.
// K64: the operator's FP64 storage -- column-packed tiles (G, P), packed
// symmetric tiles (sym_te values per tile) or near-symmetric records (nsym_rec
// values per record, nsym_nx exceptions, nsym_xidx)
__global__ void mrhs_layout_kernel(const double* __restrict__ K64, uint32_t N, uint32_t G, uint32_t P,
                                   uint32_t sym_te, uint64_t nsym_rec, uint32_t nsym_nx,
                                   const uint16_t* __restrict__ nsym_xidx, const uint32_t* __restrict__ slot_of,
                                   uint32_t n_elem, int bf16, uint32_t ld, void* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per = (bf16 && N != 60) ? (uint64_t)N * ld : (uint64_t)N * N;
    if (i >= per * n_elem) return;
    const uint64_t e = i / per;
    const uint32_t w = (uint32_t)(i - e * per);
    uint32_t r, c;
    uint64_t dst = i;
    if (N == 60) {
        r = w / N;
        c = w % N;
        dst = e * 3600 + pos60(bf16 != 0, r, c);
    } else if (bf16) {
        r = w / ld;
        c = w % ld;
    } else {
        c = w / N;
        r = w % N;
    }
    float val = 0.0f;
    if (c < N) {
        const uint64_t slot = slot_of[e];
        const uint64_t t = slot / G;
        const uint32_t g = (uint32_t)(slot - t * G), GN = G * N, NP = N / P;
        if (nsym_rec) {
            val = (float)nsym_value(K64 + slot * nsym_rec, N, nsym_nx, nsym_xidx, r, c);
        } else if (sym_te) {  // packed symmetric tiles: K[r][c] = K[min][max] of the upper triangle
            const uint32_t a = r < c ? r : c, b = r < c ? c : r;
            val = (float)K64[t * sym_te + g * (N * (N + 1) / 2) + a * N - (a * (a - 1)) / 2 + (b - a)];
        } else {
            val = (float)K64[((t * NP + c / P) * GN + g * N + r) * P + c % P];
        }
    }
    if (bf16) static_cast<__bf16*>(out)[dst] = (__bf16)val;
    else static_cast<float*>(out)[dst] = val;
}

}  // namespace
}  // namespace amx

struct amx_mrhs {
    int device = 0;
    bool bf16 = false;
    int rhs_tile = 32, wpb = 4;  // NT * 32 columns per wave, waves per block
    int chunked = 1;             // work split inside an XCD slice: 1 contiguous chunk per wave, 0 interleaved
    int scatter = AMX_SCATTER_ATOMIC;
    DevBuf ve;                   // deterministic scatter: element rows of all parts x nrhs
    struct Part {
        uint32_t n, n_elem, n_tiles, n_total;
        uint64_t ve_row0;        // first element row of this part in ve
        DevBuf K, dofs, off, pos;  // off/pos: DoF -> element rows (e*n + r), ascending
    };
    std::vector<Part> parts;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int cu_count = 256;
    int occupancy[256] = {};  // resident blocks per CU, by kernel variant
    ~amx_mrhs()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace amx {
namespace {

bool supported_n(uint32_t n) { return n == 8 || n == 10 || n == 20 || n == 60; }

template <bool B, int T>
void launch_persistent(amx_mrhs& m, const SegTable& tab, uint64_t max_slice, hipStream_t s)
{
    const int key = (B ? 100 : 0) + T * 10 + m.wpb;
    int per_cu = m.occupancy[key];
    if (per_cu <= 0) {
        AMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mrhs_persistent<B, T>, 64 * m.wpb, 0));
        per_cu = std::max(per_cu, 1);
        m.occupancy[key] = per_cu;
    }
    // blocks per XCD: enough for the largest slice, at most what is resident at once
    const uint64_t want = (max_slice + m.wpb - 1) / m.wpb;
    const uint64_t resident = (uint64_t)per_cu * std::max(m.cu_count / (int)kXcd, 1);
    const uint32_t nbx = (uint32_t)std::max<uint64_t>(1, std::min(want, resident));
    mrhs_persistent<B, T><<<dim3(nbx * kXcd), dim3(64 * m.wpb), 0, s>>>(tab);
}

// element-row buffer of the deterministic scatter (allocated outside launches,
// so a captured graph never allocates)
void prepare(amx_mrhs& m, int nrhs)
{
    if (m.scatter != AMX_SCATTER_DETERMINISTIC) return;
    uint64_t rows = 0;
    for (auto& p : m.parts) rows += (uint64_t)p.n_elem * p.n;
    const uint64_t bytes = std::max<uint64_t>(rows * (uint64_t)nrhs * 4, 4);
    if (m.ve.bytes < bytes) m.ve.alloc(bytes);
}

void launch(amx_mrhs& m, int nrhs, const float* const* U, float* const* V, hipStream_t s)
{
    SegTable tab{};
    const bool det = m.scatter == AMX_SCATTER_DETERMINISTIC;
    if (det) {
        uint64_t rows = 0;
        for (auto& p : m.parts) rows += (uint64_t)p.n_elem * p.n;
        AMX_REQUIRE(m.ve.bytes >= rows * (uint64_t)nrhs * 4, AMX_INVALID_STATE, "element-row buffer not prepared");
    }
    AMX_REQUIRE(m.parts.size() <= (size_t)kMaxSeg, AMX_NOT_IMPLEMENTED, "too many element groups in one launch");
    const int NT = m.rhs_tile / 32;
    const uint32_t ngroups = (uint32_t)((nrhs + m.rhs_tile - 1) / m.rhs_tile);
    for (auto& p : m.parts)
        AMX_REQUIRE((uint64_t)p.n_total * (uint64_t)nrhs < (1ull << 32) && nrhs < (1 << 24), AMX_NOT_IMPLEMENTED,
                    "multi-RHS products index U and V with 32 bits: n_total * nrhs must stay below 2^32");
    uint64_t max_slice = 0;
    for (size_t i = 0; i < m.parts.size(); i++) {
        auto& p = m.parts[i];
        Seg& sg = tab.s[i];
        sg.K = p.K.p;
        sg.dofs = p.dofs.as<uint32_t>();
        sg.U = U[i];
        sg.V = V[i];
        sg.n = p.n;
        sg.n_elem = p.n_elem;
        sg.n_tiles = p.n_tiles;
        sg.ngroups = ngroups;
        sg.nrhs = (uint32_t)nrhs;
        sg.dstride = (uint32_t)dof_stride((int)p.n, m.bf16);
        sg.ve = det ? m.ve.as<float>() + p.ve_row0 * (uint64_t)nrhs : nullptr;
        max_slice = std::max<uint64_t>(max_slice, ((uint64_t)p.n_tiles * ngroups + kXcd - 1) / kXcd);
        const uint64_t len = (uint64_t)p.n_total * nrhs;
        if (len && !det) zero_f32<<<(uint32_t)((len + 255) / 256), 256, 0, s>>>(V[i], len);
    }
    tab.n_seg = (uint32_t)m.parts.size();
    tab.chunked = (uint32_t)m.chunked;
    if (max_slice == 0) return;
    if (m.bf16) {
        if (NT == 1) launch_persistent<true, 1>(m, tab, max_slice, s);
        else if (NT == 2) launch_persistent<true, 2>(m, tab, max_slice, s);
        else launch_persistent<true, 4>(m, tab, max_slice, s);
    } else {
        if (NT == 1) launch_persistent<false, 1>(m, tab, max_slice, s);
        else if (NT == 2) launch_persistent<false, 2>(m, tab, max_slice, s);
        else launch_persistent<false, 4>(m, tab, max_slice, s);
    }
    AMX_HIP(hipGetLastError());
    if (det) {
        for (size_t i = 0; i < m.parts.size(); i++) {
            auto& p = m.parts[i];
            const uint64_t len = (uint64_t)p.n_total * nrhs;
            if (!len) continue;
            const float* ve = m.ve.as<float>() + p.ve_row0 * (uint64_t)nrhs;
            if (nrhs % 4 == 0 && reinterpret_cast<uintptr_t>(V[i]) % 16 == 0 &&
                reinterpret_cast<uintptr_t>(ve) % 16 == 0) {
                mrhs_gather4_kernel<<<(uint32_t)((len / 4 + 255) / 256), 256, 0, s>>>(
                    p.off.as<uint32_t>(), p.pos.as<uint32_t>(), reinterpret_cast<const float4*>(ve),
                    reinterpret_cast<float4*>(V[i]), p.n_total, (uint32_t)(nrhs / 4));
            } else {
                mrhs_gather_kernel<<<(uint32_t)((len + 255) / 256), 256, 0, s>>>(
                    p.off.as<uint32_t>(), p.pos.as<uint32_t>(), ve, V[i], p.n_total, (uint32_t)nrhs);
            }
        }
        AMX_HIP(hipGetLastError());
    }
}

}  // namespace
}  // namespace amx

extern "C" {

int amx_mrhs_create(amx_op* const* ops, int n_ops, int precision, amx_mrhs** out)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(ops && out && n_ops > 0, AMX_INVALID_ARGUMENT, "bad argument");
        AMX_REQUIRE(precision == AMX_PREC_FP32 || precision == AMX_PREC_BF16, AMX_INVALID_ARGUMENT,
                    "multi-RHS products run in FP32 or BF16");
        std::unique_ptr<amx_mrhs> m(new amx_mrhs());
        m->bf16 = precision == AMX_PREC_BF16;
        m->device = ops[0]->op.device;
        amx::DeviceGuard dg(m->device);
        hipDeviceProp_t prop;
        AMX_HIP(hipGetDeviceProperties(&prop, m->device));
        m->cu_count = prop.multiProcessorCount;
        for (int i = 0; i < n_ops; i++) {
            amx::Op& op = ops[i]->op;
            AMX_REQUIRE(op.finalized, AMX_INVALID_STATE, "operator not finalized");
            AMX_REQUIRE(!op.is_free(), AMX_NOT_IMPLEMENTED, "multi-RHS products need stored element matrices");
            AMX_REQUIRE(op.device == m->device, AMX_INVALID_ARGUMENT, "operators on different devices");
            AMX_REQUIRE(op.groups.size() == 1, AMX_NOT_IMPLEMENTED, "one element type per operator");
            const amx::Group& g = op.groups[0];
            AMX_REQUIRE(amx::supported_n(g.n), AMX_NOT_IMPLEMENTED,
                        "multi-RHS tiles exist for 8, 10, 20 and 60 DoFs per element");
            amx_mrhs::Part p;
            p.n = g.n;
            p.n_total = op.n_total;
            // elements in slot order (padding skipped)
            std::vector<uint32_t> slot_of;
            for (uint32_t sl = 0; sl < g.n_slots; sl++)
                if (g.slot_elem[sl] != amx::kPadSlot) slot_of.push_back(sl);
            p.n_elem = (uint32_t)slot_of.size();
            const int G = amx::tile_elems((int)g.n), KD = amx::kdim((int)g.n, m->bf16);
            const int DS = amx::dof_stride((int)g.n, m->bf16);
            p.n_tiles = (p.n_elem + G - 1) / G;
            AMX_REQUIRE(op.n_total < (1u << 24), AMX_NOT_IMPLEMENTED, "multi-RHS products take up to 2^24 - 1 DoFs");
            std::vector<uint32_t> dofs((size_t)p.n_tiles * DS, amx::kNone);
            for (uint32_t t = 0; t < p.n_tiles; t++) {
                if (g.n == 60) {
                    // position order (p = 32 h + s -> DoF 30 h + s; padding repeats DoF 0)
                    const uint32_t* em = &op.local_map[op.eoff[g.elems[g.slot_elem[slot_of[t]]]]];
                    for (int q = 0; q < 64; q++) dofs[(size_t)t * 64 + q] = (q % 32 < 30) ? em[30 * (q / 32) + q % 32] : em[0];
                    continue;
                }
                for (int k = 0; k < KD && k < amx::tile_rows((int)g.n); k++) {
                    const uint32_t j = t * G + k / g.n;
                    if (j >= p.n_elem) continue;
                    const uint32_t e = g.elems[g.slot_elem[slot_of[j]]];
                    dofs[(size_t)t * DS + k] = op.local_map[op.eoff[e] + k % g.n];
                }
            }
            // DoF -> element rows (e * n + r), ascending: the deterministic scatter's lists
            {
                p.ve_row0 = 0;
                for (auto& q : m->parts) p.ve_row0 += (uint64_t)q.n_elem * q.n;
                std::vector<uint32_t> off(op.n_total + 1, 0), pos((size_t)p.n_elem * g.n);
                for (uint32_t j = 0; j < p.n_elem; j++) {
                    const uint32_t* em = &op.local_map[op.eoff[g.elems[g.slot_elem[slot_of[j]]]]];
                    for (uint32_t r = 0; r < g.n; r++) off[em[r] + 1]++;
                }
                for (uint32_t d = 0; d < op.n_total; d++) off[d + 1] += off[d];
                std::vector<uint32_t> fill(off.begin(), off.end() - 1);
                for (uint32_t j = 0; j < p.n_elem; j++) {
                    const uint32_t* em = &op.local_map[op.eoff[g.elems[g.slot_elem[slot_of[j]]]]];
                    for (uint32_t r = 0; r < g.n; r++) pos[fill[em[r]]++] = j * g.n + r;
                }
                AMX_REQUIRE((uint64_t)p.n_elem * g.n < (1ull << 32), AMX_NOT_IMPLEMENTED, "too many element rows");
                p.off.alloc(off.size() * 4);
                AMX_HIP(hipMemcpy(p.off.p, off.data(), off.size() * 4, hipMemcpyHostToDevice));
                p.pos.alloc(std::max<size_t>(pos.size(), 1) * 4);
                AMX_HIP(hipMemcpy(p.pos.p, pos.data(), pos.size() * 4, hipMemcpyHostToDevice));
            }
            DevBuf dslot;
            dslot.alloc(std::max<size_t>(slot_of.size(), 1) * 4);
            AMX_HIP(hipMemcpy(dslot.p, slot_of.data(), slot_of.size() * 4, hipMemcpyHostToDevice));
            p.dofs.alloc(std::max<size_t>(dofs.size(), 1) * 4);
            AMX_HIP(hipMemcpy(p.dofs.p, dofs.data(), dofs.size() * 4, hipMemcpyHostToDevice));
            const uint32_t ld = (m->bf16 && g.n != 60) ? (uint32_t)amx::bf16_ld((int)g.n) : g.n;
            const uint64_t total = (uint64_t)p.n_elem * g.n * ld;
            p.K.alloc(std::max<uint64_t>(total, 8) * (m->bf16 ? 2 : 4));
            if (total)
                amx::mrhs_layout_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, op.stream>>>(
                    g.K[0].as<double>(), g.n, g.L[0].G, g.L[0].P,
                    g.L[0].sym ? (uint32_t)amx::tile_values(g.L[0], amx::KType::F64) : 0u,
                    g.L[0].nsym ? amx::nsym_record_bytes(g.n, g.L[0].nx) / 8u : 0u, g.L[0].nx,
                    g.nsym_xidx.as<uint16_t>(), dslot.as<uint32_t>(),
                    p.n_elem, m->bf16 ? 1 : 0, ld,
                    p.K.p);
            AMX_HIP(hipGetLastError());
            AMX_HIP(hipStreamSynchronize(op.stream));
            m->parts.push_back(std::move(p));
        }
        *out = m.release();
    });
}

int amx_mrhs_destroy(amx_mrhs* m)
{
    if (m) {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(m->device);
        delete m;
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    return AMX_SUCCESS;
}

int amx_mrhs_set_tiles(amx_mrhs* m, int rhs_tile, int waves_per_block, int work_split)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m, AMX_INVALID_ARGUMENT, "NULL handle");
        AMX_REQUIRE(rhs_tile == 32 || rhs_tile == 64 || rhs_tile == 128, AMX_INVALID_ARGUMENT, "rhs_tile: 32, 64 or 128");
        AMX_REQUIRE(waves_per_block == 1 || waves_per_block == 2 || waves_per_block == 4, AMX_INVALID_ARGUMENT,
                    "waves_per_block: 1, 2 or 4");
        AMX_REQUIRE(work_split == 0 || work_split == 1, AMX_INVALID_ARGUMENT, "work_split: 0 or 1");
        m->rhs_tile = rhs_tile;
        m->wpb = waves_per_block;
        m->chunked = work_split;
    });
}

int amx_mrhs_set_scatter(amx_mrhs* m, int scatter)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m, AMX_INVALID_ARGUMENT, "NULL handle");
        AMX_REQUIRE(scatter == AMX_SCATTER_DETERMINISTIC || scatter == AMX_SCATTER_ATOMIC, AMX_INVALID_ARGUMENT,
                    "scatter: AMX_SCATTER_DETERMINISTIC or AMX_SCATTER_ATOMIC");
        m->scatter = scatter;
    });
}

int amx_mrhs_get_tiles(const amx_mrhs* m, int* rhs_tile, int* waves_per_block, int* work_split)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m, AMX_INVALID_ARGUMENT, "NULL handle");
        if (rhs_tile) *rhs_tile = m->rhs_tile;
        if (waves_per_block) *waves_per_block = m->wpb;
        if (work_split) *work_split = m->chunked;
    });
}

int amx_mrhs_apply_dev(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m && U && V && nrhs > 0, AMX_INVALID_ARGUMENT, "bad argument");
        amx::DeviceGuard dg(m->device);
        amx::prepare(*m, nrhs);
        amx::launch(*m, nrhs, U, V, static_cast<hipStream_t>(stream));
    });
}

int amx_mrhs_autotune(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, int reps, double* results,
                      void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m && U && V && nrhs > 0 && reps > 0, AMX_INVALID_ARGUMENT, "bad argument");
        amx::DeviceGuard dg(m->device);
        hipStream_t s = static_cast<hipStream_t>(stream);
        amx::prepare(*m, nrhs);
        hipEvent_t e0, e1;
        AMX_HIP(hipEventCreate(&e0));
        AMX_HIP(hipEventCreate(&e1));
        double best = 1e30;
        int brt = 32, bw = 4, bc = 1, idx = 0;
        for (int rt : {32, 64, 128}) {
            for (int w : {1, 2, 4}) {
                for (int c : {0, 1}) {
                    double ms = 1e30;
                    if (rt <= std::max(32, ((nrhs + 31) / 32) * 32)) {
                        m->rhs_tile = rt;
                        m->wpb = w;
                        m->chunked = c;
                        amx::launch(*m, nrhs, U, V, s);  // warm-up
                        AMX_HIP(hipEventRecord(e0, s));
                        for (int r = 0; r < reps; r++) amx::launch(*m, nrhs, U, V, s);
                        AMX_HIP(hipEventRecord(e1, s));
                        AMX_HIP(hipEventSynchronize(e1));
                        float f = 0;
                        AMX_HIP(hipEventElapsedTime(&f, e0, e1));
                        ms = f / reps;
                        if (ms < best) {
                            best = ms;
                            brt = rt;
                            bw = w;
                            bc = c;
                        }
                    }
                    if (results) results[idx] = ms;  // order: rhs_tile 32/64/128 x waves 1/2/4 x split 0/1
                    idx++;
                }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        m->rhs_tile = brt;
        m->wpb = bw;
        m->chunked = bc;
    });
}

int amx_mrhs_graph_capture(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m && U && V && nrhs > 0, AMX_INVALID_ARGUMENT, "bad argument");
        AMX_REQUIRE(stream != nullptr, AMX_INVALID_ARGUMENT, "capture needs a non-default stream");
        amx::DeviceGuard dg(m->device);
        hipStream_t s = static_cast<hipStream_t>(stream);
        amx::prepare(*m, nrhs);
        if (m->exec) { (void)hipGraphExecDestroy(m->exec); m->exec = nullptr; }
        if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
        AMX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            amx::launch(*m, nrhs, U, V, s);
        } catch (...) {
            hipGraph_t g;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        AMX_HIP(hipStreamEndCapture(s, &m->graph));
        AMX_HIP(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    });
}

int amx_mrhs_graph_launch(amx_mrhs* m, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m && m->exec, AMX_INVALID_STATE, "no captured graph");
        amx::DeviceGuard dg(m->device);
        AMX_HIP(hipGraphLaunch(m->exec, static_cast<hipStream_t>(stream)));
    });
}

}  // extern "C"
