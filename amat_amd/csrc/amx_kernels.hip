// amx_kernels.hip -- gfx950 kernels of the assembly-free EbE operator.
//
// Hot path (per matvec):
//   ebe_tile_kernel  one 64-lane wave per tile of G = 64/n elements: lane
//                    (g, r) gathers u[map] (coalesced index read, cached u
//                    read), streams its row of K_e as P-column packs with
//                    non-temporal 16-byte loads, and forms ve = K_e u_e with
//                    the column values broadcast across the wave (readlane for
//                    G = 1, ds_bpermute otherwise); whole matrices are never
//                    staged in LDS, packed upper triangles of multi-element
//                    tiles are (each lane then reads its row from the wave's copy).
//   ebe_nsym_kernel  60-DoF elements held as near-symmetric FP64 records
//   psx_kernel       (FP32 / BF16: packed records with exceptions): one wave per
//                    element, the record staged in the wave's LDS region, every
//                    entry rebuilt bit for bit (amx_nsym.hpp).
//   gather_kernel    v[d] = sum of the element rows that touch d, in the
//                    reference's element order (node -> (element, row)
//                    transposed lists), a block's list range staged in LDS,
//                    with the BC rows fused.
// The *_batched kernels run the same arithmetic for many operators (the jobs
// of the solve server) in one launch.
// Blocks are remapped so that each of the 8 XCDs walks a contiguous range of
// tiles / DoFs: its L2 then holds one slab of u / ve instead of all of it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "amx_internal.hpp"
#include "amx_kernels.hpp"
#include "amx_nsym.hpp"

namespace amx {

namespace {

constexpr uint32_t kFlagBit = 0x80000000u;
constexpr uint32_t kIdxMask = 0x7fffffffu;
constexpr uint32_t kPadSlot = 0xffffffffu;

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Return a logical block index such that XCD x gets the contiguous
// logical range [x*q + min(x, r), ... + q + (x < r)), q = nb / 8, r = nb % 8.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    const uint32_t x = b & 7u, i = b >> 3;
    const uint32_t q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + i;
}

template <typename KT> struct KSize;
template <> struct KSize<double> { static constexpr int v = 8; };
template <> struct KSize<float> { static constexpr int v = 4; };
template <> struct KSize<uint16_t> { static constexpr int v = 2; };

// columns per lane load: the largest P in {8,4,2,1} dividing n with
// P * sizeof(KT) <= 16 bytes
__host__ __device__ constexpr int pack_of(int n, int ksize)
{
    return (n % 8 == 0 && 8 * ksize <= 16) ? 8
           : (n % 4 == 0 && 4 * ksize <= 16) ? 4
           : (n % 2 == 0 && 2 * ksize <= 16) ? 2
                                              : 1;
}

// packed symmetric tiles: values of one element (upper triangle), index of
// (r, c) in it, values per stored tile (G elements, padded to 16 bytes)
__host__ __device__ constexpr uint32_t sym_ns(uint32_t n) { return n * (n + 1) / 2; }
__host__ __device__ __forceinline__ uint32_t sym_idx(uint32_t n, uint32_t r, uint32_t c)
{
    const uint32_t a = r < c ? r : c, b = r < c ? c : r;
    return a * n - (a * (a - 1)) / 2 + (b - a);
}
__host__ __device__ constexpr uint32_t sym_tile_values(uint32_t n, uint32_t ksize)
{
    return (((64 / n) * sym_ns(n) * ksize + 15) / 16 * 16) / ksize;
}

__device__ __forceinline__ float bf16_to_f32(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

__device__ __forceinline__ uint16_t f64_to_bf16(double x)
{
    // round to FP32 then to BF16, both to nearest even
    const uint32_t b = __float_as_uint((float)x);
    const uint32_t bias = 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)((b + bias) >> 16);
}

template <typename KT> __device__ __forceinline__ KT from_f64(double x);
template <> __device__ __forceinline__ double from_f64<double>(double x) { return x; }
template <> __device__ __forceinline__ float from_f64<float>(double x) { return (float)x; }
template <> __device__ __forceinline__ uint16_t from_f64<uint16_t>(double x) { return f64_to_bf16(x); }

// P consecutive K values of one lane, read with one non-temporal load.
template <typename KT, int P> struct KPack;

template <int P> struct KPack<double, P> {
    typedef double vec __attribute__((ext_vector_type(P)));
    vec v;
    __device__ __forceinline__ void load(const double* p) { v = __builtin_nontemporal_load((const vec*)p); }
    __device__ __forceinline__ double get(int i) const { return v[i]; }
};
template <> struct KPack<double, 1> {
    double v;
    __device__ __forceinline__ void load(const double* p) { v = __builtin_nontemporal_load(p); }
    __device__ __forceinline__ double get(int) const { return v; }
};
template <int P> struct KPack<float, P> {
    typedef float vec __attribute__((ext_vector_type(P)));
    vec v;
    __device__ __forceinline__ void load(const float* p) { v = __builtin_nontemporal_load((const vec*)p); }
    __device__ __forceinline__ float get(int i) const { return v[i]; }
};
template <> struct KPack<float, 1> {
    float v;
    __device__ __forceinline__ void load(const float* p) { v = __builtin_nontemporal_load(p); }
    __device__ __forceinline__ float get(int) const { return v; }
};
// BF16: P values packed two per 32-bit word
template <int P> struct KPack<uint16_t, P> {
    static constexpr int W = P / 2;
    typedef uint32_t vec __attribute__((ext_vector_type(W)));
    vec v;
    __device__ __forceinline__ void load(const uint16_t* p) { v = __builtin_nontemporal_load((const vec*)p); }
    __device__ __forceinline__ float get(int i) const
    {
        const uint32_t w = v[i >> 1];
        return bf16_to_f32((i & 1) ? (w >> 16) : (w & 0xffffu));
    }
};
template <> struct KPack<uint16_t, 2> {
    uint32_t v;
    __device__ __forceinline__ void load(const uint16_t* p) { v = __builtin_nontemporal_load((const uint32_t*)p); }
    __device__ __forceinline__ float get(int i) const { return bf16_to_f32(i ? (v >> 16) : (v & 0xffffu)); }
};
template <> struct KPack<uint16_t, 1> {
    uint16_t v;
    __device__ __forceinline__ void load(const uint16_t* p) { v = __builtin_nontemporal_load(p); }
    __device__ __forceinline__ float get(int) const { return bf16_to_f32(v); }
};

// acc + k*u.  FP64: product and sum rounded separately (no contraction), the
// arithmetic of the reference's scalar GEMV (include/aMatFree.hpp:1994-1999,
// built without FMA), so the deterministic FP64 path reproduces it bit for bit.
__device__ __forceinline__ double madd(double acc, double k, double u)
{
#pragma clang fp contract(off)
    const double p = k * u;
    return acc + p;
}
__device__ __forceinline__ float madd(float acc, float k, float u) { return __builtin_fmaf(k, u, acc); }

// one stored value in the compute type
__device__ __forceinline__ double kval(double x) { return x; }
__device__ __forceinline__ float kval(float x) { return x; }
__device__ __forceinline__ float kval(uint16_t x) { return bf16_to_f32(x); }

// LDS written by this wave is visible to all its lanes
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of lane `src` (uniform, compile-time after unrolling) to all lanes
__device__ __forceinline__ double readlane(double x, int src)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane(float x, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}

template <int G, int N, typename AT> __device__ __forceinline__ AT bcast(AT x, uint32_t g, int c)
{
    if constexpr (G == 1) {
        return readlane(x, c);
    } else {
        return __shfl(x, (int)(g * N + c), 64);
    }
}

template <typename VT> __device__ __forceinline__ void atomic_add(VT* p, VT x) { unsafeAtomicAdd(p, x); }

// ---------------------------------------------------------------- EbE tile
// One wave, tile t: G elements x N rows (lane = g*N + r).
// Element row (slot, r) of the deterministic scatter is ve[ve_base + r*rs + slot*ss]:
// rs = 1, ss = N (rows of an element together) or rs = n_slots, ss = 1 (row r of
// consecutive elements together: what a wave of consecutive DoFs reads in the gather).
//
// ebe_tile_acc: lane (g, r) of tile t forms row r of K_e u_e (m: its map entry).
//
// SYM (packed symmetric tiles, G > 1): the wave copies its tile's upper
// triangles (16-byte loads, all lanes) into its LDS region ksh, then lane
// (g, r) reads K[r][c] = K[min][max] there, c ascending: the products and sums
// of the full layout, in the same order.
template <int N, typename KT> constexpr uint32_t sym_region_values()
{
    return (64 / N > 1) ? sym_tile_values(N, KSize<KT>::v) : 1u;
}

template <int N, typename KT, typename VT, bool MASKED, bool SYM = false>
__device__ __forceinline__ VT ebe_tile_acc(const KT* __restrict__ K, const uint32_t* __restrict__ map,
                                           const VT* __restrict__ u, uint32_t t, const float* __restrict__ kdiag,
                                           uint32_t& m, KT* ksh = nullptr)
{
    using AT = VT;
    constexpr int G = 64 / N;
    constexpr int GN = G * N;
    constexpr int P = pack_of(N, KSize<KT>::v);
    constexpr int NP = N / P;
    const uint32_t lane = threadIdx.x & 63u;
    // lanes past G*N mirror the last active lane (same addresses, no traffic)
    const uint32_t lc = lane < (uint32_t)GN ? lane : (uint32_t)GN - 1u;
    const uint32_t g = lc / N;
    const uint64_t row = (uint64_t)t * GN + lc;  // = slot * N + r
    AT acc = AT(0);
    if constexpr (SYM) {
        static_assert(G > 1, "packed symmetric tiles hold several elements");
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        constexpr uint32_t TE = sym_tile_values(N, KSize<KT>::v), NS = sym_ns(N);
        constexpr uint32_t CH = TE * KSize<KT>::v / 16, NCH = (CH + 63) / 64;
        const u32x4* src = reinterpret_cast<const u32x4*>(K + (uint64_t)t * TE);
        u32x4 ch[NCH];
#pragma unroll
        for (uint32_t i = 0; i < NCH; i++)
            if (lane + 64 * i < CH) ch[i] = __builtin_nontemporal_load(src + lane + 64 * i);
        m = map[row];
        VT uv = u[m & kIdxMask];
        if (MASKED && (m & kFlagBit)) uv = VT(0);
        const AT ue = uv;
        u32x4* dst = reinterpret_cast<u32x4*>(ksh);
#pragma unroll
        for (uint32_t i = 0; i < NCH; i++)
            if (lane + 64 * i < CH) dst[lane + 64 * i] = ch[i];
        wave_lds_sync();
        const uint32_t r = lc - g * N;
        const KT* ke = ksh + g * NS;
#pragma unroll
        for (int c = 0; c < N; ++c) acc = madd(acc, (AT)kval(ke[sym_idx(N, r, c)]), bcast<G, N>(ue, g, c));
        if constexpr (std::is_same<KT, uint16_t>::value) acc = madd(acc, (AT)kdiag[row], ue);
        return acc;
    }
    m = map[row];
    VT uv = u[m & kIdxMask];
    if (MASKED && (m & kFlagBit)) uv = VT(0);
    const AT ue = uv;
    const KT* kp = K + ((uint64_t)t * NP * GN + lc) * P;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        KPack<KT, P> k;
        k.load(kp + (uint64_t)j * GN * P);
#pragma unroll
        for (int p = 0; p < P; ++p) acc = madd(acc, (AT)k.get(p), bcast<G, N>(ue, g, j * P + p));
    }
    // BF16 storage keeps the diagonal in FP32, corrected for the row's
    // off-diagonal rounding (row sums preserved); the BF16 diagonal slot is 0
    if constexpr (std::is_same<KT, uint16_t>::value) acc = madd(acc, (AT)kdiag[row], ue);
    return acc;
}

// TPW consecutive packed symmetric tiles t0.. of one wave (ntw of them real,
// uniform per wave): every tile's 16-byte loads, map entries and u values are
// issued first (a wave-scope fence per tile would keep the next tile's loads
// behind it); then tile by tile the wave stages its triangles in LDS and lane
// (g, r) forms row r of K_e u_e as ebe_tile_acc does.  ksh: the wave's TPW
// regions.
template <int N, typename KT, typename VT, bool MASKED, int TPW>
__device__ __forceinline__ void ebe_tiles_sym(const KT* __restrict__ K, const uint32_t* __restrict__ map,
                                              const VT* __restrict__ u, uint32_t t0, uint32_t ntw,
                                              const float* __restrict__ kdiag, VT (&acc)[TPW], KT* ksh)
{
    constexpr int G = 64 / N;
    constexpr int GN = G * N;
    static_assert(G > 1, "packed symmetric tiles hold several elements");
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr uint32_t TE = sym_tile_values(N, KSize<KT>::v), NS = sym_ns(N);
    constexpr uint32_t CH = TE * KSize<KT>::v / 16, NCH = (CH + 63) / 64;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t lc = lane < (uint32_t)GN ? lane : (uint32_t)GN - 1u;
    const uint32_t g = lc / N, r = lc - g * N;
    constexpr bool BF = std::is_same<KT, uint16_t>::value;
    u32x4 ch[TPW][NCH];
    uint32_t m[TPW];
    VT ue[TPW];
    float kd[TPW];
#pragma unroll
    for (int i = 0; i < TPW; i++) {
        if ((uint32_t)i < ntw) {
            const u32x4* src = reinterpret_cast<const u32x4*>(K + (uint64_t)(t0 + i) * TE);
#pragma unroll
            for (uint32_t c = 0; c < NCH; c++)
                if (lane + 64 * c < CH) ch[i][c] = __builtin_nontemporal_load(src + lane + 64 * c);
            m[i] = map[(uint64_t)(t0 + i) * GN + lc];
            if constexpr (BF) kd[i] = kdiag[(uint64_t)(t0 + i) * GN + lc];
        }
    }
#pragma unroll
    for (int i = 0; i < TPW; i++) {
        if ((uint32_t)i < ntw) {
            VT uv = u[m[i] & kIdxMask];
            if (MASKED && (m[i] & kFlagBit)) uv = VT(0);
            ue[i] = uv;
        }
    }
    // tile i is staged and computed while the loads of the later tiles are in flight
#pragma unroll
    for (int i = 0; i < TPW; i++) {
        if ((uint32_t)i < ntw) {
            u32x4* dst = reinterpret_cast<u32x4*>(ksh + i * TE);
#pragma unroll
            for (uint32_t c = 0; c < NCH; c++)
                if (lane + 64 * c < CH) dst[lane + 64 * c] = ch[i][c];
            wave_lds_sync();
            const KT* ke = ksh + i * TE + g * NS;
            VT a = VT(0);
#pragma unroll
            for (int c = 0; c < N; ++c) a = madd(a, (VT)kval(ke[sym_idx(N, r, c)]), bcast<G, N>(ue[i], g, c));
            if constexpr (BF) a = madd(a, (VT)kd[i], ue[i]);
            acc[i] = a;
        }
    }
}

// one tile, stored by its own lanes
template <int N, typename KT, typename VT, bool MASKED, bool ATOMIC, bool SYM = false>
__device__ __forceinline__ void ebe_tile_body(const KT* __restrict__ K, const uint32_t* __restrict__ map,
                                              const VT* __restrict__ u, VT* __restrict__ out, uint64_t ve_base,
                                              uint32_t rs, uint32_t ss, uint32_t t, const float* __restrict__ kdiag,
                                              KT* ksh = nullptr)
{
    constexpr int G = 64 / N;
    constexpr int GN = G * N;
    uint32_t m;
    const VT acc = ebe_tile_acc<N, KT, VT, MASKED, SYM>(K, map, u, t, kdiag, m, ksh);
    const uint32_t lane = threadIdx.x & 63u;
    if (lane < (uint32_t)GN) {
        if constexpr (ATOMIC) {
            atomic_add(&out[m & kIdxMask], acc);
        } else {
            const uint32_t g = lane / N;
            const uint64_t slot = (uint64_t)t * G + g, r = lane - g * N;
            out[ve_base + r * rs + slot * ss] = acc;
        }
    }
}

// A block's 4 waves take 4 consecutive tiles (TPW consecutive tiles per wave
// for multi-element tiles with the deterministic scatter: their loads are in
// flight together).  Multi-element tiles with the deterministic scatter (rows
// stored [r][slot]) stage the block's rows in LDS and store each row index's
// consecutive slots together.
template <int N, typename KT, typename VT, bool MASKED, bool ATOMIC, int TPW, bool SYM = false>
__global__ __launch_bounds__(256) void ebe_tile_kernel(const KT* __restrict__ K, const uint32_t* __restrict__ map,
                                                      const VT* __restrict__ u, VT* __restrict__ out,
                                                      uint64_t ve_base, uint32_t rs, uint32_t ss, uint32_t tile_begin,
                                                      uint32_t n_tiles, const uint32_t* __restrict__ skip,
                                                      const float* __restrict__ kdiag)
{
    constexpr int G = 64 / N;
    constexpr uint32_t TE = sym_region_values<N, KT>();
    // packed symmetric tiles: one LDS region per tile of the block
    __shared__ __align__(16) KT ksh[SYM ? 4u * TPW * TE : 1u];
    if (skip && *skip) return;  // converged solver: nothing to do
    const uint32_t w = threadIdx.x >> 6;
    if constexpr (G == 1 || ATOMIC) {
        const uint32_t tb = xcd_block(blockIdx.x, gridDim.x) * 4u;
        if (tb + w >= n_tiles) return;  // uniform per wave
        ebe_tile_body<N, KT, VT, MASKED, ATOMIC, SYM>(K, map, u, out, ve_base, rs, ss, tile_begin + tb + w, kdiag,
                                                      ksh + (SYM ? w * TE : 0u));
    } else {
        constexpr int GN = G * N;
        constexpr uint32_t BT = 4u * TPW;  // tiles per block
        __shared__ VT st[N * BT * G];      // [r][tile in block * G + g]
        const uint32_t tb = xcd_block(blockIdx.x, gridDim.x) * BT;
        if (tb >= n_tiles) return;         // uniform per block
        const uint32_t nt = min(BT, n_tiles - tb), lane = threadIdx.x & 63u;
        VT acc[TPW];
        uint32_t m;
        if constexpr (SYM) {
            const uint32_t first = w * TPW, ntw = nt > first ? min((uint32_t)TPW, nt - first) : 0u;
            ebe_tiles_sym<N, KT, VT, MASKED, TPW>(K, map, u, tile_begin + tb + first, ntw, kdiag, acc,
                                                  ksh + first * TE);
        } else if (nt == BT) {
#pragma unroll
            for (int i = 0; i < TPW; i++)
                acc[i] = ebe_tile_acc<N, KT, VT, MASKED>(K, map, u, tile_begin + tb + w * TPW + i, kdiag, m);
        } else {
#pragma unroll
            for (int i = 0; i < TPW; i++)
                if (w * TPW + i < nt)
                    acc[i] = ebe_tile_acc<N, KT, VT, MASKED>(K, map, u, tile_begin + tb + w * TPW + i, kdiag, m);
        }
        if (lane < (uint32_t)GN) {
            const uint32_t g = lane / N, r = lane - g * N;
#pragma unroll
            for (int i = 0; i < TPW; i++)
                if (w * TPW + i < nt) st[r * (BT * G) + (w * TPW + i) * G + g] = acc[i];
        }
        __syncthreads();
        const uint32_t per_r = nt * G;
        const uint64_t slot0 = (uint64_t)(tile_begin + tb) * G;
        for (uint32_t i = threadIdx.x; i < (uint32_t)N * per_r; i += 256u) {
            const uint32_t r = i / per_r, k = i - r * per_r;
            out[ve_base + (uint64_t)r * rs + (slot0 + k) * ss] = st[r * (BT * G) + k];
        }
    }
}

// tiles per wave of the element kernel: two for multi-element tiles of FP32 /
// BF16 storage (measured: tet10 96^3 BF16 matvec 506 -> 476 us, hex8 128^3 122 ->
// 109 us; FP32 641 -> 629 and 142 -> 136 us), one for FP64 (two were 2-4 % slower)
// and for one-element tiles
template <int N, typename KT> constexpr int tiles_per_wave()
{
    return (64 / N == 1 || KSize<KT>::v == 8) ? 1 : 2;
}
// packed symmetric tiles per wave, their loads issued together: two for FP32
// and BF16 storage (tet10 96^3 matvec 0.486 -> 0.453 ms and 0.418 -> 0.402 ms;
// four: 0.486 and 0.405), one for FP64 (two: 0.788 against 0.781)
template <int N, typename KT> constexpr int sym_tiles_per_wave() { return KSize<KT>::v == 8 ? 1 : 2; }

// ------------------------------------------------ near-symmetric records
// One wave per element of a one-element FP64 tile group stored as
// near-symmetric records (amx_nsym.hpp).  The wave copies the record's U and
// X (16-byte loads, every lane) and its gathered u_e into its LDS region; lane
// r loads its delta chunks into registers.  Lane r then forms row r with c
// ascending from U[c][r] plus its delta (c < r: consecutive lanes read
// consecutive LDS words) or U[r][c] (c >= r): the products and sums of the
// whole layout in the same order, so the result is bit-identical.  The LDS
// word of (r, c) is min(off(c) - c + r, off(r) - r + c) (nsym_row); columns
// that hold an exception in some row (colany, the same for every element of
// the group) also move the lane's cursor over its X values.
template <int N> constexpr uint32_t nsym_rounds() { return (nsym_nsp(N) / 2u + 63u) / 64u + 2u; }
// exception values a record may hold: what the copy rounds hold beyond U
template <int N> constexpr uint32_t nsym_max_x() { return nsym_rounds<N>() * 128u - nsym_nsp(N); }

// LDS values of one wave's region: U, X and u_e
__host__ __device__ constexpr uint32_t nsym_region_values(uint32_t n, uint32_t nx)
{
    return nsym_nsp(n) + nsym_nxp(nx) + 64u;
}

// f(integral_constant<int, I>) for I = B .. E-1
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// The loads of one record: U and X chunks (all lanes), the lane's delta
// chunks, its map entry.  Every load is unconditional (indices clamped into the
// record), so that no wait lands between them.
template <int N> struct NsymRec {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 ch[nsym_rounds<N>()];
    u32x4 dq[nsym_nq(N)];
    uint32_t m;

    __device__ __forceinline__ void issue(const uint8_t* __restrict__ rec, uint32_t nch,
                                          const uint32_t* __restrict__ map, uint64_t slot)
    {
        const uint32_t lane = threadIdx.x & 63u, lc = lane < (uint32_t)N ? lane : (uint32_t)N - 1u;
        // the map entry first: the u gather waits for it, not for the record
        m = map[slot * N + lc];
        __builtin_amdgcn_sched_barrier(0);
        const u32x4* src = reinterpret_cast<const u32x4*>(rec);
#pragma unroll
        for (uint32_t i = 0; i < nsym_rounds<N>(); i++)
            ch[i] = __builtin_nontemporal_load(src + min(lane + 64u * i, nch - 1u));
        const u32x4* dsrc = src + nch;
#pragma unroll
        for (uint32_t q = 0; q < nsym_nq(N); q++)
            dq[q] = __builtin_nontemporal_load(dsrc + nsym_dbase(N, q) + (lc >= 16u * q + 1u ? lc - 16u * q - 1u : 0u));
        // keep these loads ahead of what follows (the u gather, the compute)
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
    }
};

template <bool MASKED>
__device__ __forceinline__ double gather_u(const double* __restrict__ u, uint32_t m)
{
    double uv = u[m & kIdxMask];
    if (MASKED && (m & kFlagBit)) uv = 0.0;
    return uv;
}

// copy record R's U and X and the gathered u_e into the wave's LDS region wl
// (waits for the record); returns the lane's delta chunks, zero for the
// chunk columns of rows r <= 16q (no lower entry there)
template <int N>
__device__ __forceinline__ void nsym_stage(const NsymRec<N>& R, uint32_t nch, uint32_t nxp, double uv, double* wl,
                                           typename NsymRec<N>::u32x4 (&dq)[nsym_nq(N)])
{
    typedef typename NsymRec<N>::u32x4 u32x4;
    const uint32_t lane = threadIdx.x & 63u, lc = lane < (uint32_t)N ? lane : (uint32_t)N - 1u;
    u32x4* dst = reinterpret_cast<u32x4*>(wl);
#pragma unroll
    for (uint32_t i = 0; i < nsym_rounds<N>(); i++)
        if (lane + 64u * i < nch) dst[lane + 64u * i] = R.ch[i];
    wl[nsym_nsp(N) + nxp + lane] = uv;
#pragma unroll
    for (uint32_t q = 0; q < nsym_nq(N); q++) dq[q] = lc < 16u * q + 1u ? u32x4{0u, 0u, 0u, 0u} : R.dq[q];
    wave_lds_sync();
}

// row r = lc of K_e u_e from the staged record, c ascending.  The LDS word
// of (r, c) is U[c][r] = off(c) + r - c for c < r and U[r][c] = off(r) + c - r
// for c >= r (off(a) = a n - a(a-1)/2); the lower form is the smaller of the
// two while c < r and the larger after (rows shrink by one value each), so
// the word is min(off(c) - c + r, off(r) - r + c): two adds and a min, with
// nothing carried from column to column.
template <int N>
__device__ __forceinline__ double nsym_row(const typename NsymRec<N>::u32x4 (&dq)[nsym_nq(N)], uint32_t nxp,
                                           uint32_t xs, uint64_t colany, const double* wl)
{
    constexpr uint32_t NSP = nsym_nsp(N);
    const uint32_t lane = threadIdx.x & 63u, lc = lane < (uint32_t)N ? lane : (uint32_t)N - 1u;
    const char* kb = reinterpret_cast<const char*>(wl);
    const double* ush = wl + NSP + nxp;
    // byte offsets in the region: 8 r and 8 (off(r) - r), the lane's next X value
    const uint32_t r8 = 8u * lc, b8 = 8u * (nsym_uidx(N, lc, lc) - lc);
    uint32_t px = 8u * (NSP + xs);
    double acc = 0.0;
    static_for<0, N>([&](auto ci) {
        constexpr int c = decltype(ci)::value;
        constexpr uint32_t A8 = 8u * (nsym_uidx(N, c, c) - c);
        const uint32_t wd = dq[c / 16][(c % 16) / 4];
        const int32_t d = ((int32_t)(wd << (24 - 8 * (c % 4)))) >> 24;
        // some row has an exception in column c (uniform) and this lane's is one
        const bool exc = ((colany >> c) & 1u) != 0 && d == -128;
        const uint32_t a = exc ? px : min(r8 + A8, b8 + 8u * c);
        px += exc ? 8u : 0u;
        uint64_t bits = *reinterpret_cast<const uint64_t*>(kb + a);
        bits = (bits & 0xffffffff00000000ull) | (uint32_t)((uint32_t)bits + (uint32_t)d);
        acc = madd(acc, __longlong_as_double((long long)bits), ush[c]);
    });
    return acc;
}

// one element: row lc of K_e u_e (m: the lane's map entry)
template <int N, bool MASKED>
__device__ __forceinline__ double ebe_nsym_acc(const uint8_t* __restrict__ rec, uint32_t nxp, uint32_t xs,
                                               uint64_t colany, const uint32_t* __restrict__ map,
                                               const double* __restrict__ u, uint64_t slot, double* wl, uint32_t& m)
{
    const uint32_t nch = (nsym_nsp(N) + nxp) / 2u;  // 16-byte chunks of U and X
    NsymRec<N> R;
    R.issue(rec, nch, map, slot);
    m = R.m;
    typename NsymRec<N>::u32x4 dq[nsym_nq(N)];
    nsym_stage<N>(R, nch, nxp, gather_u<MASKED>(u, R.m), wl, dq);
    return nsym_row<N>(dq, nxp, xs, colany, wl);
}

// one element per wave, one wave per block (more elements per wave with the
// next record in flight during the compute were slower: the registers they
// need leave two waves per SIMD instead of the 2-3 the LDS regions allow),
// dynamic LDS of nsym_region_values(N, nx) doubles
template <int N, bool MASKED, bool ATOMIC>
__global__ __launch_bounds__(64) void ebe_nsym_kernel(const uint8_t* __restrict__ K, uint64_t rec_bytes,
                                                     uint32_t nxp, const uint32_t* __restrict__ xs, uint64_t colany,
                                                     const uint32_t* __restrict__ map, const double* __restrict__ u,
                                                     double* __restrict__ out, uint64_t ve_base, uint32_t rs,
                                                     uint32_t ss, uint32_t tile_begin, uint32_t n_tiles,
                                                     const uint32_t* __restrict__ skip)
{
    extern __shared__ __align__(16) double nsym_lds[];
    if (skip && *skip) return;  // converged solver: nothing to do
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= n_tiles) return;
    const uint64_t slot = (uint64_t)tile_begin + t;
    const uint32_t lane = threadIdx.x, lc = lane < (uint32_t)N ? lane : (uint32_t)N - 1u;
    uint32_t m;
    const double acc = ebe_nsym_acc<N, MASKED>(K + slot * rec_bytes, nxp, xs[lc], colany, map, u, slot, nsym_lds, m);
    if (lane < (uint32_t)N) {
        if constexpr (ATOMIC) atomic_add(&out[m & kIdxMask], acc);
        else out[ve_base + (uint64_t)lane * rs + slot * ss] = acc;
    }
}

// ---------------------------------- packed symmetric records with exceptions
// FP32 / BF16 storage of the same 60-DoF groups (amx_nsym.hpp): one wave per
// element, the record (U then X) and u_e staged in the wave's LDS region as
// for near-symmetric records; lane r reads the value of (r, c) at
// min(off(c) - c + r, off(r) - r + c) or, where bit c of its row's exception
// mask is set (a per-group table: mask low, mask high, first X index, per
// lane), its next X value -- the values and the order of the whole layout
// (BF16: the zero diagonal slot, then the FP32 diagonal as that layout adds it).
template <int N, typename KT> constexpr uint32_t psx_rounds()
{
    return (nsym_ns(N) * (uint32_t)sizeof(KT) / 16u + 63u) / 64u + 1u;
}
template <int N, typename KT> constexpr uint32_t psx_max_x()
{
    return psx_rounds<N, KT>() * 64u * 16u / (uint32_t)sizeof(KT) - nsym_ns(N);
}
// LDS bytes of one wave's region: the record's values and u_e (64 values of VT)
template <typename VT>
__host__ __device__ constexpr uint32_t psx_region_bytes(uint32_t n, uint32_t nx, uint32_t ksize)
{
    return psx_record_values(n, nx, ksize) * ksize + 64u * (uint32_t)sizeof(VT);
}

template <int N, typename KT, typename VT, bool MASKED>
__device__ __forceinline__ VT psx_acc(const KT* __restrict__ rec, uint32_t nrv, const uint32_t* __restrict__ ptab,
                                      const uint32_t* __restrict__ map, const VT* __restrict__ u,
                                      const float* __restrict__ kdiag, uint64_t slot, char* wl, uint32_t& m)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr uint32_t NR = psx_rounds<N, KT>(), NS = nsym_ns(N);
    constexpr bool BF = std::is_same<KT, uint16_t>::value;
    const uint32_t lane = threadIdx.x & 63u, lc = lane < (uint32_t)N ? lane : (uint32_t)N - 1u;
    const uint32_t nch = nrv * (uint32_t)sizeof(KT) / 16u;
    // the map entry first (the u gather waits for it), then the record
    m = map[slot * N + lc];
    __builtin_amdgcn_sched_barrier(0);
    const u32x4* src = reinterpret_cast<const u32x4*>(rec);
    u32x4 ch[NR];
#pragma unroll
    for (uint32_t i = 0; i < NR; i++) ch[i] = __builtin_nontemporal_load(src + min(lane + 64u * i, nch - 1u));
    const uint32_t mlo = ptab[3u * lane], mhi = ptab[3u * lane + 1u], xs = ptab[3u * lane + 2u];
    float kd = 0.0f;
    if constexpr (BF) kd = kdiag[slot * N + lc];
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    VT uv = u[m & kIdxMask];
    if (MASKED && (m & kFlagBit)) uv = VT(0);
    u32x4* dst = reinterpret_cast<u32x4*>(wl);
#pragma unroll
    for (uint32_t i = 0; i < NR; i++)
        if (lane + 64u * i < nch) dst[lane + 64u * i] = ch[i];
    VT* ush = reinterpret_cast<VT*>(wl + 16u * nch);
    ush[lane] = uv;
    wave_lds_sync();
    const KT* kv = reinterpret_cast<const KT*>(wl);
    const uint32_t b = nsym_uidx(N, lc, lc) - lc;  // off(r) - r
    uint32_t px = NS + xs;
    VT acc = VT(0);
    static_for<0, N>([&](auto ci) {
        constexpr int c = decltype(ci)::value;
        const bool exc = (((c < 32 ? mlo >> (c & 31) : mhi >> (c & 31))) & 1u) != 0;
        const uint32_t a = exc ? px : min(lc + (nsym_uidx(N, c, c) - c), b + c);
        px += exc ? 1u : 0u;
        acc = madd(acc, (VT)kval(kv[a]), ush[c]);
    });
    if constexpr (BF) acc = madd(acc, (VT)kd, uv);
    return acc;
}

template <int N, typename KT, typename VT, bool MASKED, bool ATOMIC>
__global__ __launch_bounds__(64) void psx_kernel(const KT* __restrict__ K, uint32_t nrv,
                                                const uint32_t* __restrict__ ptab, const uint32_t* __restrict__ map,
                                                const VT* __restrict__ u, VT* __restrict__ out, uint64_t ve_base,
                                                uint32_t rs, uint32_t ss, uint32_t tile_begin, uint32_t n_tiles,
                                                const uint32_t* __restrict__ skip, const float* __restrict__ kdiag)
{
    extern __shared__ __align__(16) char psx_lds[];
    if (skip && *skip) return;  // converged solver: nothing to do
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= n_tiles) return;
    const uint64_t slot = (uint64_t)tile_begin + t;
    const uint32_t lane = threadIdx.x;
    uint32_t m;
    const VT acc = psx_acc<N, KT, VT, MASKED>(K + slot * nrv, nrv, ptab, map, u, kdiag, slot, psx_lds, m);
    if (lane < (uint32_t)N) {
        if constexpr (ATOMIC) atomic_add(&out[m & kIdxMask], acc);
        else out[ve_base + (uint64_t)lane * rs + slot * ss] = acc;
    }
}

// index of the entry that owns work item x: the largest j < n with prefix[j] <= x
// (x wave-uniform, so the search runs on the scalar unit)
__device__ __forceinline__ uint32_t owner_of(const uint32_t* __restrict__ prefix, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Many operators (jobs) in one launch: wave -> tile of entry owner_of(prefix).
// Same per-tile arithmetic as ebe_tile_kernel (deterministic scatter only).
// Jobs with near-symmetric records or packed records with exceptions run in
// rec_batched_kernel instead.
template <int N, typename KT, typename VT>
__global__ __launch_bounds__(256) void ebe_tile_batched_kernel(const EbeJob* __restrict__ jobs,
                                                              const uint32_t* __restrict__ prefix, uint32_t n_jobs)
{
    constexpr bool SYM_OK = sym_region_values<N, KT>() > 1u;
    __shared__ __align__(16) KT ksh[SYM_OK ? 4u * sym_region_values<N, KT>() : 1u];
    const uint32_t gt = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4u + (threadIdx.x >> 6));
    if (gt >= prefix[n_jobs]) return;
    const uint32_t j = owner_of(prefix, n_jobs, gt);
    const EbeJob& jb = jobs[j];
    if (jb.skip && *jb.skip) return;
    const uint32_t t = jb.tile_begin + (gt - prefix[j]);
    if constexpr (SYM_OK) {
        if (jb.sym) {
            KT* reg = ksh + (threadIdx.x >> 6) * sym_region_values<N, KT>();
            if (jb.masked)
                ebe_tile_body<N, KT, VT, true, false, true>(static_cast<const KT*>(jb.K), jb.map,
                                                            static_cast<const VT*>(jb.u), static_cast<VT*>(jb.out),
                                                            jb.ve_base, jb.ve_rs, jb.ve_ss, t, jb.kdiag, reg);
            else
                ebe_tile_body<N, KT, VT, false, false, true>(static_cast<const KT*>(jb.K), jb.map,
                                                             static_cast<const VT*>(jb.u), static_cast<VT*>(jb.out),
                                                             jb.ve_base, jb.ve_rs, jb.ve_ss, t, jb.kdiag, reg);
            return;
        }
    }
    if (jb.masked)
        ebe_tile_body<N, KT, VT, true, false>(static_cast<const KT*>(jb.K), jb.map, static_cast<const VT*>(jb.u),
                                              static_cast<VT*>(jb.out), jb.ve_base, jb.ve_rs, jb.ve_ss, t, jb.kdiag);
    else
        ebe_tile_body<N, KT, VT, false, false>(static_cast<const KT*>(jb.K), jb.map, static_cast<const VT*>(jb.u),
                                               static_cast<VT*>(jb.out), jb.ve_base, jb.ve_rs, jb.ve_ss, t,
                                               jb.kdiag);
}

// Many jobs whose 60-DoF groups hold near-symmetric records (FP64) or packed
// records with exceptions (FP32 / BF16), in one launch with the one-operator
// kernels' shape: one-wave blocks, block -> element of entry owner_of(prefix),
// the dynamic LDS one region sized for the largest exception count of the
// launch's jobs (so as many waves stay resident per CU as in a one-operator
// launch).  Same arithmetic as ebe_nsym_kernel / psx_kernel.  MASKED: every
// job of the launch has the identity BC form.
template <typename KT, typename VT, bool MASKED>
__global__ __launch_bounds__(64) void rec_batched_kernel(const EbeJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ prefix, uint32_t n_jobs)
{
    extern __shared__ __align__(16) char rec_lds[];
    const uint32_t gt = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x));
    if (gt >= prefix[n_jobs]) return;
    const uint32_t j = owner_of(prefix, n_jobs, gt);
    const EbeJob& jb = jobs[j];
    if (jb.skip && *jb.skip) return;
    const uint32_t t = jb.tile_begin + (gt - prefix[j]), lane = threadIdx.x;
    uint32_t m;
    VT acc;
    if constexpr (std::is_same<KT, double>::value) {
        static_assert(std::is_same<VT, double>::value, "near-symmetric records: FP64 vectors");
        const uint32_t nxp = nsym_nxp(jb.nsym_nx), lc = lane < 60u ? lane : 59u;
        const uint8_t* rec = static_cast<const uint8_t*>(jb.K) + (uint64_t)t * nsym_record_bytes_of(60, jb.nsym_nx);
        double* wl = reinterpret_cast<double*>(rec_lds);
        acc = ebe_nsym_acc<60, MASKED>(rec, nxp, jb.nsym_xs[lc], jb.nsym_colany, jb.map,
                                       static_cast<const double*>(jb.u), t, wl, m);
    } else {
        const uint32_t nrv = psx_record_values(60, jb.nsym_nx, (uint32_t)sizeof(KT));
        const KT* rec = static_cast<const KT*>(jb.K) + (uint64_t)t * nrv;
        acc = psx_acc<60, KT, VT, MASKED>(rec, nrv, jb.nsym_xs, jb.map, static_cast<const VT*>(jb.u), jb.kdiag, t,
                                          rec_lds, m);
    }
    if (lane < 60u) static_cast<VT*>(jb.out)[jb.ve_base + (uint64_t)lane * jb.ve_rs + (uint64_t)t * jb.ve_ss] = acc;
}

// Any n: one wave per element, column-major storage (G = 1, P = 1), element
// vector staged in LDS, rows strided over the lanes.
constexpr int kGenericMaxN = 1024;

template <typename KT, typename VT, bool MASKED, bool ATOMIC>
__global__ __launch_bounds__(256) void ebe_generic_kernel(const KT* __restrict__ K, const uint32_t* __restrict__ map,
                                                         const VT* __restrict__ u, VT* __restrict__ out,
                                                         uint64_t ve_base, uint32_t rs, uint32_t ss, uint32_t n,
                                                         uint32_t slot_begin, uint32_t n_slots,
                                                         const uint32_t* __restrict__ skip,
                                                         const float* __restrict__ kdiag)
{
    if (skip && *skip) return;
    __shared__ VT ue[4][kGenericMaxN];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t sloc = xcd_block(blockIdx.x, gridDim.x) * 4u + w;
    if (sloc >= n_slots) return;
    const uint64_t slot = slot_begin + sloc;
    for (uint32_t c = lane; c < n; c += 64) {
        const uint32_t m = map[slot * n + c];
        VT uv = u[m & kIdxMask];
        if (MASKED && (m & kFlagBit)) uv = VT(0);
        ue[w][c] = uv;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const KT* ke = K + slot * n * n;
    for (uint32_t r = lane; r < n; r += 64) {
        VT acc = VT(0);
        for (uint32_t c = 0; c < n; c++) {
            KPack<KT, 1> k;
            k.load(ke + (uint64_t)c * n + r);
            acc = madd(acc, (VT)k.get(0), ue[w][c]);
        }
        if constexpr (std::is_same<KT, uint16_t>::value) acc = madd(acc, (VT)kdiag[slot * n + r], ue[w][r]);
        if constexpr (ATOMIC) atomic_add(&out[map[slot * n + r] & kIdxMask], acc);
        else out[ve_base + (uint64_t)r * rs + slot * ss] = acc;
    }
}

// ------------------------------------------------------------------ gather
// A block owns 256 consecutive DoFs d0 .. d0+255, whose lists are the
// contiguous entry range [off[d0], off[d0+256]).  The block reads that range
// in windows of kGatherWindow entries: every thread loads entries tid,
// tid + 256, ... of the window (list index, then element row) with all loads
// in flight together, and the rows land in LDS; then each thread adds the
// part of its own DoF's list inside the window, in list order, to its running
// sum.  The loads no longer follow one list per lane (no dependent chains, no
// lanes idle while a neighbour walks a long list), and every sum keeps the
// reference's element order, so the result is the same bit for bit.  The list
// indices are read once per application and loaded non-temporally, so they
// do not displace the element rows, map and vectors that the next element
// kernel and gather find in the caches (hex8 128^3: element kernel 0.143 ->
// 0.135 ms, gather 49 -> 46 us).
constexpr uint32_t kGatherWindow = 2048;

template <typename VT, int BC>
__device__ __forceinline__ void gather_block(const uint32_t* __restrict__ off, const uint32_t* __restrict__ pos,
                                             const VT* __restrict__ ve, const VT* __restrict__ u,
                                             const uint8_t* __restrict__ flag, const double* __restrict__ coef,
                                             VT* __restrict__ v, uint32_t n, uint32_t d0, VT* __restrict__ buf)
{
    constexpr uint32_t PER = kGatherWindow / 256u;
    const uint32_t tid = threadIdx.x, d = d0 + tid;
    const uint32_t dn = min(256u, n - d0);
    const uint32_t e0 = off[d0], e1 = off[d0 + dn];  // uniform
    uint32_t b = 0, c = 0;
    if (d < n) b = off[d], c = off[d + 1];
    VT sum = VT(0);
    for (uint32_t w = e0; w < e1; w += kGatherWindow) {
        const uint32_t cnt = min(kGatherWindow, e1 - w);
        uint32_t p[PER];
        VT x[PER];
        // entries past the window's end re-read its last entry and are not kept
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) p[j] = __builtin_nontemporal_load(pos + w + min(tid + 256u * j, cnt - 1u));
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) x[j] = ve[p[j]];
#pragma unroll
        for (uint32_t j = 0; j < PER; j++)
            if (tid + 256u * j < cnt) buf[tid + 256u * j] = x[j];
        __syncthreads();
        const uint32_t lo = max(b, w), hi = min(c, w + cnt);
        for (uint32_t q = lo; q < hi; q++) sum += buf[q - w];
        __syncthreads();  // the window's sums precede the next window's stores
    }
    if (d >= n) return;
    if constexpr (BC == 1) {
        if (flag[d]) sum = u[d];
    } else if constexpr (BC == 2) {
        if (flag[d]) sum = (VT)((double)sum + coef[d] * (double)u[d]);  // penalty row in FP64
    }
    v[d] = sum;
}

template <typename VT, int BC>
__global__ __launch_bounds__(256) void gather_kernel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ pos,
                                                    const VT* __restrict__ ve, const VT* __restrict__ u,
                                                    const uint8_t* __restrict__ flag, const double* __restrict__ coef,
                                                    VT* __restrict__ v, uint32_t n, const uint32_t* __restrict__ skip)
{
    __shared__ VT buf[kGatherWindow];
    if (skip && *skip) return;
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x)) * 256u;
    if (d0 >= n) return;  // uniform per block
    gather_block<VT, BC>(off, pos, ve, u, flag, coef, v, n, d0, buf);
}

// many jobs in one launch: block -> 256 DoFs of entry owner_of(prefix)
template <typename VT>
__global__ __launch_bounds__(256) void gather_batched_kernel(const GatherJob* __restrict__ jobs,
                                                            const uint32_t* __restrict__ prefix, uint32_t n_jobs)
{
    __shared__ VT buf[kGatherWindow];
    const uint32_t gb = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x));
    if (gb >= prefix[n_jobs]) return;
    const uint32_t j = owner_of(prefix, n_jobs, gb);
    const GatherJob& jb = jobs[j];
    if (jb.skip && *jb.skip) return;
    const uint32_t d0 = (gb - prefix[j]) * 256u;
    if (d0 >= jb.n) return;
    const VT* ve = static_cast<const VT*>(jb.ve);
    const VT* u = static_cast<const VT*>(jb.u);
    VT* v = static_cast<VT*>(jb.v);
    if (jb.bc == 1) gather_block<VT, 1>(jb.off, jb.pos, ve, u, jb.flag, jb.coef, v, jb.n, d0, buf);
    else if (jb.bc == 2) gather_block<VT, 2>(jb.off, jb.pos, ve, u, jb.flag, jb.coef, v, jb.n, d0, buf);
    else gather_block<VT, 0>(jb.off, jb.pos, ve, u, jb.flag, jb.coef, v, jb.n, d0, buf);
}

template <typename VT, int BC>
__global__ void bc_fixup_kernel(const uint32_t* __restrict__ cdofs, uint32_t nc, const double* __restrict__ coef,
                                const VT* __restrict__ u, VT* __restrict__ v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const uint32_t d = cdofs[i];
    if constexpr (BC == 1) v[d] = u[d];
    else v[d] = (VT)((double)v[d] + coef[d] * (double)u[d]);
}

// ------------------------------------------------------------------ set-up
struct TiledIndex {
    uint32_t N, G, P;
    // destination index -> (slot, r, c)
    __device__ __forceinline__ void decompose(uint64_t idx, uint64_t& slot, uint32_t& r, uint32_t& c) const
    {
        const uint32_t p = (uint32_t)(idx % P);
        uint64_t rest = idx / P;
        const uint32_t GN = G * N;
        const uint32_t l = (uint32_t)(rest % GN);
        rest /= GN;
        const uint32_t NP = N / P;
        const uint32_t j = (uint32_t)(rest % NP);
        const uint64_t t = rest / NP;
        const uint32_t g = l / N;
        r = l - g * N;
        c = j * P + p;
        slot = t * G + g;
    }
};

template <typename KT>
__global__ void relayout_rowmajor_kernel(const double* __restrict__ src, const uint32_t* __restrict__ slot_elem,
                                         uint64_t total, TiledIndex ti, KT* __restrict__ dst)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    uint64_t slot;
    uint32_t r, c;
    ti.decompose(idx, slot, r, c);
    const uint32_t e = slot_elem[slot];
    double val = (e == kPadSlot) ? 0.0 : src[(uint64_t)e * ti.N * ti.N + (uint64_t)r * ti.N + c];
    if (std::is_same<KT, uint16_t>::value && r == c) val = 0.0;  // diagonal kept in FP32 (bf16_diag_kernel)
    dst[idx] = from_f64<KT>(val);
}

template <typename KT>
__global__ void relayout_scaled_kernel(const double* __restrict__ kref, const double* __restrict__ scale,
                                       const uint32_t* __restrict__ slot_elem, uint64_t total, TiledIndex ti,
                                       KT* __restrict__ dst)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    uint64_t slot;
    uint32_t r, c;
    ti.decompose(idx, slot, r, c);
    const uint32_t e = slot_elem[slot];
    double val = (e == kPadSlot) ? 0.0 : scale[e] * kref[(uint64_t)r * ti.N + c];
    if (std::is_same<KT, uint16_t>::value && r == c) val = 0.0;  // diagonal kept in FP32 (bf16_diag_kernel)
    dst[idx] = from_f64<KT>(val);
}

// packed symmetric tiles from row-major matrices (src) or k_ref + scale: one
// thread per stored value, padding included (zeros)
template <typename KT>
__global__ void relayout_sym_kernel(const double* __restrict__ src, const double* __restrict__ kref,
                                    const double* __restrict__ scale, const uint32_t* __restrict__ slot_elem,
                                    uint64_t total, uint32_t N, uint32_t G, uint32_t TE, KT* __restrict__ dst)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint64_t t = idx / TE;
    const uint32_t w = (uint32_t)(idx - t * TE), NS = sym_ns(N);
    double val = 0.0;
    if (w < G * NS) {
        const uint32_t g = w / NS;
        uint32_t k = w - g * NS, a = 0;
        while (k >= N - a) {  // row a of the upper triangle holds N - a values
            k -= N - a;
            a++;
        }
        const uint32_t b = a + k;
        const uint32_t e = slot_elem[t * G + g];
        if (e != kPadSlot && !(std::is_same<KT, uint16_t>::value && a == b))  // BF16: diagonal kept in FP32
            val = src ? src[(uint64_t)e * N * N + (uint64_t)a * N + b] : scale[e] * kref[(uint64_t)a * N + b];
    }
    dst[idx] = from_f64<KT>(val);
}

// ---- near-symmetric records (amx_nsym.hpp)
__device__ __forceinline__ double src_value(const double* __restrict__ src, const double* __restrict__ kref,
                                            const double* __restrict__ scale, uint32_t e, uint32_t N, uint32_t r,
                                            uint32_t c)
{
    return src ? src[(uint64_t)e * N * N + (uint64_t)r * N + c] : scale[e] * kref[(uint64_t)r * N + c];
}

// lower entry k of a row-major strict lower triangle -> (r, c), r > c
__device__ __forceinline__ void lower_rc(uint32_t k, uint32_t& r, uint32_t& c)
{
    r = (uint32_t)((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
    while (r * (r - 1u) / 2u > k) r--;
    while ((r + 1u) * r / 2u <= k) r++;
    c = k - r * (r - 1u) / 2u;
}

// K[r][c] coded from K[c][r] by a byte on the low word (same high word,
// difference in [-127, 127])
__device__ __forceinline__ bool nsym_codable(double lower, double upper, int32_t& d)
{
    const uint64_t a = (uint64_t)__double_as_longlong(lower), b = (uint64_t)__double_as_longlong(upper);
    const int64_t diff = (int64_t)(a & 0xffffffffull) - (int64_t)(b & 0xffffffffull);
    d = (int32_t)diff;
    return (a >> 32) == (b >> 32) && diff >= -127 && diff <= 127;
}

// flags[r*N + c] = 1 for the lower entries that some matrix of the group
// cannot code (one thread per slot and lower entry; every writer stores 1):
// kt F64 -- the byte code of near-symmetric records; F32 / BF16 -- the
// rounded entry differs from the rounded transposed entry
__global__ void nsym_scan_kernel(const double* __restrict__ src, const double* __restrict__ kref,
                                 const double* __restrict__ scale, const uint32_t* __restrict__ slot_elem,
                                 uint64_t total, uint32_t N, int kt, uint32_t* __restrict__ flags)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint32_t NL = N * (N - 1u) / 2u;
    const uint64_t slot = idx / NL;
    const uint32_t e = slot_elem[slot];
    if (e == kPadSlot) return;
    uint32_t r, c;
    lower_rc((uint32_t)(idx - slot * NL), r, c);
    const double lo = src_value(src, kref, scale, e, N, r, c), up = src_value(src, kref, scale, e, N, c, r);
    bool ok;
    if (kt == (int)KType::F32) {
        ok = __float_as_uint(from_f64<float>(lo)) == __float_as_uint(from_f64<float>(up));
    } else if (kt == (int)KType::BF16) {
        ok = from_f64<uint16_t>(lo) == from_f64<uint16_t>(up);
    } else {
        int32_t d;
        ok = nsym_codable(lo, up, d);
    }
    if (!ok) flags[r * N + c] = 1u;
}

// packed symmetric records with exceptions, one thread per stored value;
// xrc[x] = r*N + c of exception x
template <typename KT>
__global__ void relayout_psx_kernel(const double* __restrict__ src, const double* __restrict__ kref,
                                    const double* __restrict__ scale, const uint32_t* __restrict__ slot_elem,
                                    uint64_t total, uint32_t N, uint32_t nx, uint32_t nrv,
                                    const uint16_t* __restrict__ xrc, KT* __restrict__ dst)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint64_t slot = idx / nrv;
    const uint32_t v = (uint32_t)(idx - slot * nrv);
    const uint32_t e = slot_elem[slot];
    double val = 0.0;
    if (e != kPadSlot) {
        if (v < nsym_ns(N)) {
            uint32_t a = 0, k = v;
            while (k >= N - a) {  // row a of the upper triangle holds N - a values
                k -= N - a;
                a++;
            }
            // BF16: the diagonal is kept in FP32 (bf16_diag_kernel), its slot is 0
            if (!(std::is_same<KT, uint16_t>::value && k == 0)) val = src_value(src, kref, scale, e, N, a, a + k);
        } else if (v < nsym_ns(N) + nx) {
            const uint32_t rc = xrc[v - nsym_ns(N)];
            val = src_value(src, kref, scale, e, N, rc / N, rc % N);
        }
    }
    dst[idx] = from_f64<KT>(val);
}

// records, one thread per 32-bit word.  xrc[x]: r*N + c of exception x;
// xidx[r*N + c]: its index, 0xffff elsewhere.
__global__ void relayout_nsym_kernel(const double* __restrict__ src, const double* __restrict__ kref,
                                     const double* __restrict__ scale, const uint32_t* __restrict__ slot_elem,
                                     uint64_t total, uint32_t N, uint32_t nx, const uint16_t* __restrict__ xrc,
                                     const uint16_t* __restrict__ xidx, uint32_t* __restrict__ dst)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint32_t RW = (uint32_t)(nsym_record_bytes_of(N, nx) / 4u);
    const uint64_t slot = idx / RW;
    const uint32_t w = (uint32_t)(idx - slot * RW);
    const uint32_t e = slot_elem[slot];
    const uint32_t NSP = nsym_nsp(N), UXW = 2u * (NSP + nsym_nxp(nx));
    uint32_t out = 0u;
    if (e != kPadSlot) {
        if (w < UXW) {  // U and X values, two words each
            const uint32_t v = w / 2u;
            uint64_t bits = 0;
            if (v < nsym_ns(N)) {
                uint32_t a = 0, k = v;
                while (k >= N - a) {  // row a of the upper triangle holds N - a values
                    k -= N - a;
                    a++;
                }
                bits = (uint64_t)__double_as_longlong(src_value(src, kref, scale, e, N, a, a + k));
            } else if (v >= NSP && v < NSP + nx) {
                const uint32_t rc = xrc[v - NSP];
                bits = (uint64_t)__double_as_longlong(src_value(src, kref, scale, e, N, rc / N, rc % N));
                bits = (bits & 0xffffffff00000000ull) | (uint32_t)((uint32_t)bits + 128u);
            }
            out = (w & 1u) ? (uint32_t)(bits >> 32) : (uint32_t)bits;
        } else {  // delta bytes, four per word
            const uint32_t chunk = (w - UXW) / 4u, b0 = ((w - UXW) % 4u) * 4u;
            uint32_t q = 0;
            while (chunk >= nsym_dbase(N, q + 1u)) q++;
            const uint32_t r = chunk - nsym_dbase(N, q) + 16u * q + 1u;
            for (uint32_t b = 0; b < 4u; b++) {
                const uint32_t c = 16u * q + b0 + b;
                int32_t d = 0;
                if (c < r) {
                    if (xidx[r * N + c] != 0xffffu) {
                        d = -128;
                    } else {
                        (void)nsym_codable(src_value(src, kref, scale, e, N, r, c),
                                           src_value(src, kref, scale, e, N, c, r), d);
                    }
                }
                out |= ((uint32_t)d & 0xffu) << (8u * b);
            }
        }
    }
    dst[idx] = out;
}

// diagonal entries of near-symmetric records into ve (row (slot, r) at ve[r*rs + slot*ss])
__global__ void nsym_diag_kernel(const double* __restrict__ K, uint32_t N, uint64_t rec_values, uint64_t n_rows,
                                 uint32_t rs, uint32_t ss, double* __restrict__ ve)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // slot * N + r
    if (i >= n_rows) return;
    const uint64_t slot = i / N;
    const uint32_t r = (uint32_t)(i - slot * N);
    ve[(uint64_t)r * rs + slot * ss] = K[slot * rec_values + nsym_uidx(N, r, r)];
}

// FP32 diagonal of the BF16 storage: d_r = K_rr + sum_{j != r} (K_rj - bf16(K_rj)),
// so that every row of the stored matrix sums to the FP64 row sum (constants
// stay in the null space of a Laplacian element); symmetric K stays symmetric.
__global__ void bf16_diag_kernel(const double* __restrict__ src, const double* __restrict__ kref,
                                 const double* __restrict__ scale, const uint32_t* __restrict__ slot_elem,
                                 uint32_t N, uint64_t n_rows, float* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // slot * N + r
    if (i >= n_rows) return;
    const uint64_t slot = i / N;
    const uint32_t r = (uint32_t)(i - slot * N);
    const uint32_t e = slot_elem[slot];
    if (e == kPadSlot) {
        out[i] = 0.0f;
        return;
    }
    double d = 0.0;
    for (uint32_t j = 0; j < N; j++) {
        const double kv = src ? src[(uint64_t)e * N * N + (uint64_t)r * N + j] : scale[e] * kref[(uint64_t)r * N + j];
        d += (j == r) ? kv : kv - (double)bf16_to_f32(f64_to_bf16(kv));
    }
    out[i] = (float)d;
}

__global__ void extract_diag_kernel(const double* __restrict__ K, TiledIndex ti, uint32_t sym_te, uint64_t n_rows,
                                    uint32_t rs, uint32_t ss, double* __restrict__ ve)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // slot * N + r
    if (i >= n_rows) return;
    const uint64_t slot = i / ti.N;
    const uint32_t r = (uint32_t)(i - slot * ti.N);
    const uint64_t t = slot / ti.G;
    const uint32_t g = (uint32_t)(slot - t * ti.G);
    const uint32_t GN = ti.G * ti.N, NP = ti.N / ti.P;
    const uint64_t k = sym_te ? t * sym_te + g * sym_ns(ti.N) + sym_idx(ti.N, r, r)
                              : ((t * NP + r / ti.P) * GN + g * ti.N + r) * ti.P + r % ti.P;
    ve[(uint64_t)r * rs + slot * ss] = K[k];
}

template <typename VT> __global__ void fill_kernel(VT* __restrict__ p, uint64_t n, VT value)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void coef_kernel(const double* __restrict__ diag, const uint32_t* __restrict__ cdofs, uint32_t nc,
                            double penalty, double* __restrict__ c64)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const uint32_t d = cdofs[i];
    c64[d] = penalty - diag[d];
}

__global__ void set_entries_kernel(double* __restrict__ u, const uint32_t* __restrict__ ids, uint32_t n, double value)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[ids[i]] = value;
}

__global__ void pack_kernel(double* __restrict__ dst, const double* __restrict__ src,
                            const uint32_t* __restrict__ ids, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[ids[i]];
}

__global__ void halo_accumulate_kernel(double* __restrict__ v, const double* __restrict__ recv,
                                       const uint32_t* __restrict__ dst, const uint32_t* __restrict__ off,
                                       const uint32_t* __restrict__ src, uint32_t n_unique)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_unique) return;
    double s = v[dst[i]];
    for (uint32_t q = off[i]; q < off[i + 1]; ++q) s += recv[src[q]];
    v[dst[i]] = s;
}

__global__ void d2f_kernel(const double* __restrict__ a, float* __restrict__ b, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void f2d_kernel(const float* __restrict__ a, double* __restrict__ b, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

template <int N, typename KT, typename VT>
void launch_tile(bool masked, bool atomic, bool sym, const EbeArgs& a, hipStream_t s)
{
    constexpr int TPW = tiles_per_wave<N, KT>(), STPW = sym_tiles_per_wave<N, KT>();
    const uint32_t per_block = atomic ? 4u : 4u * (sym ? STPW : TPW);
    const dim3 grid(blocks_for(a.n_tiles, per_block)), block(256);
    const KT* K = static_cast<const KT*>(a.K);
    const VT* u = static_cast<const VT*>(a.u);
    VT* out = static_cast<VT*>(a.out);
#define AMX_TILE(M, A, T, S)                                                                                       \
    ebe_tile_kernel<N, KT, VT, M, A, T, S><<<grid, block, 0, s>>>(K, a.map, u, out, a.ve_base, a.ve_rs, a.ve_ss,   \
                                                                   a.tile_begin, a.n_tiles, a.skip, a.kdiag)
    if constexpr (N == 8 || N == 10) {  // packed symmetric tiles (sym_supported)
        if (sym) {
            if (masked) {
                if (atomic) AMX_TILE(true, true, 1, true);
                else AMX_TILE(true, false, STPW, true);
            } else {
                if (atomic) AMX_TILE(false, true, 1, true);
                else AMX_TILE(false, false, STPW, true);
            }
            return;
        }
    }
    AMX_REQUIRE(!sym, AMX_INVALID_STATE, "no packed symmetric tile kernel for n = " + std::to_string(N));
    if (masked) {
        if (atomic) AMX_TILE(true, true, 1, false);
        else AMX_TILE(true, false, TPW, false);
    } else {
        if (atomic) AMX_TILE(false, true, 1, false);
        else AMX_TILE(false, false, TPW, false);
    }
#undef AMX_TILE
}


template <int N>
void launch_nsym_n(const Layout& L, bool masked, bool atomic, const EbeArgs& a, hipStream_t s)
{
    const uint32_t nxp = nsym_nxp(L.nx);
    const dim3 grid(a.n_tiles), block(64);
    const size_t lds = nsym_region_values(N, L.nx) * sizeof(double);
    const uint8_t* K = static_cast<const uint8_t*>(a.K);
    const uint64_t rb = nsym_record_bytes_of(N, L.nx);
    const double* u = static_cast<const double*>(a.u);
    double* out = static_cast<double*>(a.out);
#define AMX_NSYM(M, A)                                                                                             \
    ebe_nsym_kernel<N, M, A><<<grid, block, lds, s>>>(K, rb, nxp, a.nsym_xs, a.nsym_colany, a.map, u, out,        \
                                                      a.ve_base, a.ve_rs, a.ve_ss, a.tile_begin, a.n_tiles, a.skip)
    if (masked) {
        if (atomic) AMX_NSYM(true, true);
        else AMX_NSYM(true, false);
    } else {
        if (atomic) AMX_NSYM(false, true);
        else AMX_NSYM(false, false);
    }
#undef AMX_NSYM
}

template <int N, typename KT, typename VT>
void launch_psx(const Layout& L, bool masked, bool atomic, const EbeArgs& a, hipStream_t s)
{
    const uint32_t nrv = psx_record_values(N, L.nx, sizeof(KT));
    const dim3 grid(a.n_tiles), block(64);
    const size_t lds = psx_region_bytes<VT>(N, L.nx, sizeof(KT));
    const KT* K = static_cast<const KT*>(a.K);
    const VT* u = static_cast<const VT*>(a.u);
    VT* out = static_cast<VT*>(a.out);
#define AMX_PSX(M, A)                                                                                              \
    psx_kernel<N, KT, VT, M, A><<<grid, block, lds, s>>>(K, nrv, a.nsym_xs, a.map, u, out, a.ve_base, a.ve_rs,      \
                                                         a.ve_ss, a.tile_begin, a.n_tiles, a.skip, a.kdiag)
    if (masked) {
        if (atomic) AMX_PSX(true, true);
        else AMX_PSX(true, false);
    } else {
        if (atomic) AMX_PSX(false, true);
        else AMX_PSX(false, false);
    }
#undef AMX_PSX
}

template <typename KT, typename VT>
void launch_generic(uint32_t n, bool masked, bool atomic, const EbeArgs& a, hipStream_t s)
{
    const dim3 grid(blocks_for(a.n_tiles, 4)), block(256);
    const KT* K = static_cast<const KT*>(a.K);
    const VT* u = static_cast<const VT*>(a.u);
    VT* out = static_cast<VT*>(a.out);
    if (masked) {
        if (atomic) ebe_generic_kernel<KT, VT, true, true><<<grid, block, 0, s>>>(K, a.map, u, out, a.ve_base, a.ve_rs, a.ve_ss, n, a.tile_begin, a.n_tiles, a.skip, a.kdiag);
        else ebe_generic_kernel<KT, VT, true, false><<<grid, block, 0, s>>>(K, a.map, u, out, a.ve_base, a.ve_rs, a.ve_ss, n, a.tile_begin, a.n_tiles, a.skip, a.kdiag);
    } else {
        if (atomic) ebe_generic_kernel<KT, VT, false, true><<<grid, block, 0, s>>>(K, a.map, u, out, a.ve_base, a.ve_rs, a.ve_ss, n, a.tile_begin, a.n_tiles, a.skip, a.kdiag);
        else ebe_generic_kernel<KT, VT, false, false><<<grid, block, 0, s>>>(K, a.map, u, out, a.ve_base, a.ve_rs, a.ve_ss, n, a.tile_begin, a.n_tiles, a.skip, a.kdiag);
    }
}

template <typename KT, typename VT>
void dispatch_n(const Layout& L, bool masked, bool atomic, const EbeArgs& a, hipStream_t s)
{
    if (!L.specialized) return launch_generic<KT, VT>(L.n, masked, atomic, a, s);
    const bool sy = L.sym;
    switch (L.n) {
    case 4: return launch_tile<4, KT, VT>(masked, atomic, sy, a, s);
    case 8: return launch_tile<8, KT, VT>(masked, atomic, sy, a, s);
    case 10: return launch_tile<10, KT, VT>(masked, atomic, sy, a, s);
    case 12: return launch_tile<12, KT, VT>(masked, atomic, sy, a, s);
    case 20: return launch_tile<20, KT, VT>(masked, atomic, sy, a, s);
    case 24: return launch_tile<24, KT, VT>(masked, atomic, sy, a, s);
    case 30: return launch_tile<30, KT, VT>(masked, atomic, sy, a, s);
    case 60: return launch_tile<60, KT, VT>(masked, atomic, sy, a, s);
    default: fail(AMX_INVALID_STATE, "no tile kernel for n = " + std::to_string(L.n));
    }
}

bool has_tile_kernel(uint32_t n)
{
    switch (n) {
    case 4: case 8: case 10: case 12: case 20: case 24: case 30: case 60: return true;
    default: return false;
    }
}

}  // namespace

size_t ktype_size(KType kt) { return kt == KType::F64 ? 8 : kt == KType::F32 ? 4 : 2; }

bool sym_supported(uint32_t n) { return n == 8 || n == 10; }

bool nsym_supported(uint32_t n) { return n == 60; }

uint32_t nsym_max_exceptions(uint32_t n, KType kt)
{
    AMX_REQUIRE(nsym_supported(n), AMX_INVALID_ARGUMENT, "no near-symmetric kernel for n = " + std::to_string(n));
    return kt == KType::F64 ? nsym_max_x<60>() : kt == KType::F32 ? psx_max_x<60, float>() : psx_max_x<60, uint16_t>();
}

uint64_t nsym_record_bytes(uint32_t n, uint32_t nx) { return nsym_record_bytes_of(n, nx); }

uint64_t tile_values(const Layout& L, KType kt)
{
    if (L.nsym && kt == KType::F64) return nsym_record_bytes_of(L.n, L.nx) / 8u;
    if (L.nsym) return psx_record_values(L.n, L.nx, (uint32_t)ktype_size(kt));
    if (L.sym) return sym_tile_values(L.n, (uint32_t)ktype_size(kt));
    return (uint64_t)L.G * L.n * L.n;
}

Layout choose_layout(uint32_t n, KType kt)
{
    Layout L;
    L.n = n;
    if (has_tile_kernel(n)) {
        L.specialized = true;
        L.G = 64 / n;
        L.P = (uint32_t)pack_of((int)n, (int)ktype_size(kt));
    } else {
        AMX_REQUIRE(n <= (uint32_t)kGenericMaxN, AMX_NOT_IMPLEMENTED,
                    "elements with more than " + std::to_string(kGenericMaxN) + " DoFs");
        L.specialized = false;
        L.G = 1;
        L.P = 1;
    }
    return L;
}

void launch_relayout_sym(const double* src, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                         uint32_t n_slots, const Layout& L, KType kt, void* dst, hipStream_t s)
{
    const uint32_t TE = (uint32_t)tile_values(L, kt);
    const uint64_t total = (uint64_t)(n_slots / L.G) * TE;
    if (total == 0) return;
    const uint32_t nb = blocks_for(total, 256);
    switch (kt) {
    case KType::F64: relayout_sym_kernel<double><<<nb, 256, 0, s>>>(src, k_ref, scale, slot_elem, total, L.n, L.G, TE, (double*)dst); break;
    case KType::F32: relayout_sym_kernel<float><<<nb, 256, 0, s>>>(src, k_ref, scale, slot_elem, total, L.n, L.G, TE, (float*)dst); break;
    case KType::BF16: relayout_sym_kernel<uint16_t><<<nb, 256, 0, s>>>(src, k_ref, scale, slot_elem, total, L.n, L.G, TE, (uint16_t*)dst); break;
    }
    AMX_HIP(hipGetLastError());
}

void launch_relayout_from_rowmajor(const double* src, const uint32_t* slot_elem, uint32_t n_slots, const Layout& L,
                                   KType kt, void* dst, hipStream_t s)
{
    if (L.sym) return launch_relayout_sym(src, nullptr, nullptr, slot_elem, n_slots, L, kt, dst, s);
    const uint64_t total = (uint64_t)n_slots * L.n * L.n;
    if (total == 0) return;
    const TiledIndex ti{L.n, L.G, L.P};
    const uint32_t nb = blocks_for(total, 256);
    switch (kt) {
    case KType::F64: relayout_rowmajor_kernel<double><<<nb, 256, 0, s>>>(src, slot_elem, total, ti, (double*)dst); break;
    case KType::F32: relayout_rowmajor_kernel<float><<<nb, 256, 0, s>>>(src, slot_elem, total, ti, (float*)dst); break;
    case KType::BF16: relayout_rowmajor_kernel<uint16_t><<<nb, 256, 0, s>>>(src, slot_elem, total, ti, (uint16_t*)dst); break;
    }
    AMX_HIP(hipGetLastError());
}

void launch_relayout_scaled(const double* k_ref, const double* scale, const uint32_t* slot_elem, uint32_t n_slots,
                            const Layout& L, KType kt, void* dst, hipStream_t s)
{
    if (L.sym) return launch_relayout_sym(nullptr, k_ref, scale, slot_elem, n_slots, L, kt, dst, s);
    const uint64_t total = (uint64_t)n_slots * L.n * L.n;
    if (total == 0) return;
    const TiledIndex ti{L.n, L.G, L.P};
    const uint32_t nb = blocks_for(total, 256);
    switch (kt) {
    case KType::F64: relayout_scaled_kernel<double><<<nb, 256, 0, s>>>(k_ref, scale, slot_elem, total, ti, (double*)dst); break;
    case KType::F32: relayout_scaled_kernel<float><<<nb, 256, 0, s>>>(k_ref, scale, slot_elem, total, ti, (float*)dst); break;
    case KType::BF16: relayout_scaled_kernel<uint16_t><<<nb, 256, 0, s>>>(k_ref, scale, slot_elem, total, ti, (uint16_t*)dst); break;
    }
    AMX_HIP(hipGetLastError());
}

void launch_bf16_diag(const double* src_rowmajor, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                      uint32_t n_slots, uint32_t n, float* out, hipStream_t s)
{
    const uint64_t rows = (uint64_t)n_slots * n;
    if (rows == 0) return;
    bf16_diag_kernel<<<blocks_for(rows, 256), 256, 0, s>>>(src_rowmajor, k_ref, scale, slot_elem, n, rows, out);
    AMX_HIP(hipGetLastError());
}

void launch_nsym_scan(const double* src, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                      uint32_t n_slots, uint32_t n, KType kt, uint32_t* flags, hipStream_t s)
{
    const uint64_t total = (uint64_t)n_slots * (n * (n - 1u) / 2u);
    if (total == 0) return;
    nsym_scan_kernel<<<blocks_for(total, 256), 256, 0, s>>>(src, k_ref, scale, slot_elem, total, n, (int)kt, flags);
    AMX_HIP(hipGetLastError());
}

void launch_relayout_psx(const double* src, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                         uint32_t n_slots, const Layout& L, KType kt, const uint16_t* xrc, void* dst, hipStream_t s)
{
    const uint32_t nrv = psx_record_values(L.n, L.nx, (uint32_t)ktype_size(kt));
    const uint64_t total = (uint64_t)n_slots * nrv;
    if (total == 0) return;
    const uint32_t nb = blocks_for(total, 256);
    if (kt == KType::F32)
        relayout_psx_kernel<float><<<nb, 256, 0, s>>>(src, k_ref, scale, slot_elem, total, L.n, L.nx, nrv, xrc,
                                                      static_cast<float*>(dst));
    else
        relayout_psx_kernel<uint16_t><<<nb, 256, 0, s>>>(src, k_ref, scale, slot_elem, total, L.n, L.nx, nrv, xrc,
                                                         static_cast<uint16_t*>(dst));
    AMX_HIP(hipGetLastError());
}

void launch_relayout_nsym(const double* src, const double* k_ref, const double* scale, const uint32_t* slot_elem,
                          uint32_t n_slots, const Layout& L, const uint16_t* xrc, const uint16_t* xidx, void* dst,
                          hipStream_t s)
{
    const uint64_t total = (uint64_t)n_slots * (nsym_record_bytes_of(L.n, L.nx) / 4u);
    if (total == 0) return;
    relayout_nsym_kernel<<<blocks_for(total, 256), 256, 0, s>>>(src, k_ref, scale, slot_elem, total, L.n, L.nx, xrc,
                                                                xidx, static_cast<uint32_t*>(dst));
    AMX_HIP(hipGetLastError());
}

void launch_extract_diag(const double* K, const Layout& L, uint32_t n_slots, uint32_t rs, uint32_t ss, double* ve,
                         hipStream_t s)
{
    const uint64_t rows = (uint64_t)n_slots * L.n;
    if (rows == 0) return;
    if (L.nsym) {
        nsym_diag_kernel<<<blocks_for(rows, 256), 256, 0, s>>>(K, L.n, nsym_record_bytes_of(L.n, L.nx) / 8u, rows,
                                                               rs, ss, ve);
        AMX_HIP(hipGetLastError());
        return;
    }
    const uint32_t te = L.sym ? (uint32_t)tile_values(L, KType::F64) : 0u;
    extract_diag_kernel<<<blocks_for(rows, 256), 256, 0, s>>>(K, TiledIndex{L.n, L.G, L.P}, te, rows, rs, ss, ve);
    AMX_HIP(hipGetLastError());
}

void launch_ebe(const Layout& L, KType kt, bool vt_double, bool masked, bool atomic, const EbeArgs& a, hipStream_t s)
{
    if (a.n_tiles == 0) return;
    if (L.nsym) {
        AMX_REQUIRE(L.n == 60 && (kt != KType::F64 || vt_double), AMX_INVALID_STATE,
                    "near-symmetric records hold matrices of 60 DoFs");
        if (kt == KType::F64) launch_nsym_n<60>(L, masked, atomic, a, s);
        else if (kt == KType::F32 && vt_double) launch_psx<60, float, double>(L, masked, atomic, a, s);
        else if (kt == KType::F32) launch_psx<60, float, float>(L, masked, atomic, a, s);
        else if (vt_double) launch_psx<60, uint16_t, double>(L, masked, atomic, a, s);
        else launch_psx<60, uint16_t, float>(L, masked, atomic, a, s);
        AMX_HIP(hipGetLastError());
        return;
    }
    if (vt_double) {
        // FP64 vectors: FP64 storage, or FP32 / BF16 storage widened on load
        if (kt == KType::F64) dispatch_n<double, double>(L, masked, atomic, a, s);
        else if (kt == KType::F32) dispatch_n<float, double>(L, masked, atomic, a, s);
        else dispatch_n<uint16_t, double>(L, masked, atomic, a, s);
    } else if (kt == KType::F32) {
        dispatch_n<float, float>(L, masked, atomic, a, s);
    } else if (kt == KType::BF16) {
        dispatch_n<uint16_t, float>(L, masked, atomic, a, s);
    } else {
        fail(AMX_INVALID_ARGUMENT, "FP32 vectors need FP32 or BF16 element matrices");
    }
    AMX_HIP(hipGetLastError());
}

void launch_gather(bool vt_double, int bc, const uint32_t* off, const uint32_t* pos, const void* ve, const void* u,
                   const uint8_t* flag, const double* coef, void* v, uint32_t n, hipStream_t s, const uint32_t* skip)
{
    if (n == 0) return;
    const uint32_t nb = blocks_for(n, 256);
#define AMX_GATHER(VT, BC)                                                                                         \
    gather_kernel<VT, BC><<<nb, 256, 0, s>>>(off, pos, (const VT*)ve, (const VT*)u, flag, coef, (VT*)v, n, skip)
    if (vt_double) {
        if (bc == 0) AMX_GATHER(double, 0);
        else if (bc == 1) AMX_GATHER(double, 1);
        else AMX_GATHER(double, 2);
    } else {
        if (bc == 0) AMX_GATHER(float, 0);
        else if (bc == 1) AMX_GATHER(float, 1);
        else AMX_GATHER(float, 2);
    }
#undef AMX_GATHER
    AMX_HIP(hipGetLastError());
}

void launch_ebe_batched(uint32_t n, KType kt, bool vt_double, const EbeJob* jobs, const uint32_t* prefix,
                        uint32_t n_jobs, uint32_t total_tiles, hipStream_t s, bool records, uint32_t max_nx,
                        bool masked)
{
    if (total_tiles == 0 || n_jobs == 0) return;
    if (records) {  // every job: 60-DoF records, one element per tile
        AMX_REQUIRE(n == 60 && (kt != KType::F64 || vt_double), AMX_INVALID_STATE,
                    "near-symmetric records hold matrices of 60 DoFs with FP64 vectors");
        AMX_REQUIRE(max_nx <= nsym_max_exceptions(60, kt), AMX_INVALID_STATE, "more exceptions than a record holds");
        const uint32_t ks = kt == KType::F64 ? 8u : kt == KType::F32 ? 4u : 2u;
        const size_t lds = kt == KType::F64 ? nsym_region_values(60, max_nx) * sizeof(double)
                           : vt_double      ? psx_region_bytes<double>(60, max_nx, ks)
                                            : psx_region_bytes<float>(60, max_nx, ks);
#define AMX_REC_B(KTT, VTT)                                                                                        \
    do {                                                                                                           \
        if (masked) rec_batched_kernel<KTT, VTT, true><<<total_tiles, 64, lds, s>>>(jobs, prefix, n_jobs);         \
        else rec_batched_kernel<KTT, VTT, false><<<total_tiles, 64, lds, s>>>(jobs, prefix, n_jobs);               \
    } while (0)
        if (kt == KType::F64) AMX_REC_B(double, double);
        else if (kt == KType::F32 && vt_double) AMX_REC_B(float, double);
        else if (kt == KType::F32) AMX_REC_B(float, float);
        else if (vt_double) AMX_REC_B(uint16_t, double);
        else AMX_REC_B(uint16_t, float);
#undef AMX_REC_B
        AMX_HIP(hipGetLastError());
        return;
    }
    const uint32_t nb = blocks_for(total_tiles, 4);
    const size_t lds = 0;
#define AMX_EBE_B(NN, KTT, VTT) ebe_tile_batched_kernel<NN, KTT, VTT><<<nb, 256, lds, s>>>(jobs, prefix, n_jobs)
#define AMX_EBE_N(KTT, VTT)                                                                                        \
    switch (n) {                                                                                                   \
    case 4: AMX_EBE_B(4, KTT, VTT); break;                                                                         \
    case 8: AMX_EBE_B(8, KTT, VTT); break;                                                                         \
    case 10: AMX_EBE_B(10, KTT, VTT); break;                                                                       \
    case 12: AMX_EBE_B(12, KTT, VTT); break;                                                                       \
    case 20: AMX_EBE_B(20, KTT, VTT); break;                                                                       \
    case 24: AMX_EBE_B(24, KTT, VTT); break;                                                                       \
    case 30: AMX_EBE_B(30, KTT, VTT); break;                                                                       \
    case 60: AMX_EBE_B(60, KTT, VTT); break;                                                                       \
    default: fail(AMX_NOT_IMPLEMENTED, "batched launches need a tile kernel (n = " + std::to_string(n) + ")");   \
    }
    if (vt_double) {
        if (kt == KType::F64) {
            AMX_EBE_N(double, double);
        } else if (kt == KType::F32) {
            AMX_EBE_N(float, double);
        } else {
            AMX_EBE_N(uint16_t, double);
        }
    } else if (kt == KType::F32) {
        AMX_EBE_N(float, float);
    } else {
        AMX_EBE_N(uint16_t, float);
    }
#undef AMX_EBE_N
#undef AMX_EBE_B
    AMX_HIP(hipGetLastError());
}

void launch_gather_batched(bool vt_double, const GatherJob* jobs, const uint32_t* prefix, uint32_t n_jobs,
                           uint32_t total_blocks, hipStream_t s)
{
    if (total_blocks == 0 || n_jobs == 0) return;
    if (vt_double) gather_batched_kernel<double><<<total_blocks, 256, 0, s>>>(jobs, prefix, n_jobs);
    else gather_batched_kernel<float><<<total_blocks, 256, 0, s>>>(jobs, prefix, n_jobs);
    AMX_HIP(hipGetLastError());
}

bool batched_n(uint32_t n) { return has_tile_kernel(n); }

void launch_bc_fixup(bool vt_double, int bc, const uint32_t* cdofs, uint32_t nc, const double* coef, const void* u,
                     void* v, hipStream_t s)
{
    if (nc == 0 || bc == 0) return;
    const uint32_t nb = blocks_for(nc, 256);
    if (vt_double) {
        if (bc == 1) bc_fixup_kernel<double, 1><<<nb, 256, 0, s>>>(cdofs, nc, coef, (const double*)u, (double*)v);
        else bc_fixup_kernel<double, 2><<<nb, 256, 0, s>>>(cdofs, nc, coef, (const double*)u, (double*)v);
    } else {
        if (bc == 1) bc_fixup_kernel<float, 1><<<nb, 256, 0, s>>>(cdofs, nc, coef, (const float*)u, (float*)v);
        else bc_fixup_kernel<float, 2><<<nb, 256, 0, s>>>(cdofs, nc, coef, (const float*)u, (float*)v);
    }
    AMX_HIP(hipGetLastError());
}

void launch_fill(void* p, uint64_t n, bool vt_double, double value, hipStream_t s)
{
    if (n == 0) return;
    if (vt_double) fill_kernel<double><<<blocks_for(n, 256), 256, 0, s>>>((double*)p, n, value);
    else fill_kernel<float><<<blocks_for(n, 256), 256, 0, s>>>((float*)p, n, (float)value);
    AMX_HIP(hipGetLastError());
}

void launch_coef(const double* diag, const uint32_t* cdofs, uint32_t nc, double penalty, double* coef64, hipStream_t s)
{
    if (nc == 0) return;
    coef_kernel<<<blocks_for(nc, 256), 256, 0, s>>>(diag, cdofs, nc, penalty, coef64);
    AMX_HIP(hipGetLastError());
}

void launch_set_entries(double* u, const uint32_t* ids, uint32_t n, double value, hipStream_t s)
{
    if (n == 0) return;
    set_entries_kernel<<<blocks_for(n, 256), 256, 0, s>>>(u, ids, n, value);
    AMX_HIP(hipGetLastError());
}

void launch_pack(double* dst, const double* src, const uint32_t* ids, uint32_t n, hipStream_t s)
{
    if (n == 0) return;
    pack_kernel<<<blocks_for(n, 256), 256, 0, s>>>(dst, src, ids, n);
    AMX_HIP(hipGetLastError());
}

void launch_halo_accumulate(double* v, const double* recv, const uint32_t* dst, const uint32_t* off,
                            const uint32_t* src, uint32_t n_unique, hipStream_t s)
{
    if (n_unique == 0) return;
    halo_accumulate_kernel<<<blocks_for(n_unique, 256), 256, 0, s>>>(v, recv, dst, off, src, n_unique);
    AMX_HIP(hipGetLastError());
}

void launch_d2f(const double* a, float* b, uint64_t n, hipStream_t s)
{
    if (n == 0) return;
    d2f_kernel<<<blocks_for(n, 256), 256, 0, s>>>(a, b, n);
    AMX_HIP(hipGetLastError());
}

void launch_f2d(const float* a, double* b, uint64_t n, hipStream_t s)
{
    if (n == 0) return;
    f2d_kernel<<<blocks_for(n, 256), 256, 0, s>>>(a, b, n);
    AMX_HIP(hipGetLastError());
}

}  // namespace amx

// ---------------------------------------------------------- measurement aid
namespace amx {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// read-only stream: 16-byte non-temporal loads, grid-stride; the xor keeps
// the loads live and the (practically never taken) store keeps the xor
__global__ __launch_bounds__(256) void read_stream_kernel(const u32x4* __restrict__ p, uint64_t n16,
                                                         uint32_t* __restrict__ out)
{
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256u) {
        const u32x4 v = __builtin_nontemporal_load(p + i);
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9e3779b9u) out[blockIdx.x] = acc;
}

}  // namespace
}  // namespace amx

extern "C" int amx_measure_read_bandwidth(int device, uint64_t bytes, int iters, double* bytes_per_s)
{
    return amx::guarded(__func__, [&] {
        AMX_REQUIRE(bytes_per_s && iters > 0 && bytes >= (1u << 20), AMX_INVALID_ARGUMENT,
                    "need bytes >= 1 MiB, iters > 0 and an output");
        int ndev = 0;
        AMX_HIP(hipGetDeviceCount(&ndev));
        AMX_REQUIRE(device >= 0 && device < ndev, AMX_INVALID_ARGUMENT, "device not present");
        int prev = -1;
        AMX_HIP(hipGetDevice(&prev));
        AMX_HIP(hipSetDevice(device));
        struct Restore {
            int d;
            ~Restore() { if (d >= 0) (void)hipSetDevice(d); }
        } restore{prev};
        const uint64_t n16 = bytes / 16;
        amx::DevBuf buf, out;
        buf.alloc(n16 * 16);
        int cus = 0;
        AMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        const uint32_t grid = (uint32_t)std::max(1, cus) * 8u;
        out.alloc((size_t)grid * sizeof(uint32_t));
        AMX_HIP(hipMemset(buf.p, 0, buf.bytes));
        hipStream_t s = nullptr;
        AMX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        AMX_HIP(hipEventCreate(&e0));
        AMX_HIP(hipEventCreate(&e1));
        amx::read_stream_kernel<<<grid, 256, 0, s>>>(buf.as<amx::u32x4>(), n16, out.as<uint32_t>());  // warm-up
        AMX_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < iters; i++)
            amx::read_stream_kernel<<<grid, 256, 0, s>>>(buf.as<amx::u32x4>(), n16, out.as<uint32_t>());
        AMX_HIP(hipEventRecord(e1, s));
        AMX_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        AMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        *bytes_per_s = (double)n16 * 16.0 * iters / (ms * 1e-3);
    });
}
