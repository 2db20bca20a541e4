// amx_mrhs.hip -- multi-RHS element products on MFMA: V = sum_e P_e^T K_e P_e U
// for U, V of nrhs columns (DoF-major, columns contiguous), one persistent
// launch over several operators (meshes of different element types).
//
// A wave owns one element tile x one RHS tile (32 * NT columns).  Elements
// with n > 32 DoFs are one tile of ceil(n/32) 32-row M-tiles; smaller ones are
// packed block-diagonally, G = 32/n elements per 32-row tile (4 hex8), whose
// off-diagonal zeros are produced in registers, never stored.
//   FP32: v_mfma_f32_32x32x2_f32, A fragment lane l = A[l%32][k0 + l/32]
//         read from column-major K_e (contiguous rows for fixed k);
//   BF16: v_mfma_f32_32x32x16_bf16, A fragment = 8 consecutive columns of one
//         row, read as one 16-byte load from row-major K_e (row stride
//         rounded up to 8); U is rounded to BF16 on load.
// C fragment (both): lane l, register i -> row 8(i/4) + 4(l/32) + i%4, column l%32.
// The element rows are added into V with FP32 atomics (order not fixed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "amx_op.hpp"

using amx::DevBuf;
using amx::fail;
using amx::guarded;

namespace amx {
namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kMaxSeg = 8;

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

struct Seg {
    const void* K;          // FP32: [elem][c][r]; BF16: [elem][r][ld]
    const uint32_t* dofs;   // [tile][KD] local DoF ids (kNone: padding)
    const float* U;
    float* V;
    uint32_t n, n_elem, n_tiles, ngroups, nrhs, work_begin;
};

struct SegTable {
    Seg s[kMaxSeg];
    uint32_t n_seg, total_work;
};

__host__ __device__ constexpr int tile_elems(int n) { return n > 32 ? 1 : 32 / n; }
__host__ __device__ constexpr int tile_rows(int n) { return n > 32 ? n : (32 / n) * n; }
__host__ __device__ constexpr int mtiles(int n) { return (tile_rows(n) + 31) / 32; }
// K dimension of a tile padded to the MFMA step (2 for FP32, 16 for BF16)
__host__ __device__ constexpr int kdim(int n, bool bf16) { return bf16 ? (tile_rows(n) + 15) / 16 * 16 : (tile_rows(n) + 1) / 2 * 2; }
__host__ __device__ constexpr int bf16_ld(int n) { return (n + 7) / 8 * 8; }

__device__ __forceinline__ uint32_t rl(uint32_t x, int lane) { return (uint32_t)__builtin_amdgcn_readlane((int)x, lane); }

__device__ __forceinline__ unsigned short f2bf(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (unsigned short)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

// one (element tile, RHS tile) of a segment
template <int N, bool BF16, int NT>
__device__ __forceinline__ void tile_product(const Seg& sg, uint32_t t, uint32_t ng)
{
    constexpr int G = tile_elems(N), ROWS = tile_rows(N), MT = mtiles(N), KD = kdim(N, BF16);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t half = lane >> 5, col_l = lane & 31u;
    // the tile's DoF list, one entry per lane (KD <= 64)
    const uint32_t mydof = lane < (uint32_t)KD ? sg.dofs[(uint64_t)t * KD + lane] : kNone;
    f32x16 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int q = 0; q < NT; q++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[m][q][i] = 0.0f;
    const uint32_t col0 = ng * 32u * NT;
    const uint64_t e0 = (uint64_t)t * G;

    if constexpr (!BF16) {
        const float* K = static_cast<const float*>(sg.K);
#pragma unroll 2
        for (int ks = 0; ks < KD / 2; ks++) {
            const int k = 2 * ks + (int)half;  // this lane's K index
            float a[MT];
#pragma unroll
            for (int m = 0; m < MT; m++) {
                const int row = m * 32 + (int)col_l;
                float v = 0.0f;
                if (row < ROWS && k < ROWS) {
                    const int g = (N > 32) ? 0 : row / N, r = (N > 32) ? row : row % N;
                    const int gk = (N > 32) ? 0 : k / N, c = (N > 32) ? k : k % N;
                    if (g == gk && e0 + g < sg.n_elem) v = K[(e0 + g) * N * N + (uint64_t)c * N + r];
                }
                a[m] = v;
            }
            const uint32_t d0 = rl(mydof, 2 * ks), d1 = rl(mydof, 2 * ks + 1);
            const uint32_t d = half ? d1 : d0;
#pragma unroll
            for (int q = 0; q < NT; q++) {
                const uint32_t col = col0 + q * 32 + col_l;
                const float b = (d != kNone && col < sg.nrhs) ? sg.U[(uint64_t)d * sg.nrhs + col] : 0.0f;
#pragma unroll
                for (int m = 0; m < MT; m++) acc[m][q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[m], b, acc[m][q], 0, 0, 0);
            }
        }
    } else {
        const unsigned short* K = static_cast<const unsigned short*>(sg.K);
        constexpr int LD = bf16_ld(N);
#pragma unroll
        for (int ks = 0; ks < KD / 16; ks++) {
            const int kb = 16 * ks + 8 * (int)half;  // first of this lane's 8 K indices
            u16x8 a[MT];
#pragma unroll
            for (int m = 0; m < MT; m++) {
                const int row = m * 32 + (int)col_l;
                u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
                if (row < ROWS && kb < ROWS) {
                    const int g = (N > 32) ? 0 : row / N, r = (N > 32) ? row : row % N;
                    if (N > 32) {
                        if (e0 < sg.n_elem) {
                            // 8 columns kb..kb+7 of row r; row stride LD (multiple of 8) keeps them 16-byte aligned
                            v = *reinterpret_cast<const u16x8*>(K + e0 * N * LD + (uint64_t)r * LD + kb);
                        }
                    } else {
                        // block-diagonal: columns kb..kb+7 belong to element kb / N only when N % 8 == 0
#pragma unroll
                        for (int j = 0; j < 8; j++) {
                            const int k = kb + j;
                            if (k < ROWS && k / N == g && e0 + g < sg.n_elem)
                                v[j] = K[(e0 + g) * N * LD + (uint64_t)r * LD + (k % N)];
                        }
                    }
                }
                a[m] = v;
            }
#pragma unroll
            for (int q = 0; q < NT; q++) {
                const uint32_t col = col0 + q * 32 + col_l;
                u16x8 b;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const uint32_t d = __shfl(mydof, kb + j, 64);
                    const float x = (d != kNone && col < sg.nrhs) ? sg.U[(uint64_t)d * sg.nrhs + col] : 0.0f;
                    b[j] = f2bf(x);
                }
#pragma unroll
                for (int m = 0; m < MT; m++)
                    acc[m][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a[m]),
                                                                         __builtin_bit_cast(bf16x8, b), acc[m][q], 0, 0, 0);
            }
        }
    }
    // add the element rows into V
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int row = m * 32 + 8 * (i / 4) + 4 * (int)half + (i % 4);
            const uint32_t d = __shfl(mydof, row < 64 ? row : 63, 64);
            if (row < ROWS && d != kNone) {
#pragma unroll
                for (int q = 0; q < NT; q++) {
                    const uint32_t col = col0 + q * 32 + col_l;
                    if (col < sg.nrhs) unsafeAtomicAdd(&sg.V[(uint64_t)d * sg.nrhs + col], acc[m][q][i]);
                }
            }
        }
}

template <bool BF16, int NT>
__global__ __launch_bounds__(256) void mrhs_persistent(SegTable tab)
{
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < tab.total_work; w += waves) {
        uint32_t si = 0;
        while (si + 1 < tab.n_seg && w >= tab.s[si + 1].work_begin) si++;
        const Seg& sg = tab.s[si];
        const uint32_t local = w - sg.work_begin;
        const uint32_t t = local / sg.ngroups, ng = local - t * sg.ngroups;
        switch (sg.n) {
        case 8: tile_product<8, BF16, NT>(sg, t, ng); break;
        case 10: tile_product<10, BF16, NT>(sg, t, ng); break;
        case 20: tile_product<20, BF16, NT>(sg, t, ng); break;
        case 60: tile_product<60, BF16, NT>(sg, t, ng); break;
        default: break;
        }
    }
}

__global__ void zero_f32(float* p, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// per-element matrices for the MFMA kernels from the tiled FP64 storage
__global__ void mrhs_layout_kernel(const double* __restrict__ K64, uint32_t N, uint32_t G, uint32_t P,
                                   const uint32_t* __restrict__ slot_of, uint32_t n_elem, int bf16, uint32_t ld,
                                   void* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per = bf16 ? (uint64_t)N * ld : (uint64_t)N * N;
    if (i >= per * n_elem) return;
    const uint64_t e = i / per;
    const uint32_t w = (uint32_t)(i - e * per);
    uint32_t r, c;
    if (bf16) { r = w / ld; c = w % ld; } else { c = w / N; r = w % N; }
    float val = 0.0f;
    if (c < N) {
        const uint64_t slot = slot_of[e];
        const uint64_t t = slot / G;
        const uint32_t g = (uint32_t)(slot - t * G), GN = G * N, NP = N / P;
        val = (float)K64[((t * NP + c / P) * GN + g * N + r) * P + c % P];
    }
    if (bf16) static_cast<unsigned short*>(out)[i] = f2bf(val);
    else static_cast<float*>(out)[i] = val;
}

}  // namespace
}  // namespace amx

struct amx_mrhs {
    int device = 0;
    bool bf16 = false;
    int rhs_tile = 32, wpb = 4;  // NT * 32 columns per wave, waves per block
    struct Part {
        uint32_t n, n_elem, n_tiles, n_total;
        DevBuf K, dofs;
    };
    std::vector<Part> parts;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int cu_count = 256;
    ~amx_mrhs()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace amx {
namespace {

bool supported_n(uint32_t n) { return n == 8 || n == 10 || n == 20 || n == 60; }

void launch(amx_mrhs& m, int nrhs, const float* const* U, float* const* V, hipStream_t s)
{
    SegTable tab{};
    uint32_t work = 0;
    AMX_REQUIRE(m.parts.size() <= (size_t)kMaxSeg, AMX_NOT_IMPLEMENTED, "too many element groups in one launch");
    const int NT = m.rhs_tile / 32;
    const uint32_t ngroups = (uint32_t)((nrhs + m.rhs_tile - 1) / m.rhs_tile);
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
        sg.work_begin = work;
        work += p.n_tiles * ngroups;
        const uint64_t len = (uint64_t)p.n_total * nrhs;
        if (len) zero_f32<<<(uint32_t)((len + 255) / 256), 256, 0, s>>>(V[i], len);
    }
    tab.n_seg = (uint32_t)m.parts.size();
    tab.total_work = work;
    if (work == 0) return;
    const uint32_t waves_needed = work;
    uint32_t blocks = std::min<uint32_t>((waves_needed + m.wpb - 1) / m.wpb, (uint32_t)m.cu_count * 8);
    const dim3 grid(blocks), block(64 * m.wpb);
#define AMX_MRHS_LAUNCH(B, T) mrhs_persistent<B, T><<<grid, block, 0, s>>>(tab)
    if (m.bf16) {
        if (NT == 1) AMX_MRHS_LAUNCH(true, 1);
        else if (NT == 2) AMX_MRHS_LAUNCH(true, 2);
        else AMX_MRHS_LAUNCH(true, 4);
    } else {
        if (NT == 1) AMX_MRHS_LAUNCH(false, 1);
        else if (NT == 2) AMX_MRHS_LAUNCH(false, 2);
        else AMX_MRHS_LAUNCH(false, 4);
    }
#undef AMX_MRHS_LAUNCH
    AMX_HIP(hipGetLastError());
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
            p.n_tiles = (p.n_elem + G - 1) / G;
            std::vector<uint32_t> dofs((size_t)p.n_tiles * KD, amx::kNone);
            for (uint32_t t = 0; t < p.n_tiles; t++)
                for (int k = 0; k < KD && k < amx::tile_rows((int)g.n); k++) {
                    const uint32_t j = t * G + (g.n > 32 ? 0 : k / g.n);
                    if (j >= p.n_elem) continue;
                    const uint32_t e = g.elems[g.slot_elem[slot_of[j]]];
                    dofs[(size_t)t * KD + k] = op.local_map[op.eoff[e] + (g.n > 32 ? k : k % g.n)];
                }
            DevBuf dslot;
            dslot.alloc(std::max<size_t>(slot_of.size(), 1) * 4);
            AMX_HIP(hipMemcpy(dslot.p, slot_of.data(), slot_of.size() * 4, hipMemcpyHostToDevice));
            p.dofs.alloc(std::max<size_t>(dofs.size(), 1) * 4);
            AMX_HIP(hipMemcpy(p.dofs.p, dofs.data(), dofs.size() * 4, hipMemcpyHostToDevice));
            const uint32_t ld = m->bf16 ? (uint32_t)amx::bf16_ld((int)g.n) : g.n;
            const uint64_t total = (uint64_t)p.n_elem * g.n * ld;
            p.K.alloc(std::max<uint64_t>(total, 8) * (m->bf16 ? 2 : 4));
            if (total)
                amx::mrhs_layout_kernel<<<(uint32_t)((total + 255) / 256), 256, 0, op.stream>>>(
                    g.K[0].as<double>(), g.n, g.L[0].G, g.L[0].P, dslot.as<uint32_t>(), p.n_elem, m->bf16 ? 1 : 0, ld,
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

int amx_mrhs_set_tiles(amx_mrhs* m, int rhs_tile, int waves_per_block)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m, AMX_INVALID_ARGUMENT, "NULL handle");
        AMX_REQUIRE(rhs_tile == 32 || rhs_tile == 64 || rhs_tile == 128, AMX_INVALID_ARGUMENT, "rhs_tile: 32, 64 or 128");
        AMX_REQUIRE(waves_per_block == 1 || waves_per_block == 2 || waves_per_block == 4, AMX_INVALID_ARGUMENT,
                    "waves_per_block: 1, 2 or 4");
        m->rhs_tile = rhs_tile;
        m->wpb = waves_per_block;
    });
}

int amx_mrhs_get_tiles(const amx_mrhs* m, int* rhs_tile, int* waves_per_block)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m, AMX_INVALID_ARGUMENT, "NULL handle");
        if (rhs_tile) *rhs_tile = m->rhs_tile;
        if (waves_per_block) *waves_per_block = m->wpb;
    });
}

int amx_mrhs_apply_dev(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m && U && V && nrhs > 0, AMX_INVALID_ARGUMENT, "bad argument");
        amx::DeviceGuard dg(m->device);
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
        hipEvent_t e0, e1;
        AMX_HIP(hipEventCreate(&e0));
        AMX_HIP(hipEventCreate(&e1));
        double best = 1e30;
        int brt = 32, bw = 4, idx = 0;
        for (int rt : {32, 64, 128}) {
            for (int w : {1, 2, 4}) {
                double ms = 1e30;
                if (rt <= std::max(32, ((nrhs + 31) / 32) * 32)) {
                    m->rhs_tile = rt;
                    m->wpb = w;
                    amx::launch(*m, nrhs, U, V, s);  // warm-up
                    AMX_HIP(hipEventRecord(e0, s));
                    for (int r = 0; r < reps; r++) amx::launch(*m, nrhs, U, V, s);
                    AMX_HIP(hipEventRecord(e1, s));
                    AMX_HIP(hipEventSynchronize(e1));
                    float f = 0;
                    AMX_HIP(hipEventElapsedTime(&f, e0, e1));
                    ms = f / reps;
                    if (ms < best) { best = ms; brt = rt; bw = w; }
                }
                if (results) results[idx] = ms;  // order: rhs_tile 32/64/128 x waves 1/2/4
                idx++;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        m->rhs_tile = brt;
        m->wpb = bw;
    });
}

int amx_mrhs_graph_capture(amx_mrhs* m, int nrhs, const float* const* U, float* const* V, void* stream)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(m && U && V && nrhs > 0, AMX_INVALID_ARGUMENT, "bad argument");
        AMX_REQUIRE(stream != nullptr, AMX_INVALID_ARGUMENT, "capture needs a non-default stream");
        amx::DeviceGuard dg(m->device);
        hipStream_t s = static_cast<hipStream_t>(stream);
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
