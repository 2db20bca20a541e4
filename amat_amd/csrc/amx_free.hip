// amx_free.hip -- method 2 (FREE): element contributions computed from node
// coordinates at every application; no element matrices are stored.
//
// The reference's FREE method calls a user function that builds K_e from the
// element's coordinates before each element GEMV (aMat::set_element_matrix_
// function, include/aMat.hpp:114-116, called from aMatFree.hpp:2449-2476;
// the examples pass ke_hex20_iso through computeElemMat,
// examples/src/example07a.cpp:54-70).  Here the element kinds of the examples
// are built in and applied without forming K_e, with the quadrature, shape
// functions and constitutive matrix of ke_hex8 / ke_tet10 / ke_hex20_iso
// (examples/src/ke_matrix.cpp), so f = s_e K_e u_e up to rounding:
//
//   J[k][d] = sum_a x_ak dN_ad,  inv = J^-1 (inv[d][k] = dxi_d/dx_k),
//   w = |J| w_q s_e,  dN_ad = dN_a/dxi_d at the quadrature point.
//   Laplace (one DoF per node):
//     h_d = sum_a u_a dN_ad,  p_d = w sum_k inv[d][k] (sum_d' h_d' inv[d'][k]),
//     f_a += sum_d dN_ad p_d.
//   Isotropic elasticity (DoF 3a+i, Voigt xx,yy,zz,xy,yz,xz, engineering shear):
//     H_id = sum_a u_ai dN_ad,  grad u = H inv,  sigma = w C eps(u),
//     P_di = sum_k inv[d][k] sigma_ik,  f_ai += sum_d dN_ad P_di.
//
// Outputs go to the element rows of the deterministic scatter (or atomics), so
// the maps, BC forms and gather are those of the stored-matrix kernels.  With
// DIAG the kernels write diag(s_e K_e) instead (Jacobi, penalty, trace).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <type_traits>

#include "amx_internal.hpp"
#include "amx_kernels.hpp"

namespace amx {

namespace {

constexpr uint32_t kFlag = 0x80000000u, kMask = 0x7fffffffu;
constexpr int kMaxDevices = 64;

__device__ __forceinline__ void lds_sync_wave()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// two-value LDS accesses of the compute type RT (FP64 or FP32)
template <typename RT> struct Vec2;
template <> struct Vec2<double> {
    using type = double2;
    __device__ static type make(double a, double b) { return make_double2(a, b); }
};
template <> struct Vec2<float> {
    using type = float2;
    __device__ static type make(float a, float b) { return make_float2(a, b); }
};

// det J and inv[d][k] = cofactor(J)[k][d] / det (cofactors as in the host
// element routines, amx_fem.cpp jacobian; one division)
template <typename RT> __device__ __forceinline__ RT jac_inv_r(const RT J[3][3], RT inv[3][3])
{
    const RT c00 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
    const RT c10 = J[2][1] * J[0][2] - J[0][1] * J[2][2];
    const RT c20 = J[0][1] * J[1][2] - J[1][1] * J[0][2];
    const RT det = J[0][0] * c00 + J[1][0] * c10 + J[2][0] * c20;
    const RT r = RT(1) / det;
    inv[0][0] = c00 * r;
    inv[0][1] = c10 * r;
    inv[0][2] = c20 * r;
    inv[1][0] = (J[1][2] * J[2][0] - J[2][2] * J[1][0]) * r;
    inv[1][1] = (J[2][2] * J[0][0] - J[0][2] * J[2][0]) * r;
    inv[1][2] = (J[0][2] * J[1][0] - J[1][2] * J[0][0]) * r;
    inv[2][0] = (J[2][1] * J[1][0] - J[1][1] * J[2][0]) * r;
    inv[2][1] = (J[0][1] * J[2][0] - J[2][1] * J[0][0]) * r;
    inv[2][2] = (J[1][1] * J[0][0] - J[0][1] * J[1][0]) * r;
    return det;
}

// ---- Laplace elements: one lane per element slot ------------------------
// xe: [NN*3][stride] (coordinate-major, slot fastest), qp: [nq][NN*3+1]
// (dN_ad at 3a+d, weight last; uniform across lanes -> scalar loads).
// RT: compute, coordinate and vector type (FP64; FP32 for the inner solves of
// the mixed-precision solvers, reading an FP32 copy of the coordinates).
// Factors and the table stay FP64 in memory and are rounded on load.
template <int NN, typename RT, bool MASKED, bool ATOMIC, bool DIAG>
__global__ __launch_bounds__(256) void free_laplace_kernel(const RT* __restrict__ xe, const double* __restrict__ scale,
                                                          const uint32_t* __restrict__ map, const RT* __restrict__ u,
                                                          RT* __restrict__ out, uint64_t ve_base, uint32_t rs,
                                                          uint32_t ss, uint32_t slot_begin, uint32_t n_slots,
                                                          uint32_t stride, const double* __restrict__ qp, uint32_t nq,
                                                          const uint32_t* __restrict__ skip)
{
    if (skip && *skip) return;
    const uint32_t sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= n_slots) return;
    const uint64_t slot = (uint64_t)slot_begin + sl;
    RT x[NN * 3], ue[NN], f[NN];
    uint32_t m[NN];
#pragma unroll
    for (int i = 0; i < NN * 3; i++) x[i] = (RT)xe[(uint64_t)i * stride + slot];
#pragma unroll
    for (int a = 0; a < NN; a++) {
        m[a] = map[slot * NN + a];
        ue[a] = RT(0);
        if constexpr (!DIAG) {
            const RT v = u[m[a] & kMask];
            ue[a] = (MASKED && (m[a] & kFlag)) ? RT(0) : v;
        }
        f[a] = RT(0);
    }
    const RT sc = (RT)scale[slot];
    for (uint32_t q = 0; q < nq; q++) {
        const double* t = qp + (uint64_t)q * (NN * 3 + 1);
        RT J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, h[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const RT dn = (RT)t[3 * a + d];
#pragma unroll
                for (int k = 0; k < 3; k++) J[k][d] += x[3 * a + k] * dn;
                if constexpr (!DIAG) h[d] += ue[a] * dn;
            }
        RT inv[3][3];
        const RT det = jac_inv_r(J, inv);
        RT w = det * (RT)t[NN * 3] * sc;
        if constexpr (DIAG) {
            // a non-positive Jacobian at any point marks the element (NaN diagonal);
            // finalize rejects it, as the element routines do
            if (!(det > RT(0))) w = (RT)__builtin_nan("");
#pragma unroll
            for (int a = 0; a < NN; a++) {
                RT s = RT(0);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const RT g = (RT)t[3 * a] * inv[0][k] + (RT)t[3 * a + 1] * inv[1][k] + (RT)t[3 * a + 2] * inv[2][k];
                    s += g * g;
                }
                f[a] += s * w;
            }
        } else {
            RT gu[3], p[3];
#pragma unroll
            for (int k = 0; k < 3; k++) gu[k] = (h[0] * inv[0][k] + h[1] * inv[1][k] + h[2] * inv[2][k]) * w;
#pragma unroll
            for (int d = 0; d < 3; d++) p[d] = inv[d][0] * gu[0] + inv[d][1] * gu[1] + inv[d][2] * gu[2];
#pragma unroll
            for (int a = 0; a < NN; a++) f[a] += (RT)t[3 * a] * p[0] + (RT)t[3 * a + 1] * p[1] + (RT)t[3 * a + 2] * p[2];
        }
    }
#pragma unroll
    for (int a = 0; a < NN; a++) {
        if constexpr (ATOMIC) unsafeAtomicAdd(&out[m[a] & kMask], f[a]);
        else out[ve_base + (uint64_t)a * rs + slot * ss] = f[a];
    }
}

// ---- hex20 isotropic elasticity -----------------------------------------
// A wave takes EW = 64/QP element slots at a time; in phase 1 lane
// l = e*QP + q holds element e at quadrature point q (QP: the point count
// padded to 8, 32 or 64; padding points carry weight 0) and writes P (or inv
// and w for DIAG) to an LDS slot; in phase 2 lane o = 3a+i < 60 sums the
// output o of each element over the points.  xe: [slot][60], qp: [QP][61].
// The grid is persistent: a block loads the dN table into LDS once and its
// waves loop over element groups.  LDS strides avoid bank conflicts of the
// 16-byte reads: dN rows [q][a][4] at kTabRow doubles (2*kTabRow = 4 * odd
// dwords, so 8 rows hit 8 distinct bank groups), node values [e][a][x y z
// u_x u_y u_z] at kElemRow doubles per element (likewise); the P slots
// overlay the node values once phase 1 has read them.
constexpr int kH20Slot = 12;   // S[4i+d] = P_di (DIAG: S[4k+d] = inv[d][k]), S[3] = w
// LDS strides per compute type: FP64 reads two values per 16-byte access
// (dN rows of 82, node values x y z u_x u_y u_z at 6 per node, 122 per
// element); FP32 reads a whole dN row quadruple, a P row or a node's
// x y z u_x in one 16-byte access (rows of 84, 8 values per node: x y z u_x
// u_y u_z and 2 pad, 164 per element)
template <typename RT> struct H20Lds;
template <> struct H20Lds<double> { static constexpr int tab_row = 82, elem_row = 122, node = 6; };
template <> struct H20Lds<float> { static constexpr int tab_row = 84, elem_row = 164, node = 8; };

template <int QP, typename RT> constexpr int h20_wave_values()
{
    return (64 / QP) * H20Lds<RT>::elem_row > 64 * kH20Slot ? (64 / QP) * H20Lds<RT>::elem_row : 64 * kH20Slot;
}

template <int QP, typename RT, bool MASKED, bool ATOMIC, bool DIAG>
__global__ __launch_bounds__(256) void free_hex20_kernel(const RT* __restrict__ xe, const double* __restrict__ scale,
                                                        const uint32_t* __restrict__ map, const RT* __restrict__ u,
                                                        RT* __restrict__ out, uint64_t ve_base, uint32_t rs,
                                                        uint32_t ss, uint32_t slot_begin, uint32_t n_slots,
                                                        const double* __restrict__ qp, uint32_t nq,
                                                        double c1, double c2, double c3,
                                                        const uint32_t* __restrict__ skip)
{
    using RT2 = typename Vec2<RT>::type;
    constexpr int EW = 64 / QP, WD = h20_wave_values<QP, RT>();
    constexpr int kTabRow = H20Lds<RT>::tab_row, kElemRow = H20Lds<RT>::elem_row, NS = H20Lds<RT>::node;
    constexpr bool F32 = std::is_same<RT, float>::value;
    __shared__ __align__(16) RT tab[QP * kTabRow];
    __shared__ RT wq[QP];
    __shared__ __align__(16) RT work[4][WD];
    const RT k1 = (RT)c1, k2 = (RT)c2, k3 = (RT)c3;
    if (skip && *skip) return;
    for (uint32_t i = threadIdx.x; i < QP * 80u; i += blockDim.x) {
        const uint32_t q = i / 80u, r = i - q * 80u, a = r >> 2, d = r & 3u;
        tab[q * kTabRow + r] = d < 3 ? (RT)qp[q * 61u + 3u * a + d] : RT(0);
    }
    for (uint32_t q = threadIdx.x; q < QP; q += blockDim.x) wq[q] = (RT)qp[q * 61u + 60u];
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    RT* W = work[wv];
    const uint32_t n_groups = (n_slots + EW - 1) / EW;
    const uint32_t pa = lane < 60 ? lane / 3u : 19u, pi = lane < 60 ? lane - 3u * pa : 0u;

    // the next group's coordinates, map entries, element values and factor are
    // loaded into registers while the current group is computed: lane value j
    // is entry idx = lane + 64 j of the group's EW x 60 block
    constexpr int NL = (EW * 60 + 63) / 64;
    RT px[NL], pu[NL], ps = RT(0);
    uint32_t pm[NL];
    auto issue = [&](uint32_t grp) {
        const uint32_t first = grp * EW, n_here = min((uint32_t)EW, n_slots - first);
        const uint64_t slot0 = (uint64_t)slot_begin + first;
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const uint32_t idx = lane + 64u * j, e = idx / 60u;
            const bool ok = idx < EW * 60u && e < n_here;
            // branch-free: entries outside the group load the group's first entry
            const uint64_t src = (slot0 + (ok ? e : 0u)) * 60u + (ok ? idx - 60u * e : 0u);
            const RT xv = (RT)xe[src];
            const uint32_t mv = DIAG ? 0u : map[src];
            px[j] = ok ? xv : RT(0);
            pm[j] = ok ? mv : 0u;
        }
        const uint32_t e = lane / QP;
        const RT sv = (RT)scale[slot0 + min(e, n_here - 1u)];
        ps = e < n_here ? sv : RT(0);
    };
    auto gather = [&]() {
#pragma unroll
        for (int j = 0; j < NL; j++) {
            RT v = RT(0);
            if constexpr (!DIAG) {
                v = u[pm[j] & kMask];  // entries outside the group read u[0]; their elements are not stored
                if (MASKED && (pm[j] & kFlag)) v = RT(0);
            }
            pu[j] = v;
        }
    };

    uint32_t grp = blockIdx.x * 4u + wv;  // uniform per wave
    if (grp < n_groups) {
        issue(grp);
        gather();
    }
    for (; grp < n_groups; grp += gridDim.x * 4u) {
        const uint32_t first = grp * EW;
        const uint32_t n_here = min((uint32_t)EW, n_slots - first);
        const uint64_t slot0 = (uint64_t)slot_begin + first;
        const uint32_t next = grp + gridDim.x * 4u;

        // element e's node a at W[e*kElemRow + NS a]: x y z, then u
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const uint32_t idx = lane + 64u * j;
            if (idx < EW * 60u) {
                const uint32_t e = idx / 60u, r = idx - e * 60u, a = r / 3u, k = r - 3u * a;
                W[e * kElemRow + NS * a + k] = px[j];
                W[e * kElemRow + NS * a + 3u + k] = pu[j];
            }
        }
        const RT sc_here = ps;
        lds_sync_wave();
        if (next < n_groups) issue(next);

        // phase 1
        {
            const uint32_t e = lane / QP, q = lane % QP;
            const RT* XU = W + e * kElemRow;
            const RT* T = tab + q * kTabRow;
            RT J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 4
            for (int a = 0; a < 20; a++) {
                RT dn[3], xs[3], us[3];
                if constexpr (F32) {
                    const float4 t4 = *reinterpret_cast<const float4*>(T + 4 * a);
                    const float4 x4 = *reinterpret_cast<const float4*>(XU + 8 * a);
                    dn[0] = t4.x, dn[1] = t4.y, dn[2] = t4.z;
                    xs[0] = x4.x, xs[1] = x4.y, xs[2] = x4.z;
                    us[0] = x4.w;
                    if constexpr (!DIAG) {
                        const float2 u2 = *reinterpret_cast<const float2*>(XU + 8 * a + 4);
                        us[1] = u2.x, us[2] = u2.y;
                    }
                } else {
                    const RT2 t01 = *reinterpret_cast<const RT2*>(T + 4 * a);
                    const RT2 p0 = *reinterpret_cast<const RT2*>(XU + 6 * a);
                    const RT2 p1 = *reinterpret_cast<const RT2*>(XU + 6 * a + 2);
                    dn[0] = t01.x, dn[1] = t01.y, dn[2] = T[4 * a + 2];
                    xs[0] = p0.x, xs[1] = p0.y, xs[2] = p1.x;
                    us[0] = p1.y;
                    if constexpr (!DIAG) {
                        const RT2 p2 = *reinterpret_cast<const RT2*>(XU + 6 * a + 4);
                        us[1] = p2.x, us[2] = p2.y;
                    }
                }
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int d = 0; d < 3; d++) J[k][d] += xs[k] * dn[d];
                if constexpr (!DIAG) {
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int d = 0; d < 3; d++) H[k][d] += us[k] * dn[d];
                }
            }
            RT inv[3][3];
            const RT det = jac_inv_r(J, inv);
            RT w = e < n_here ? det * wq[q] * sc_here : RT(0);
            // DIAG: a non-positive Jacobian at a real point marks the element (NaN diagonal)
            if (DIAG && e < n_here && q < nq && !(det > RT(0))) w = (RT)__builtin_nan("");
            RT S[kH20Slot];
            if constexpr (DIAG) {
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int d = 0; d < 3; d++) S[4 * k + d] = inv[d][k];
            } else {
                RT g[3][3];  // g[i][k] = du_i/dx_k
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        g[i][k] = H[i][0] * inv[0][k] + H[i][1] * inv[1][k] + H[i][2] * inv[2][k];
                const RT exx = g[0][0] * w, eyy = g[1][1] * w, ezz = g[2][2] * w;
                RT sg[3][3];
                sg[0][0] = k1 * exx + k2 * eyy + k2 * ezz;
                sg[1][1] = k2 * exx + k1 * eyy + k2 * ezz;
                sg[2][2] = k2 * exx + k2 * eyy + k1 * ezz;
                sg[0][1] = sg[1][0] = (k3 * w) * (g[0][1] + g[1][0]);
                sg[1][2] = sg[2][1] = (k3 * w) * (g[1][2] + g[2][1]);
                sg[0][2] = sg[2][0] = (k3 * w) * (g[0][2] + g[2][0]);
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int d = 0; d < 3; d++)
                        S[4 * i + d] = inv[d][0] * sg[i][0] + inv[d][1] * sg[i][1] + inv[d][2] * sg[i][2];
            }
            S[3] = w;
            S[7] = S[11] = RT(0);
            lds_sync_wave();  // every lane's reads of the node values precede the slot writes that overlay them
            RT* slotp = W + lane * kH20Slot;
#pragma unroll
            for (int j = 0; j < kH20Slot; j += 2) *reinterpret_cast<RT2*>(slotp + j) = Vec2<RT>::make(S[j], S[j + 1]);
        }
        lds_sync_wave();
        if (next < n_groups) gather();

        // phase 2: f[e] = sum_q dN_q(a) . P_q(:, i) of element e
        RT f[EW];
#pragma unroll
        for (int e = 0; e < EW; e++) f[e] = RT(0);
        auto point = [&](uint32_t e, uint32_t q, RT d0, RT d1, RT d2) {
            const RT* S = W + (e * QP + q) * kH20Slot;
            if constexpr (DIAG) {
                RT g[3];
#pragma unroll
                for (int k = 0; k < 3; k++) g[k] = d0 * S[4 * k] + d1 * S[4 * k + 1] + d2 * S[4 * k + 2];
                const RT gi = pi == 0 ? g[0] : (pi == 1 ? g[1] : g[2]);
                const RT gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
                f[e] += (k1 * gi * gi + k3 * (gg - gi * gi)) * S[3];
            } else {
                if constexpr (F32) {
                    const float4 p4 = *reinterpret_cast<const float4*>(S + 4 * pi);
                    f[e] += d0 * p4.x + d1 * p4.y + d2 * p4.z;
                } else {
                    const RT2 p01 = *reinterpret_cast<const RT2*>(S + 4 * pi);
                    const RT p2 = S[4 * pi + 2];
                    f[e] += d0 * p01.x + d1 * p01.y + d2 * p2;
                }
            }
        };
        for (uint32_t q = 0; q < nq; q++) {
            RT t0, t1, t2;
            if constexpr (F32) {
                const float4 t4 = *reinterpret_cast<const float4*>(tab + q * kTabRow + 4 * pa);
                t0 = t4.x, t1 = t4.y, t2 = t4.z;
            } else {
                const RT2 t01 = *reinterpret_cast<const RT2*>(tab + q * kTabRow + 4 * pa);
                t0 = t01.x, t1 = t01.y, t2 = tab[q * kTabRow + 4 * pa + 2];
            }
#pragma unroll
            for (uint32_t e = 0; e < EW; e++) point(e, q, t0, t1, t2);
        }
        if (lane < 60) {
#pragma unroll
            for (uint32_t e = 0; e < EW; e++) {
                if (e < n_here) {
                    const uint64_t slot = slot0 + e;
                    if constexpr (ATOMIC) unsafeAtomicAdd(&out[map[slot * 60u + lane] & kMask], f[e]);
                    else out[ve_base + (uint64_t)lane * rs + slot * ss] = f[e];
                }
            }
        }
        lds_sync_wave();  // phase-2 reads of the slots precede the next group's node writes
    }
}

// coordinates in the compute type: FP64, or the FP32 copy built at finalize
template <typename RT> const RT* coords(const FreeArgs& a)
{
    if constexpr (std::is_same<RT, float>::value) {
        AMX_REQUIRE(a.xe32 != nullptr, AMX_INVALID_STATE, "no FP32 coordinates");
        return a.xe32;
    } else {
        return a.xe;
    }
}

template <int QP, typename RT>
void launch_hex20(bool masked, bool atomic, bool diag, const FreeArgs& a, hipStream_t s)
{
    constexpr uint32_t per_block = 4u * (64u / QP);
    // resident blocks per device: CUs x blocks per CU (the LDS bounds it); a
    // benign race between threads computes the same value
    static std::atomic<int> resident[kMaxDevices];
    int dev = 0;
    AMX_HIP(hipGetDevice(&dev));
    AMX_REQUIRE(dev >= 0 && dev < kMaxDevices, AMX_NOT_IMPLEMENTED, "device ordinal beyond the supported range");
    int res = resident[dev].load(std::memory_order_relaxed);
    if (res == 0) {
        int cus = 0, per_cu = 0;
        AMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        AMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, free_hex20_kernel<QP, RT, false, false, false>,
                                                             256, 0));
        res = std::max(1, cus * std::max(per_cu, 1));
        resident[dev].store(res, std::memory_order_relaxed);
    }
    const uint32_t need = (a.n_slots + per_block - 1) / per_block;
    const dim3 grid(std::min<uint32_t>(need, (uint32_t)res)), block(256);
    const RT* x = coords<RT>(a);
    const RT* u = static_cast<const RT*>(a.u);
    RT* out = static_cast<RT*>(a.out);
#define AMX_H20(M, A, D)                                                                                           \
    free_hex20_kernel<QP, RT, M, A, D><<<grid, block, 0, s>>>(x, a.scale, a.map, u, out, a.ve_base, a.ve_rs,     \
                                                              a.ve_ss, a.slot_begin, a.n_slots, a.qp, a.nq, a.c1,   \
                                                              a.c2, a.c3, a.skip)
    if constexpr (std::is_same<RT, double>::value) {
        if (diag) {
            AMX_H20(false, false, true);
            return;
        }
    }
    if (masked) { if (atomic) AMX_H20(true, true, false); else AMX_H20(true, false, false); }
    else { if (atomic) AMX_H20(false, true, false); else AMX_H20(false, false, false); }
#undef AMX_H20
}

template <int NN, typename RT>
void launch_laplace(bool masked, bool atomic, bool diag, const FreeArgs& a, hipStream_t s)
{
    const dim3 grid((a.n_slots + 255) / 256), block(256);
    const RT* x = coords<RT>(a);
    const RT* u = static_cast<const RT*>(a.u);
    RT* out = static_cast<RT*>(a.out);
#define AMX_LAP(M, A, D)                                                                                           \
    free_laplace_kernel<NN, RT, M, A, D><<<grid, block, 0, s>>>(x, a.scale, a.map, u, out, a.ve_base, a.ve_rs,  \
                                                                a.ve_ss, a.slot_begin, a.n_slots, a.stride, a.qp,   \
                                                                a.nq, a.skip)
    if constexpr (std::is_same<RT, double>::value) {
        if (diag) {
            AMX_LAP(false, false, true);
            return;
        }
    }
    if (masked) { if (atomic) AMX_LAP(true, true, false); else AMX_LAP(true, false, false); }
    else { if (atomic) AMX_LAP(false, true, false); else AMX_LAP(false, false, false); }
#undef AMX_LAP
}

template <typename RT>
void launch_free_t(int kind, bool masked, bool atomic, bool diag, const FreeArgs& a, hipStream_t s)
{
    switch (kind) {
    case AMX_ELEM_HEX8_LAPLACE: launch_laplace<8, RT>(masked, atomic, diag, a, s); break;
    case AMX_ELEM_TET10_LAPLACE: launch_laplace<10, RT>(masked, atomic, diag, a, s); break;
    case AMX_ELEM_HEX20_ELASTIC: {
        AMX_REQUIRE(a.nq <= 64, AMX_NOT_IMPLEMENTED, "hex20: at most 64 quadrature points");
        const uint32_t qp = free_padded_points(kind, a.nq);
        if (qp == 8) launch_hex20<8, RT>(masked, atomic, diag, a, s);
        else if (qp == 32) launch_hex20<32, RT>(masked, atomic, diag, a, s);
        else launch_hex20<64, RT>(masked, atomic, diag, a, s);
        break;
    }
    default: fail(AMX_INVALID_ARGUMENT, "unknown element kind " + std::to_string(kind));
    }
}

}  // namespace

uint32_t free_padded_points(int kind, uint32_t nq)
{
    if (kind != AMX_ELEM_HEX20_ELASTIC) return nq;
    return nq <= 8 ? 8 : (nq <= 32 ? 32 : 64);
}

void launch_free(int kind, bool masked, bool atomic, bool diag, const FreeArgs& a, hipStream_t s)
{
    if (a.n_slots == 0) return;
    AMX_REQUIRE(!(diag && a.f32), AMX_INVALID_ARGUMENT, "the FREE diagonal is computed in FP64");
    if (a.f32) launch_free_t<float>(kind, masked, atomic, diag, a, s);
    else launch_free_t<double>(kind, masked, atomic, diag, a, s);
    AMX_HIP(hipGetLastError());
}

}  // namespace amx
