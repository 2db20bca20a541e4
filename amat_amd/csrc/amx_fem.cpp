// amx_fem.cpp -- element matrices and load vectors of the examples' element
// types (hex8 Laplacian, hex20 isotropic elasticity, tet10 Laplacian and
// consistent-mass load), FP64 on the host.
//
// These produce the inputs of the EbE operator.  Node orderings, the
// collapsed-cube quadrature of the tetrahedron and the Gauss-Legendre rule
// follow the reference (examples/include/shapeFunc.hpp,
// examples/src/ke_matrix.cpp, examples/include/integration.hpp) so that the
// same coordinates give the same matrices (checked against fixtures written by
// the reference sources to ~1e-15 relative).
#include <algorithm>
#include <cmath>

#include "amx_internal.hpp"

using amx::guarded;

namespace {

// Gauss-Legendre points/weights on [-1,1] by Newton's method on the Legendre
// recurrence, stopping at |dz| < 3e-14 (examples/include/integration.hpp:81-113).
void gauss_legendre(uint32_t n, double* pw)
{
    const uint32_t m = (n + 1) / 2;
    for (uint32_t i = 1; i <= m; i++) {
        double z = std::cos(M_PI * (i - 0.25) / (n + 0.5)), dz, pn, dpn;
        do {
            double p_prev = 0.0;
            pn = 1.0;
            for (uint32_t j = 1; j <= n; j++) {
                const double p_next = ((2.0 * j - 1.0) * z * pn - (j - 1.0) * p_prev) / j;
                p_prev = pn;
                pn = p_next;
            }
            dpn = n * (z * pn - p_prev) / (z * z - 1.0);
            dz = pn / dpn;
            z -= dz;
        } while (std::fabs(dz) >= 3.0e-14);
        const double w = 2.0 / ((1.0 - z * z) * dpn * dpn);
        pw[2 * (i - 1)] = -z;
        pw[2 * (i - 1) + 1] = w;
        pw[2 * (n - i)] = z;
        pw[2 * (n - i) + 1] = w;
    }
}

// Corner sign table shared by hex8 and the hex20 corners: local node order
// 0..3 bottom face counter-clockwise from (-1,-1,-1), 4..7 the top face
// (examples/include/shapeFunc.hpp:137-146, 200-232).
constexpr int kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// dN/dxi for hex8 at xi: dN[3*a + d]
void dshape_hex8(const double* xi, double* dN)
{
    for (int a = 0; a < 8; a++) {
        const double s0 = kCorner[a][0], s1 = kCorner[a][1], s2 = kCorner[a][2];
        dN[3 * a + 0] = s0 * (1.0 + s1 * xi[1]) * (1.0 + s2 * xi[2]) / 8.0;
        dN[3 * a + 1] = s1 * (1.0 + s0 * xi[0]) * (1.0 + s2 * xi[2]) / 8.0;
        dN[3 * a + 2] = s2 * (1.0 + s0 * xi[0]) * (1.0 + s1 * xi[1]) / 8.0;
    }
}

// hex20 serendipity: corners as hex8; mid-edge nodes 8..19 with the edge
// direction d (the coordinate that is 0 at the node) and the signs of the two
// other coordinates (examples/include/shapeFunc.hpp:200-269, 525-602).
struct MidEdge { int dir; int s[3]; };
constexpr MidEdge kMid[12] = {
    {0, {0, -1, -1}}, {1, {1, 0, -1}}, {0, {0, 1, -1}}, {1, {-1, 0, -1}},  // 8..11 bottom
    {0, {0, -1, 1}},  {1, {1, 0, 1}},  {0, {0, 1, 1}},  {1, {-1, 0, 1}},   // 12..15 top
    {2, {-1, -1, 0}}, {2, {1, -1, 0}}, {2, {1, 1, 0}},  {2, {-1, 1, 0}}};  // 16..19 vertical

void dshape_hex20(const double* x, double* dN)
{
    for (int a = 0; a < 8; a++) {
        const double s[3] = {(double)kCorner[a][0], (double)kCorner[a][1], (double)kCorner[a][2]};
        const double f[3] = {1.0 + s[0] * x[0], 1.0 + s[1] * x[1], 1.0 + s[2] * x[2]};
        const double q = s[0] * x[0] + s[1] * x[1] + s[2] * x[2] - 2.0;
        // N = f0 f1 f2 q / 8  ->  dN/dx_d = (s_d q + s_d f_d) prod_{other} f / 8
        for (int d = 0; d < 3; d++) {
            const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            dN[3 * a + d] = 0.125 * f[d1] * f[d2] * s[d] * (q + f[d]);
        }
    }
    for (int b = 0; b < 12; b++) {
        const MidEdge& me = kMid[b];
        const int d = me.dir, d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        const double g1 = 1.0 + me.s[d1] * x[d1], g2 = 1.0 + me.s[d2] * x[d2];
        const double bub = 1.0 - x[d] * x[d];
        double* o = dN + 3 * (8 + b);
        o[d] = 0.25 * (-2.0 * x[d]) * g1 * g2;
        o[d1] = 0.25 * bub * me.s[d1] * g2;
        o[d2] = 0.25 * bub * me.s[d2] * g1;
    }
}

// tet10 in barycentrics L0 = 1-x-y-z, L1 = x, L2 = y, L3 = z; vertices 0..3,
// mid-edge nodes 4:(0,1) 5:(1,2) 6:(2,0) 7:(0,3) 8:(1,3) 9:(2,3)
// (examples/include/shapeFunc.hpp:178-199, 384-448).
constexpr int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
constexpr double kdL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

void shape_tet10(const double* x, double* N, double* dN)
{
    const double L[4] = {1.0 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
    for (int a = 0; a < 4; a++) {
        if (N) N[a] = L[a] * (2.0 * L[a] - 1.0);
        for (int d = 0; d < 3; d++) dN[3 * a + d] = (4.0 * L[a] - 1.0) * kdL[a][d];
    }
    for (int b = 0; b < 6; b++) {
        const int i = kTetEdge[b][0], j = kTetEdge[b][1];
        if (N) N[4 + b] = 4.0 * L[i] * L[j];
        for (int d = 0; d < 3; d++) dN[3 * (4 + b) + d] = 4.0 * (kdL[i][d] * L[j] + L[i] * kdL[j][d]);
    }
}

// Jacobian of the map xi -> x for nn nodes; returns det and writes the
// inverse-transposed rows so that dN/dx_k = sum_d dN/dxi_d * inv[d][k].
double jacobian(int nn, const double* xe, const double* dN, double inv[3][3])
{
    double J[3][3] = {{0}};  // J[k][d] = dx_k / dxi_d
    for (int a = 0; a < nn; a++)
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 3; d++) J[k][d] += xe[3 * a + k] * dN[3 * a + d];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) +
                       J[1][0] * (J[2][1] * J[0][2] - J[0][1] * J[2][2]) +
                       J[2][0] * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
    // inv[d][k] = dxi_d / dx_k = cofactor(J)[k][d] / det
    inv[0][0] = (J[2][2] * J[1][1] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[2][2] * J[0][1]) / det;
    inv[0][2] = (J[1][2] * J[0][1] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[2][2] * J[1][0]) / det;
    inv[1][1] = (J[2][2] * J[0][0] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[1][2] * J[0][0]) / det;
    inv[2][0] = (J[2][1] * J[1][0] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[2][1] * J[0][0]) / det;
    inv[2][2] = (J[1][1] * J[0][0] - J[0][1] * J[1][0]) / det;
    return det;
}

void physical_gradients(int nn, const double* dN, const double inv[3][3], double* g /* g[3*a+k] */)
{
    for (int a = 0; a < nn; a++)
        for (int k = 0; k < 3; k++)
            g[3 * a + k] = dN[3 * a] * inv[0][k] + dN[3 * a + 1] * inv[1][k] + dN[3 * a + 2] * inv[2][k];
}

void ke_hex8(const double* xe, uint32_t ngt, double* ke)
{
    double pw[2 * 16];
    gauss_legendre(ngt, pw);
    for (int i = 0; i < 64; i++) ke[i] = 0.0;
    double dN[24], g[24], inv[3][3];
    for (uint32_t i = 0; i < ngt; i++)
        for (uint32_t j = 0; j < ngt; j++)
            for (uint32_t k = 0; k < ngt; k++) {
                const double xi[3] = {pw[2 * i], pw[2 * j], pw[2 * k]};
                dshape_hex8(xi, dN);
                const double det = jacobian(8, xe, dN, inv);
                if (!(det > 0.0)) amx::fail(AMX_INVALID_ARGUMENT, "ke_hex8: non-positive Jacobian");
                physical_gradients(8, dN, inv, g);
                const double w = det * pw[2 * k + 1] * pw[2 * j + 1] * pw[2 * i + 1];
                for (int r = 0; r < 8; r++)
                    for (int c = 0; c < 8; c++)
                        ke[r * 8 + c] += (g[3 * r] * g[3 * c] + g[3 * r + 1] * g[3 * c + 1] + g[3 * r + 2] * g[3 * c + 2]) * w;
            }
}

void ke_hex20_iso(const double* xe, double E, double nu, uint32_t ngt, double* ke)
{
    constexpr int NN = 20, ND = 60;
    double pw[2 * 16];
    gauss_legendre(ngt, pw);
    for (int i = 0; i < ND * ND; i++) ke[i] = 0.0;
    // isotropic constitutive matrix in Voigt order xx,yy,zz,xy,yz,xz
    const double c1 = E * (1.0 - nu) / (1.0 + nu) / (1.0 - 2.0 * nu);
    const double c2 = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);
    const double c3 = 0.5 * E / (1.0 + nu);
    double dN[3 * NN], g[3 * NN], inv[3][3];
    static thread_local double B[6][ND], CB[6][ND];
    for (uint32_t i = 0; i < ngt; i++)
        for (uint32_t j = 0; j < ngt; j++)
            for (uint32_t k = 0; k < ngt; k++) {
                const double xi[3] = {pw[2 * i], pw[2 * j], pw[2 * k]};
                dshape_hex20(xi, dN);
                const double det = jacobian(NN, xe, dN, inv);
                if (!(det > 0.0)) amx::fail(AMX_INVALID_ARGUMENT, "ke_hex20_iso: non-positive Jacobian");
                physical_gradients(NN, dN, inv, g);
                std::memset(B, 0, sizeof(B));
                for (int a = 0; a < NN; a++) {
                    const double gx = g[3 * a], gy = g[3 * a + 1], gz = g[3 * a + 2];
                    B[0][3 * a] = gx;
                    B[1][3 * a + 1] = gy;
                    B[2][3 * a + 2] = gz;
                    B[3][3 * a] = gy;
                    B[3][3 * a + 1] = gx;
                    B[4][3 * a + 1] = gz;
                    B[4][3 * a + 2] = gy;
                    B[5][3 * a] = gz;
                    B[5][3 * a + 2] = gx;
                }
                for (int c = 0; c < ND; c++) {
                    CB[0][c] = c1 * B[0][c] + c2 * B[1][c] + c2 * B[2][c];
                    CB[1][c] = c2 * B[0][c] + c1 * B[1][c] + c2 * B[2][c];
                    CB[2][c] = c2 * B[0][c] + c2 * B[1][c] + c1 * B[2][c];
                    CB[3][c] = c3 * B[3][c];
                    CB[4][c] = c3 * B[4][c];
                    CB[5][c] = c3 * B[5][c];
                }
                const double w = det * pw[2 * k + 1] * pw[2 * j + 1] * pw[2 * i + 1];
                for (int r = 0; r < ND; r++)
                    for (int c = 0; c < ND; c++) {
                        double s = 0.0;
                        for (int q = 0; q < 6; q++) s += B[q][r] * CB[q][c];
                        ke[r * ND + c] += s * w;
                    }
            }
}

// Collapsed-cube (Duffy) rule on the unit tetrahedron used by the reference
// (examples/src/ke_matrix.cpp:362-365): (xi,eta,zeta) in [-1,1]^3 ->
// x = (1+xi)(1+eta)(1+zeta)/8, y = (1+xi)(1+eta)(1-zeta)/8, z = (1+xi)(1-eta)/4,
// weight (1+xi)^2 (1+eta) / 64 * w_xi w_eta w_zeta.
template <typename F> void tet_rule(uint32_t ngt, F&& f)
{
    double pw[2 * 16];
    gauss_legendre(ngt, pw);
    for (uint32_t i = 0; i < ngt; i++)
        for (uint32_t j = 0; j < ngt; j++)
            for (uint32_t k = 0; k < ngt; k++) {
                const double a = pw[2 * i], b = pw[2 * j], c = pw[2 * k];
                const double cm = ((1 + a) * (1 + a) * (1 + b) / 64) * pw[2 * i + 1] * pw[2 * j + 1] * pw[2 * k + 1];
                const double x[3] = {(1 + a) * (1 + b) * (1 + c) / 8, (1 + a) * (1 + b) * (1 - c) / 8, (1 + a) * (1 - b) / 4};
                f(x, cm);
            }
}

void ke_tet10(const double* xe, uint32_t ngt, double* ke)
{
    for (int i = 0; i < 100; i++) ke[i] = 0.0;
    double dN[30], g[30], inv[3][3];
    tet_rule(ngt, [&](const double* x, double cm) {
        shape_tet10(x, nullptr, dN);
        const double det = jacobian(10, xe, dN, inv);
        if (!(det > 0.0)) amx::fail(AMX_INVALID_ARGUMENT, "ke_tet10: non-positive Jacobian");
        physical_gradients(10, dN, inv, g);
        const double w = det * cm;
        for (int r = 0; r < 10; r++)
            for (int c = 0; c < 10; c++)
                ke[r * 10 + c] += (g[3 * r] * g[3 * c] + g[3 * r + 1] * g[3 * c + 1] + g[3 * r + 2] * g[3 * c + 2]) * w;
    });
}

void fe_tet10(const double* xe, const double* be, uint32_t ngt, double* fe)
{
    for (int i = 0; i < 10; i++) fe[i] = 0.0;
    double N[10], dN[30], inv[3][3];
    tet_rule(ngt, [&](const double* x, double cm) {
        shape_tet10(x, N, dN);
        const double det = jacobian(10, xe, dN, inv);
        if (!(det > 0.0)) amx::fail(AMX_INVALID_ARGUMENT, "fe_tet10: non-positive Jacobian");
        double nb = 0.0;
        for (int j = 0; j < 10; j++) nb += N[j] * be[j];
        for (int i = 0; i < 10; i++) fe[i] += N[i] * nb * det * cm;
    });
}

void check_ngt(uint32_t ngt)
{
    AMX_REQUIRE(ngt >= 1 && ngt <= 16, AMX_INVALID_ARGUMENT, "ngt must be in [1, 16]");
}

}  // namespace

namespace amx {

// Quadrature table of the FREE kernels: per point, dN_a/dxi_d at 3a+d and the
// weight at 3*nodes; the points and their order are those of ke_hex8 /
// ke_hex20_iso (i, j, k loops, k fastest) and of the tetrahedral rule of
// ke_tet10.  Rows from nq up to `rows` repeat the first point with weight 0.
std::vector<double> free_quadrature(int kind, uint32_t ngt, uint32_t& nq, uint32_t rows_min)
{
    check_ngt(ngt);
    const int nn = kind == AMX_ELEM_HEX8_LAPLACE ? 8 : (kind == AMX_ELEM_HEX20_ELASTIC ? 20 : 10);
    const uint32_t w = 3 * nn + 1;
    std::vector<double> t;
    if (kind == AMX_ELEM_TET10_LAPLACE) {
        tet_rule(ngt, [&](const double* x, double cm) {
            const size_t o = t.size();
            t.resize(o + w);
            shape_tet10(x, nullptr, t.data() + o);
            t[o + 3 * nn] = cm;
        });
    } else {
        double pw[2 * 16];
        gauss_legendre(ngt, pw);
        for (uint32_t i = 0; i < ngt; i++)
            for (uint32_t j = 0; j < ngt; j++)
                for (uint32_t k = 0; k < ngt; k++) {
                    const double xi[3] = {pw[2 * i], pw[2 * j], pw[2 * k]};
                    const size_t o = t.size();
                    t.resize(o + w);
                    if (nn == 8) dshape_hex8(xi, t.data() + o);
                    else dshape_hex20(xi, t.data() + o);
                    t[o + 3 * nn] = pw[2 * k + 1] * pw[2 * j + 1] * pw[2 * i + 1];
                }
    }
    nq = (uint32_t)(t.size() / w);
    for (uint32_t r = nq; r < rows_min; r++) {
        const size_t o = t.size();
        t.resize(o + w);
        std::copy(t.begin(), t.begin() + w, t.begin() + o);
        t[o + 3 * nn] = 0.0;
    }
    return t;
}

}  // namespace amx

extern "C" {

int amx_fem_gauss(uint32_t n, double* pts_n_wts)
{
    return guarded(__func__, [&] { check_ngt(n); gauss_legendre(n, pts_n_wts); });
}

int amx_fem_ke_hex8(const double* xe, uint32_t ngt, double* ke)
{
    return guarded(__func__, [&] { check_ngt(ngt); ke_hex8(xe, ngt, ke); });
}

int amx_fem_ke_hex20_iso(const double* xe, double E, double nu, uint32_t ngt, double* ke)
{
    return guarded(__func__, [&] { check_ngt(ngt); ke_hex20_iso(xe, E, nu, ngt, ke); });
}

int amx_fem_ke_tet10(const double* xe, uint32_t ngt, double* ke)
{
    return guarded(__func__, [&] { check_ngt(ngt); ke_tet10(xe, ngt, ke); });
}

int amx_fem_fe_tet10(const double* xe, const double* be, uint32_t ngt, double* fe)
{
    return guarded(__func__, [&] { check_ngt(ngt); fe_tet10(xe, be, ngt, fe); });
}

int amx_fem_ke_tet10_batch(uint32_t n_elem, const double* coords, uint32_t ngt, double* ke)
{
    return guarded(__func__, [&] {
        check_ngt(ngt);
        int bad = -1;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < (int64_t)n_elem; e++) {
            try {
                ke_tet10(coords + 30 * e, ngt, ke + 100 * e);
            } catch (...) {
#pragma omp critical
                bad = (int)e;
            }
        }
        AMX_REQUIRE(bad < 0, AMX_INVALID_ARGUMENT, "ke_tet10: non-positive Jacobian in element " + std::to_string(bad));
    });
}

int amx_fem_fe_tet10_batch(uint32_t n_elem, const double* coords, const double* be, uint32_t ngt, double* fe)
{
    return guarded(__func__, [&] {
        check_ngt(ngt);
        int bad = -1;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < (int64_t)n_elem; e++) {
            try {
                fe_tet10(coords + 30 * e, be + 10 * e, ngt, fe + 10 * e);
            } catch (...) {
#pragma omp critical
                bad = (int)e;
            }
        }
        AMX_REQUIRE(bad < 0, AMX_INVALID_ARGUMENT, "fe_tet10: non-positive Jacobian in element " + std::to_string(bad));
    });
}

}  // extern "C"
