// ref_fem_wrap.cpp -- C entry points over the reference element kernels.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (oracle/Makefile, target `ref`) together
// with the reference sources where they lie (examples/src/ke_matrix.cpp,
// examples/src/fe_vector.cpp, the vendored Eigen) into oracle/_ref/; used only
// to write the golden fixtures under tests/golden/.  This file holds no
// reference code, only calls into it.
#include "fe_vector.hpp"
#include "integration.hpp"
#include "ke_matrix.hpp"

extern "C" {

// integration<double>(n).Pts_n_Wts (examples/include/integration.hpp:81-113)
void ref_gauss(unsigned n, double* pts_n_wts)
{
    integration<double> g(n);
    for (unsigned i = 0; i < 2 * n; i++) pts_n_wts[i] = g.Pts_n_Wts[i];
}

// ke_hex8 (examples/src/ke_matrix.cpp:436-526), row-major 8x8
void ref_ke_hex8(double* ke, const double* xe, unsigned ngt)
{
    integration<double> g(ngt);
    ke_hex8(ke, xe, g.Pts_n_Wts, ngt);
}

// ke_hex8_eig (examples/src/ke_matrix.cpp:530-619), written out row-major
void ref_ke_hex8_eig(double* ke, const double* xe, unsigned ngt)
{
    integration<double> g(ngt);
    Eigen::Matrix<double, 8, 8> k;
    double x[24];
    for (int i = 0; i < 24; i++) x[i] = xe[i];
    ke_hex8_eig(k, x, g.Pts_n_Wts, ngt);
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) ke[r * 8 + c] = k(r, c);
}

// ke_hex20_iso (examples/src/ke_matrix.cpp:1573-1740), row-major 60x60
void ref_ke_hex20_iso(double* ke, const double* xe, double E, double nu, unsigned ngt)
{
    integration<double> g(ngt);
    double x[60];
    for (int i = 0; i < 60; i++) x[i] = xe[i];
    ke_hex20_iso(ke, x, E, nu, g.Pts_n_Wts, ngt);
}

// ke_tet10 (examples/src/ke_matrix.cpp:329-433), row-major 10x10
void ref_ke_tet10(double* ke, const double* xe, unsigned ngt)
{
    integration<double> g(ngt);
    ke_tet10(ke, xe, g.Pts_n_Wts, ngt);
}

// fe_tet10 (examples/src/fe_vector.cpp:86-165)
void ref_fe_tet10(double* fe, const double* xe, const double* be, unsigned ngt)
{
    integration<double> g(ngt);
    Eigen::Matrix<double, Eigen::Dynamic, 1> f(10);
    fe_tet10(f, xe, be, g.Pts_n_Wts, ngt);
    for (int i = 0; i < 10; i++) fe[i] = f(i);
}
}
