// dropin_example.cpp -- uses the header-only par::aMatFree drop-in the way
// examples/src/example03a.cpp uses the reference (hex8 Poisson, identity BCs
// on the boundary, GPU_PURE), then checks the GPU results against the oracle.
//
// Test program: built by amat_amd/build.py, run by tests/test_gpu_dropin.py.
// Prints one line "dropin <case> maxdiff_matvec=... maxdiff_mf=..." per case
// and exits non-zero on any mismatch.  With a Gmsh file as argument it also
// runs the set-up of examples/src/example05a.cpp:135-246 (tet10 Laplace read
// by GmshMesh) on one rank; with a second argument, the driver pattern of
// examples/src/example07a.cpp (the operator held through the par::aMat base,
// dump_mat into that file, get_timer).
#include <cmath>
#include <cstdio>
#include <vector>

#include "amat_amd/GmshMesh.hpp"
#include "amat_amd/aMatFree.hpp"
#include "amx_oracle.h"

struct Mat8 {  // anything with rows(), cols(), operator()(r, c) works
    double a[64];
    long rows() const { return 8; }
    long cols() const { return 8; }
    double operator()(long r, long c) const { return a[r * 8 + c]; }
};

static int run_case(unsigned Nex, unsigned Ney, unsigned Nez, const char* dump_path = nullptr)
{
    using Maps = par::Maps<double, unsigned long, unsigned int>;
    using AFM = par::aMatFree<double, unsigned long, unsigned int>;
    const unsigned nelem = Nex * Ney * Nez, nnode = (Nex + 1) * (Ney + 1) * (Nez + 1);
    const double hx = 1.0 / Nex, hy = 1.0 / Ney, hz = 1.0 / Nez;
    std::vector<std::vector<unsigned>> map(nelem, std::vector<unsigned>(8));
    std::vector<unsigned*> mapp(nelem);
    std::vector<unsigned> dpe(nelem, 8);
    for (unsigned k = 0; k < Nez; k++)
        for (unsigned j = 0; j < Ney; j++)
            for (unsigned i = 0; i < Nex; i++) {
                const unsigned e = Nex * Ney * k + Nex * j + i;
                const unsigned n0 = i + j * (Nex + 1) + k * (Nex + 1) * (Ney + 1);
                const unsigned sx = Nex + 1, sy = (Nex + 1) * (Ney + 1);
                unsigned c[8] = {n0, n0 + 1, n0 + sx + 1, n0 + sx, n0 + sy, n0 + sy + 1, n0 + sy + sx + 1, n0 + sy + sx};
                for (int q = 0; q < 8; q++) map[e][q] = c[q];
                mapp[e] = map[e].data();
            }
    std::vector<unsigned long> l2g(nnode);
    for (unsigned i = 0; i < nnode; i++) l2g[i] = i;
    std::vector<unsigned long> cons;
    for (unsigned n = 0; n < nnode; n++) {
        const unsigned i = n % (Nex + 1), j = (n / (Nex + 1)) % (Ney + 1), k = n / ((Nex + 1) * (Ney + 1));
        if (i == 0 || j == 0 || k == 0 || i == Nex || j == Ney || k == Nez) cons.push_back(n);
    }
    std::vector<double> vals(cons.size(), 0.0);

    Maps maps;
    maps.set_map(nelem, mapp.data(), dpe.data(), nnode, l2g.data(), 0, nnode - 1, nnode);
    maps.set_bdr_map(cons.data(), vals.data(), (unsigned)cons.size());
    AFM A(maps, par::BC_METH::BC_IMATRIX);
    A.set_matfree_type(par::MATFREE_TYPE::GPU_PURE);
    Mat8 ke;
    std::vector<double> Kall((size_t)nelem * 64);
    for (unsigned e = 0; e < nelem; e++) {
        double xe[24];
        for (int q = 0; q < 8; q++) {
            const unsigned n = map[e][q];
            xe[3 * q] = (n % (Nex + 1)) * hx;
            xe[3 * q + 1] = ((n / (Nex + 1)) % (Ney + 1)) * hy;
            xe[3 * q + 2] = (n / ((Nex + 1) * (Ney + 1))) * hz;
        }
        par::detail::check(amx_fem_ke_hex8(xe, 2, ke.a), "ke_hex8");
        for (int q = 0; q < 64; q++) Kall[(size_t)e * 64 + q] = ke.a[q];
        A.set_element_matrix(e, ke, 0, 0, 1);
    }
    A.finalize();

    std::vector<double> u(nnode), v(nnode), vref(nnode), vm(nnode), vmref(nnode);
    for (unsigned i = 0; i < nnode; i++) u[i] = std::sin(0.37 * i) + 0.1 * std::cos(1.3 * i);
    A.matvec(v.data(), u.data(), true);
    A.matmult_mf(vm.data(), u.data());
    // the timers the examples print after their GPU matvecs (example07a.cpp:573)
    long double scatter_time = 0, gather_time = 0, mv_time = 0, mvTotal_time = 0;
    A.get_timer(&scatter_time, &gather_time, &mv_time, &mvTotal_time);
    const bool timers = scatter_time > 0 && gather_time > 0 && mv_time > 0 &&
                        mvTotal_time >= mv_time && mvTotal_time < 60;

    std::vector<unsigned> flat, order(nelem);
    for (unsigned e = 0; e < nelem; e++) {
        flat.insert(flat.end(), map[e].begin(), map[e].end());
        order[e] = e;
    }
    orc_matvec_ghosted_omp(nelem, dpe.data(), flat.data(), Kall.data(), 0, order.data(), nelem, nullptr, 0, u.data(),
                           vref.data(), nnode, 1);
    std::vector<uint64_t> oc(cons.begin(), cons.end());
    orc_matmult_mf(nelem, dpe.data(), flat.data(), Kall.data(), nnode, 0, nnode, oc.data(), (uint32_t)oc.size(), 0,
                   u.data(), vmref.data(), 1);
    double d1 = 0, d2 = 0;
    for (unsigned i = 0; i < nnode; i++) {
        d1 = std::fmax(d1, std::fabs(v[i] - vref[i]));
        d2 = std::fmax(d2, std::fabs(vm[i] - vmref[i]));
    }
    std::printf("dropin hex8_%ux%ux%u maxdiff_matvec=%.3e maxdiff_mf=%.3e timers=%s\n", Nex, Ney, Nez, d1, d2,
                timers ? "ok" : "bad");
    int bad = (d1 == 0.0 && d2 == 0.0 && timers) ? 0 : 1;
    if (!dump_path) return bad;

    // the driver pattern of examples/src/example07a.cpp:729-745, 796-815, 880,
    // 903-905, 981, 1128: the operator held through the par::aMat base
    typedef par::aMat<AFM, double, unsigned long, unsigned int> aMatFreeBase;
    aMatFreeBase* stMatFree = new AFM(maps, par::BC_METH::BC_IMATRIX);
    stMatFree->set_matfree_type(par::MATFREE_TYPE::GPU_PURE);
    stMatFree->set_num_streams(1);
    for (unsigned e = 0; e < nelem; e++) {
        for (int q = 0; q < 64; q++) ke.a[q] = Kall[(size_t)e * 64 + q];
        stMatFree->set_element_matrix(e, ke, 0, 0, 1);
    }
    stMatFree->finalize();
    std::vector<double> vb(nnode);
    stMatFree->matvec(vb.data(), u.data(), true);
    stMatFree->dump_mat(dump_path);
    long double st = 0, gt = 0, mt = 0, tt = 0;
    stMatFree->get_timer(&st, &gt, &mt, &tt);
    const bool base_ok = vb == v && st > 0 && gt > 0 && mt > 0 && tt >= mt &&
                         stMatFree->get_matrix_type() == par::MATRIX_TYPE::MATRIX_FREE;
    delete stMatFree;
    // the dump: symmetric bit for bit (symmetric K_e, the same element order in
    // both entries) and its product with u equals matvec
    std::vector<std::vector<std::pair<unsigned, double>>> rows(nnode);
    std::FILE* f = std::fopen(dump_path, "r");
    size_t n_entries = 0;
    bool sym = f != nullptr;
    if (f) {
        unsigned long long r, c;
        double val;
        while (std::fscanf(f, "%llu %llu %lg", &r, &c, &val) == 3) {
            if (r >= nnode || c >= nnode) { sym = false; break; }
            rows[r].push_back({(unsigned)c, val});
            n_entries++;
        }
        std::fclose(f);
    }
    double err = 0, vmax = 0;
    for (unsigned r = 0; r < nnode && sym; r++) {
        double y = 0;
        for (auto& e : rows[r]) {
            y += e.second * u[e.first];
            bool found = false;
            for (auto& t : rows[e.first]) found |= (t.first == r && t.second == e.second);
            sym &= found;
        }
        err = std::fmax(err, std::fabs(y - v[r]));
        vmax = std::fmax(vmax, std::fabs(v[r]));
    }
    const double rel = err / vmax;
    std::printf("dropin base_pointer_%ux%ux%u entries=%zu symmetric=%s dump_apply_rel=%.3e base=%s\n", Nex, Ney, Nez,
                n_entries, sym ? "ok" : "bad", rel, base_ok ? "ok" : "bad");
    bad |= (sym && base_ok && n_entries > 0 && rel <= 1e-13) ? 0 : 1;
    return bad;
}

// method 2 (FREE) the way examples/src/example07a.cpp:54-70 sets it up: an
// element function reads the application's mesh and fills ke and xe
static struct {
    unsigned Nex, Ney, Nez;
    std::vector<std::vector<unsigned>> map;
} g_app;

static void node_xyz(unsigned n, double* x)
{
    const unsigned sx = g_app.Nex + 1, sy = (g_app.Nex + 1) * (g_app.Ney + 1);
    // a smooth distortion keeps the elements distinct (and the Jacobians positive)
    const double X = (double)(n % sx) / g_app.Nex, Y = (double)((n / sx) % (g_app.Ney + 1)) / g_app.Ney,
                 Z = (double)(n / sy) / g_app.Nez;
    x[0] = X + 0.03 * std::sin(3.0 * Y + Z);
    x[1] = Y + 0.03 * std::sin(2.0 * Z + X);
    x[2] = Z + 0.03 * std::sin(2.5 * X + Y);
}

static void compute_elem_mat(unsigned eid, double* ke, double* xe)
{
    for (int q = 0; q < 8; q++) node_xyz(g_app.map[eid][q], xe + 3 * q);
    par::detail::check(amx_fem_ke_hex8(xe, 2, ke), "ke_hex8");
}

static int run_free_case(unsigned Nex, unsigned Ney, unsigned Nez)
{
    using Maps = par::Maps<double, unsigned long, unsigned int>;
    using AFM = par::aMatFree<double, unsigned long, unsigned int>;
    g_app.Nex = Nex;
    g_app.Ney = Ney;
    g_app.Nez = Nez;
    const unsigned nelem = Nex * Ney * Nez, nnode = (Nex + 1) * (Ney + 1) * (Nez + 1);
    const unsigned sx = Nex + 1, sy = (Nex + 1) * (Ney + 1);
    g_app.map.assign(nelem, std::vector<unsigned>(8));
    std::vector<unsigned*> mapp(nelem);
    std::vector<unsigned> dpe(nelem, 8), flat, order(nelem);
    for (unsigned k = 0; k < Nez; k++)
        for (unsigned j = 0; j < Ney; j++)
            for (unsigned i = 0; i < Nex; i++) {
                const unsigned e = Nex * Ney * k + Nex * j + i, n0 = i + j * sx + k * sy;
                const unsigned c[8] = {n0, n0 + 1, n0 + sx + 1, n0 + sx, n0 + sy, n0 + sy + 1, n0 + sy + sx + 1,
                                       n0 + sy + sx};
                for (int q = 0; q < 8; q++) g_app.map[e][q] = c[q];
                mapp[e] = g_app.map[e].data();
                flat.insert(flat.end(), c, c + 8);
                order[e] = e;
            }
    std::vector<unsigned long> l2g(nnode);
    for (unsigned i = 0; i < nnode; i++) l2g[i] = i;
    std::vector<unsigned long> cons;
    std::vector<double> vals;
    Maps maps;
    maps.set_map(nelem, mapp.data(), dpe.data(), nnode, l2g.data(), 0, nnode - 1, nnode);
    maps.set_bdr_map(cons.data(), vals.data(), 0);

    // (a) the reference's callback form
    AFM A(maps, par::BC_METH::BC_IMATRIX);
    A.set_matfree_type(par::MATFREE_TYPE::FREE);
    A.set_element_matrix_function(&compute_elem_mat);
    A.finalize();
    // (b) the in-kernel form on the same coordinates
    std::vector<double> xall((size_t)nelem * 24), Kall((size_t)nelem * 64);
    for (unsigned e = 0; e < nelem; e++) compute_elem_mat(e, &Kall[(size_t)e * 64], &xall[(size_t)e * 24]);
    AFM B(maps, par::BC_METH::BC_IMATRIX);
    B.set_matfree_type(par::MATFREE_TYPE::FREE);
    B.set_element_function(AMX_ELEM_HEX8_LAPLACE, xall.data(), nullptr, 1.0, 0.3, 2);
    B.finalize();

    std::vector<double> u(nnode), va(nnode), vb(nnode), vref(nnode);
    for (unsigned i = 0; i < nnode; i++) u[i] = std::sin(0.37 * i) + 0.1 * std::cos(1.3 * i);
    A.matvec(va.data(), u.data(), true);
    B.matvec(vb.data(), u.data(), true);
    orc_matvec_ghosted_omp(nelem, dpe.data(), flat.data(), Kall.data(), 0, order.data(), nelem, nullptr, 0, u.data(),
                           vref.data(), nnode, 1);
    double da = 0, db = 0, nr = 0;
    for (unsigned i = 0; i < nnode; i++) {
        da = std::fmax(da, std::fabs(va[i] - vref[i]));
        db = std::fmax(db, std::fabs(vb[i] - vref[i]));
        nr = std::fmax(nr, std::fabs(vref[i]));
    }
    std::printf("dropin free_callback_%ux%ux%u maxdiff_matvec=%.3e\n", Nex, Ney, Nez, da);
    std::printf("dropin free_inkernel_%ux%ux%u relmax=%.3e\n", Nex, Ney, Nez, db / nr);
    return (da == 0.0 && db <= 1e-13 * nr) ? 0 : 1;
}

struct RowView {  // a row-major n x n block
    const double* a;
    long n;
    long rows() const { return n; }
    long cols() const { return n; }
    double operator()(long r, long c) const { return a[r * n + c]; }
};

static void zero_displacement(const double*, double* uvw) { *uvw = 0.0; }

static int run_gmsh_case(const char* path)
{
    using Maps = par::Maps<double, unsigned long, unsigned int>;
    using AFM = par::aMatFree<double, unsigned long, unsigned int>;
    GmshMesh mesh(0, 1, 1);
    mesh.setGmshMesh(path);
    const unsigned int nelem_owned = mesh.get_nOwnedElems();
    unsigned int** localDofMap = mesh.get_localDofMap();
    unsigned int* nDofsPerElem = mesh.get_nDofsPerElem();
    const unsigned int nLocalDofs = mesh.get_nLocalDofs();
    unsigned long* local2GlobalDofMap = mesh.get_local2GlobalDofMap();
    const unsigned long startGlobalDof = mesh.get_startGlobalDof();
    const unsigned long endGlobalDof = mesh.get_endGlobalDof();
    const unsigned long nDofsTotal = mesh.get_nDofsTotal();
    const unsigned int nConstraints = mesh.get_n_constraints();
    unsigned long* constrainedDofs = mesh.get_constrained_dofs();
    double* prescribedValues = mesh.get_prescribed_values(&zero_displacement);
    unsigned int* nNodesPerElem = mesh.get_nNodesPerElem();
    unsigned int** localMap = mesh.get_localMap();

    Maps meshMaps;
    meshMaps.set_map(nelem_owned, localDofMap, nDofsPerElem, nLocalDofs, local2GlobalDofMap, startGlobalDof,
                     endGlobalDof, nDofsTotal);
    meshMaps.set_bdr_map(constrainedDofs, prescribedValues, nConstraints);
    AFM A(meshMaps, par::BC_METH::BC_IMATRIX);
    A.set_matfree_type(par::MATFREE_TYPE::GPU_PURE);
    std::vector<double> Kall((size_t)nelem_owned * 100);
    std::vector<unsigned> flat, order(nelem_owned);
    for (unsigned eid = 0; eid < nelem_owned; eid++) {
        if (nNodesPerElem[eid] != 10) {
            std::printf("dropin gmsh: element %u has %u nodes\n", eid, nNodesPerElem[eid]);
            return 1;
        }
        double xe[30];
        for (unsigned nid = 0; nid < 10; nid++) {
            const unsigned localNid = localMap[eid][nid];
            xe[nid * 3] = mesh.get_x(localNid);
            xe[nid * 3 + 1] = mesh.get_y(localNid);
            xe[nid * 3 + 2] = mesh.get_z(localNid);
        }
        par::detail::check(amx_fem_ke_tet10(xe, 2, &Kall[(size_t)eid * 100]), "ke_tet10");
        A.set_element_matrix(eid, RowView{&Kall[(size_t)eid * 100], 10}, 0, 0, 1);
        flat.insert(flat.end(), localDofMap[eid], localDofMap[eid] + 10);
        order[eid] = eid;
    }
    A.finalize();
    std::vector<double> u(nLocalDofs), v(nLocalDofs), vref(nLocalDofs), vm(nLocalDofs), vmref(nLocalDofs);
    for (unsigned i = 0; i < nLocalDofs; i++) u[i] = std::sin(0.41 * i) + 0.2 * std::cos(0.9 * i);
    A.matvec(v.data(), u.data(), true);
    A.matmult_mf(vm.data(), u.data());
    orc_matvec_ghosted_omp(nelem_owned, nDofsPerElem, flat.data(), Kall.data(), 0, order.data(), nelem_owned, nullptr,
                           0, u.data(), vref.data(), nLocalDofs, 1);
    std::vector<uint64_t> oc(constrainedDofs, constrainedDofs + nConstraints);
    orc_matmult_mf(nelem_owned, nDofsPerElem, flat.data(), Kall.data(), nLocalDofs, 0, nLocalDofs, oc.data(),
                   (uint32_t)oc.size(), 0, u.data(), vmref.data(), 1);
    double d1 = 0, d2 = 0;
    for (unsigned i = 0; i < nLocalDofs; i++) {
        d1 = std::fmax(d1, std::fabs(v[i] - vref[i]));
        d2 = std::fmax(d2, std::fabs(vm[i] - vmref[i]));
    }
    std::printf("dropin gmsh_tet10_%u maxdiff_matvec=%.3e maxdiff_mf=%.3e constraints=%u\n", nelem_owned, d1, d2,
                nConstraints);
    return (d1 == 0.0 && d2 == 0.0) ? 0 : 1;
}

int main(int argc, char** argv)
{
    int bad = 0;
    try {
        if (argc > 1) bad |= run_gmsh_case(argv[1]);
        bad |= run_case(4, 5, 6, argc > 2 ? argv[2] : nullptr);
        bad |= run_case(16, 16, 16);
        bad |= run_free_case(7, 6, 5);
        // error path: methods 1/2 belong to the reference's CPU loops
        par::Maps<double, unsigned long, unsigned int> m;
        unsigned row[8] = {0, 1, 2, 3, 4, 5, 6, 7}, *rows[1] = {row}, d8[1] = {8};
        unsigned long l2g[8] = {0, 1, 2, 3, 4, 5, 6, 7};
        m.set_map(1, rows, d8, 8, l2g, 0, 7, 8);
        par::aMatFree<double, unsigned long, unsigned int> A(m);
        try {
            A.set_matfree_type(par::MATFREE_TYPE::QUASI_FREE);
            std::printf("dropin error-path: missing NOT_IMPLEMENTED\n");
            bad = 1;
        } catch (const par::amx_error& e) {
            std::printf("dropin error-path code=%d\n", e.code);
            bad |= (e.code != AMX_NOT_IMPLEMENTED);
        }
    } catch (const std::exception& e) {
        std::printf("dropin exception: %s\n", e.what());
        return 2;
    }
    return bad;
}
