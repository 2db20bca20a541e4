// petsc_shell_example.cpp -- the PETSc shell-matrix path of the C++ drop-in
// (include/amat_amd/aMatFree.hpp built with BUILD_WITH_PETSC) driven the way
// the reference's solver driver drives it: examples/include/solve.hpp:31-83
// takes matrix.get_comm() and matrix.get_matrix(), lets KSPCG + PCJACOBI call
// the shell's MATOP_MULT and MATOP_GET_DIAGONAL, then fetches the context
// with MatShellGetContext and deletes it.  PETSc is not installed: the
// program compiles against the test double tests/cpp/petsc_stub/petsc.h, and
// the Krylov loop below stands in for KSPCG + PCJACOBI (Jacobi-preconditioned
// CG on the shell's MatMult, dot products summed over the ranks with
// MPI_Allreduce).  MPI is real (MPICH): run it directly for one rank or under
// mpiexec for several; the mesh is a structured hex8 Poisson cube cut into
// z-slabs as examples/src/example03a.cpp:190-255 cuts it, and the Maps take
// MPI_COMM_WORLD, so set_map runs the ghost collectives and the operator's
// halo exchanges are MPI messages.
//
// Usage: petsc_shell_example OUTDIR [Nex Ney Nez]
// Each rank writes OUTDIR/{u,v,d,x,kv}.r<rank>.bin (owned rows, float64) and
// prints one line "rank R owned_begin B n_owned N ..." with the checks it made.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "amat_amd/aMatFree.hpp"

using DT = double;
using GI = unsigned long;
using LI = unsigned int;

static void write_bin(const std::string& path, const std::vector<double>& a)
{
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(a.data(), sizeof(double), a.size(), f) != a.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

static double dot(const std::vector<double>& a, const std::vector<double>& b, MPI_Comm comm)
{
    double s = 0.0, t = 0.0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    MPI_Allreduce(&s, &t, 1, MPI_DOUBLE, MPI_SUM, comm);
    return t;
}

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUTDIR [Nex Ney Nez]\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const unsigned Nex = argc > 4 ? std::atoi(argv[2]) : 6, Ney = argc > 4 ? std::atoi(argv[3]) : 5,
                   Nez = argc > 4 ? std::atoi(argv[4]) : 8;
    const unsigned sx = Nex + 1, S = (Nex + 1) * (Ney + 1);
    const GI n_global = (GI)S * (Nez + 1);
    // z-slab of element layers [emin, emax]; node planes (emin, emax+1] owned
    // (rank 0 also owns plane 0); plane emin of rank > 0 is its pre-ghost layer
    const unsigned nzmin = Nez / size, rem = Nez % size;
    const unsigned nel_z = nzmin + ((unsigned)rank < rem ? 1 : 0);
    const unsigned emin = rank * nzmin + std::min<unsigned>(rank, rem), emax = emin + nel_z - 1;
    const GI owned_begin = rank == 0 ? 0 : (GI)S * (emin + 1), owned_end = (GI)S * (emax + 2) - 1;
    const LI n_pre = rank == 0 ? 0 : S;
    const LI n_owned = (LI)(owned_end - owned_begin + 1), n_all = n_pre + n_owned;
    auto local_of = [&](GI g) -> LI { return g < owned_begin ? (LI)(g - (GI)S * emin) : n_pre + (LI)(g - owned_begin); };
    std::vector<GI> l2g(n_all);
    for (LI l = 0; l < n_all; l++) l2g[l] = l < n_pre ? (GI)S * emin + l : owned_begin + (l - n_pre);

    const LI nelem = Nex * Ney * nel_z;
    std::vector<std::vector<LI>> emap(nelem, std::vector<LI>(8));
    std::vector<LI*> emapp(nelem);
    std::vector<LI> dpe(nelem, 8);
    std::vector<double> escale(nelem);
    for (unsigned k = 0; k < nel_z; k++)
        for (unsigned j = 0; j < Ney; j++)
            for (unsigned i = 0; i < Nex; i++) {
                const LI e = Nex * Ney * k + Nex * j + i;
                const GI n0 = i + (GI)j * sx + (GI)(k + emin) * S;
                const GI c[8] = {n0, n0 + 1, n0 + sx + 1, n0 + sx, n0 + S, n0 + S + 1, n0 + S + sx + 1, n0 + S + sx};
                for (int q = 0; q < 8; q++) emap[e][q] = local_of(c[q]);
                emapp[e] = emap[e].data();
                const LI eg = Nex * Ney * (k + emin) + Nex * j + i;  // global element id
                escale[e] = 0.5 + 0.25 * (eg % 7);
            }
    std::vector<GI> cons;
    for (LI l = 0; l < n_all; l++) {
        const GI g = l2g[l];
        const unsigned i = g % sx, j = (g / sx) % (Ney + 1), k = (unsigned)(g / S);
        if (i == 0 || j == 0 || k == 0 || i == Nex || j == Ney || k == Nez) cons.push_back(g);
    }
    std::vector<DT> cvals(cons.size(), 0.0);

    // the reference set-up (example03a.cpp:411-515): Maps on the communicator
    par::Maps<DT, GI, LI> maps(MPI_COMM_WORLD);
    maps.set_map(nelem, emapp.data(), dpe.data(), n_all, l2g.data(), owned_begin, owned_end, n_global);
    maps.set_bdr_map(cons.data(), cvals.data(), (LI)cons.size());
    par::aMatFree<DT, GI, LI> A(maps, par::BC_METH::BC_IMATRIX);
    A.set_matfree_type(par::MATFREE_TYPE::GPU_PURE);
    const double h[3] = {1.0 / Nex, 1.0 / Ney, 1.0 / Nez};
    const double unit[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    double xe[24], kref[64];
    for (int a = 0; a < 8; a++)
        for (int d = 0; d < 3; d++) xe[3 * a + d] = unit[a][d] * h[d];
    if (amx_fem_ke_hex8(xe, 2, kref) != AMX_SUCCESS) return 3;
    struct Ke {
        double a[64];
        long rows() const { return 8; }
        long cols() const { return 8; }
        double operator()(long r, long c) const { return a[r * 8 + c]; }
    } ke;
    for (LI e = 0; e < nelem; e++) {
        for (int q = 0; q < 64; q++) ke.a[q] = escale[e] * kref[q];
        A.set_element_matrix(e, ke, 0, 0, 1);
    }
    A.finalize();

    // ---- par::solve (examples/include/solve.hpp:31-83), step by step
    MPI_Comm comm = A.get_comm();
    Mat M = A.get_matrix();
    auto vec = [&](Vec* v) { VecCreateMPI(comm, n_owned, (PetscInt)n_global, v); };
    Vec u, v, d;
    vec(&u);
    vec(&v);
    vec(&d);
    std::vector<double> uo(n_owned);
    for (LI i = 0; i < n_owned; i++) {
        const GI g = owned_begin + i;
        uo[i] = (double)((long)((g * 7919) % 2001) - 1000) / 1000.0;
    }
    u->a = uo;
    int bad = 0;
    if (MatMult(M, u, v) != 0) bad |= 1;             // KSP's MatMult -> MATOP_MULT -> MatMult_mf
    if (MatGetDiagonal(M, d) != 0) bad |= 2;         // PCJACOBI's MatGetDiagonal
    Mat blk = nullptr;
    const int blk_rc = MatGetDiagonalBlock(M, &blk);  // PCBJACOBI would ask for this
    if (blk_rc != PETSC_ERR_SUP) bad |= 4;
    if (u->borrowed || v->borrowed || d->borrowed) bad |= 8;

    // Jacobi-PCG on the shell (b = 1 on free rows, 0 on constrained rows)
    std::vector<double> b(n_owned), x(n_owned, 0.0), r, z(n_owned), p, dinv(n_owned);
    for (LI i = 0; i < n_owned; i++) {
        b[i] = 1.0;
        dinv[i] = 1.0 / d->a[i];
    }
    for (GI g : cons)
        if (g >= owned_begin && g <= owned_end) b[g - owned_begin] = 0.0;
    r = b;
    for (LI i = 0; i < n_owned; i++) z[i] = dinv[i] * r[i];
    p = z;
    double rz = dot(r, z, comm);
    const double bb = dot(b, b, comm);
    int it = 0;
    Vec pv, qv;
    vec(&pv);
    vec(&qv);
    for (; it < 5000; it++) {
        pv->a = p;
        if (MatMult(M, pv, qv) != 0) {
            bad |= 16;
            break;
        }
        const double alpha = rz / dot(p, qv->a, comm);
        for (LI i = 0; i < n_owned; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * qv->a[i];
        }
        if (dot(r, r, comm) <= 1e-20 * bb) {
            it++;
            break;
        }
        for (LI i = 0; i < n_owned; i++) z[i] = dinv[i] * r[i];
        const double rz2 = dot(r, z, comm);
        for (LI i = 0; i < n_owned; i++) p[i] = z[i] + (rz2 / rz) * p[i];
        rz = rz2;
    }
    // clean-up of par::solve: the shell's context comes back and is deleted
    par::aMatCTX<DT, GI, LI>* ctx = nullptr;
    MatShellGetContext(M, &ctx);
    if (!ctx || ctx->aMatPtr != &A) bad |= 32;
    delete ctx;

    // the host-array matvec on owned-length arrays (isGhosted = false):
    // local_to_ghost, the exchanges, ghost_to_local
    std::vector<double> kv(n_owned);
    A.matvec(kv.data(), uo.data(), false);

    const std::string tag = ".r" + std::to_string(rank) + ".bin";
    write_bin(out + "/u" + tag, uo);
    write_bin(out + "/v" + tag, v->a);
    write_bin(out + "/d" + tag, d->a);
    write_bin(out + "/x" + tag, x);
    write_bin(out + "/kv" + tag, kv);
    std::printf("rank %d owned_begin %lu n_owned %u size %d n_pre %u cg_iterations %d bjacobi_rc %d bad %d\n", rank,
                (unsigned long)owned_begin, n_owned, size, n_pre, it, blk_rc, bad);
    VecDestroy(&u);
    VecDestroy(&v);
    VecDestroy(&d);
    VecDestroy(&pv);
    VecDestroy(&qv);
    MPI_Finalize();
    return bad ? 1 : 0;
}
