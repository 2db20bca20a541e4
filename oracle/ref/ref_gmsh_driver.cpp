// ref_gmsh_driver.cpp -- runs the reference GmshMesh on a Gmsh file.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
// examples/src/GmshMesh.cpp of the reference tree where it lies, run under
// MPICH's mpiexec by tests/golden/make_golden.py to write golden fixtures.
// This file holds no reference code: it calls GmshMesh::setGmshMesh and dumps
// the getters the examples use (examples/src/example05a.cpp:135-155).
//
// usage: mpiexec -n P ref_gmsh_driver <mesh.msh> <dofs per node> <dir>
// writes <dir>/gmsh_r<rank>.bin (layout in tests/golden/make_golden.py).
// Local-id arrays of the reference are read only at the ids its element maps
// name: its local numbering leaves ids it never fills (GmshMesh.cpp:541-545).
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include <mpi.h>

#include "GmshMesh.hpp"

static unsigned int g_dpn = 1;

// a prescribed value that tells the coordinates it was evaluated at apart
static void displacement(const double* xyz, double* uvw)
{
    for (unsigned int j = 0; j < g_dpn; j++) uvw[j] = xyz[0] + 10.0 * xyz[1] + 100.0 * xyz[2] + 1000.0 * j;
}

template <typename T> static void wr_vec(FILE* f, const std::vector<T>& v)
{
    uint64_t n = v.size();
    fwrite(&n, sizeof(n), 1, f);
    if (n) fwrite(v.data(), sizeof(T), v.size(), f);
}

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    if (argc < 4) {
        fprintf(stderr, "usage: ref_gmsh_driver mesh.msh dofs_per_node dir\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    g_dpn = (unsigned int)std::stoul(argv[2]);
    const std::string dir = argv[3];

    GmshMesh mesh(MPI_COMM_WORLD, g_dpn);
    mesh.setGmshMesh(argv[1]);
    const unsigned int ne = mesh.get_nOwnedElems();
    unsigned int* npe = mesh.get_nNodesPerElem();
    unsigned int** lmap = mesh.get_localMap();
    unsigned long** gmap = mesh.get_globalMap();
    unsigned int** ldof = mesh.get_localDofMap();
    unsigned long** gdof = mesh.get_globalDofMap();
    unsigned int* ndpe = mesh.get_nDofsPerElem();
    unsigned long* l2g = mesh.get_local2GlobalMap();
    unsigned long* l2gdof = mesh.get_local2GlobalDofMap();
    const unsigned int nc = mesh.get_n_constraints();
    unsigned long* cd = mesh.get_constrained_dofs();
    double* pv = mesh.get_prescribed_values(&displacement);

    std::vector<uint64_t> scalars = {ne,
                                     mesh.get_nOwnedDofs(),
                                     mesh.get_nLocalDofs(),
                                     mesh.get_nLocalNodes(),
                                     nc,
                                     mesh.get_startGlobalDof(),
                                     mesh.get_endGlobalDof(),
                                     mesh.get_nDofsTotal()};
    std::vector<uint32_t> v_npe, v_ndpe, v_lmap, v_ldof;
    std::vector<uint64_t> v_gmap, v_gdof, v_l2g, v_l2gdof, v_cd;
    std::vector<double> v_xyz, v_pv;
    for (unsigned int e = 0; e < ne; e++) {
        v_npe.push_back(npe[e]);
        v_ndpe.push_back(ndpe[e]);
        for (unsigned int n = 0; n < npe[e]; n++) {
            v_lmap.push_back(lmap[e][n]);
            v_gmap.push_back(gmap[e][n]);
            v_l2g.push_back(l2g[lmap[e][n]]);
            v_xyz.push_back(mesh.get_x(lmap[e][n]));
            v_xyz.push_back(mesh.get_y(lmap[e][n]));
            v_xyz.push_back(mesh.get_z(lmap[e][n]));
        }
        for (unsigned int k = 0; k < ndpe[e]; k++) {
            v_ldof.push_back(ldof[e][k]);
            v_gdof.push_back(gdof[e][k]);
            v_l2gdof.push_back(l2gdof[ldof[e][k]]);
        }
    }
    for (unsigned int i = 0; i < nc; i++) {
        v_cd.push_back(cd[i]);
        v_pv.push_back(pv[i]);
    }

    FILE* o = fopen((dir + "/gmsh_r" + std::to_string(rank) + ".bin").c_str(), "wb");
    wr_vec(o, scalars);
    wr_vec(o, v_npe);
    wr_vec(o, v_ndpe);
    wr_vec(o, v_lmap);
    wr_vec(o, v_gmap);
    wr_vec(o, v_ldof);
    wr_vec(o, v_gdof);
    wr_vec(o, v_l2g);
    wr_vec(o, v_l2gdof);
    wr_vec(o, v_xyz);
    wr_vec(o, v_cd);
    wr_vec(o, v_pv);
    fclose(o);
    MPI_Finalize();
    return 0;
}
