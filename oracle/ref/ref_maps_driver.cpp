// ref_maps_driver.cpp -- runs the reference par::Maps on per-rank input files.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
// include/maps.hpp of the reference tree where it lies, run under MPICH's
// mpiexec by tests/golden/make_golden.py to write golden fixtures.  This file
// holds no reference code: it reads inputs, calls Maps::set_map and
// Maps::set_bdr_map (include/maps.hpp:398-451, 828-880) and dumps the getters.
//
// usage: mpiexec -n P ref_maps_driver <dir>   reads <dir>/in_r<rank>.bin,
// writes <dir>/out_r<rank>.bin (layouts in tests/golden/make_golden.py).
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include <mpi.h>

#include "maps.hpp"

template <typename T> static void rd(FILE* f, T* p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); MPI_Abort(MPI_COMM_WORLD, 1); }
}
template <typename T> static void wr(FILE* f, const T* p, size_t n)
{
    if (n) fwrite(p, sizeof(T), n, f);
}
template <typename T> static void wr_vec(FILE* f, const std::vector<T>& v)
{
    uint64_t n = v.size();
    wr(f, &n, 1);
    wr(f, v.data(), v.size());
}

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    std::string dir = argc > 1 ? argv[1] : ".";
    FILE* f = fopen((dir + "/in_r" + std::to_string(rank) + ".bin").c_str(), "rb");
    if (!f) { fprintf(stderr, "no input\n"); MPI_Abort(MPI_COMM_WORLD, 1); }
    uint32_t hdr[3];
    uint64_t rng[3];
    rd(f, hdr, 3);  // n_elem, n_all_dofs, n_constraints
    rd(f, rng, 3);  // owned begin, owned end, n_global
    const uint32_t ne = hdr[0], nall = hdr[1], nc = hdr[2];
    std::vector<unsigned int> dpe(ne);
    rd(f, dpe.data(), ne);
    size_t tot = 0;
    for (auto d : dpe) tot += d;
    std::vector<unsigned int> flat(tot);
    rd(f, flat.data(), tot);
    std::vector<unsigned long> l2g(nall);
    rd(f, l2g.data(), nall);
    std::vector<unsigned long> cd(nc);
    std::vector<double> cv(nc);
    rd(f, cd.data(), nc);
    rd(f, cv.data(), nc);
    fclose(f);

    std::vector<unsigned int*> e2r(ne);
    size_t k = 0;
    for (uint32_t e = 0; e < ne; e++) { e2r[e] = flat.data() + k; k += dpe[e]; }

    par::Maps<double, unsigned long, unsigned int> maps(MPI_COMM_WORLD);
    maps.set_map(ne, e2r.data(), dpe.data(), nall, l2g.data(), rng[0], rng[1], rng[2]);
    maps.set_bdr_map(cd.data(), cv.data(), nc);

    FILE* o = fopen((dir + "/out_r" + std::to_string(rank) + ".bin").c_str(), "wb");
    uint32_t counts[4] = {maps.get_NumPreGhostDofs(), maps.get_NumDofs(), maps.get_NumPostGhostDofs(),
                          maps.get_NumDofsTotal()};
    wr(o, counts, 4);
    std::vector<unsigned int> lmap(tot);
    std::vector<unsigned long> gmap(tot);
    std::vector<unsigned int> bdr(tot);
    std::vector<double> pv(tot);
    k = 0;
    for (uint32_t e = 0; e < ne; e++)
        for (uint32_t i = 0; i < dpe[e]; i++, k++) {
            lmap[k] = maps.get_LocalMap()[e][i];
            gmap[k] = maps.get_Map()[e][i];
            bdr[k] = maps.get_BdrMap()[e][i];
            pv[k] = maps.get_PresValMap()[e][i];
        }
    wr_vec(o, lmap);
    wr_vec(o, gmap);
    std::vector<unsigned long> l2gout(maps.get_Local2Global(), maps.get_Local2Global() + maps.get_NumDofsTotal());
    wr_vec(o, l2gout);
    wr_vec(o, maps.get_localID2Rank());
    wr_vec(o, maps.get_SendDofCounts());
    wr_vec(o, maps.get_SendDofOffset());
    wr_vec(o, maps.get_SendDofIds());
    wr_vec(o, maps.get_SendRankIds());
    wr_vec(o, maps.get_RecvDofCounts());
    wr_vec(o, maps.get_RecvDofOffset());
    wr_vec(o, maps.get_RecvRankIds());
    wr_vec(o, maps.get_independentElem());
    wr_vec(o, maps.get_dependentElem());
    wr_vec(o, bdr);
    wr_vec(o, pv);
    wr_vec(o, maps.get_ownedConstrainedDofs());
    wr_vec(o, maps.get_ownedPrescribedValues());
    wr_vec(o, maps.get_ownedFreeDofs());
    wr_vec(o, maps.get_LocalDofScan());
    fclose(o);
    MPI_Finalize();
    return 0;
}
