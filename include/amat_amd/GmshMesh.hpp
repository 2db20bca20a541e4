/*
 * amat_amd/GmshMesh.hpp -- header-only C++ drop-in for the GmshMesh helper of
 * paralab/aMat's examples (examples/include/GmshMesh.hpp:40-226) over the C
 * ABI of amat_amd.h (amx_gmsh_*).
 *
 * Example code that reads a Gmsh mesh compiles with the same calls
 * (examples/src/example05a.cpp:135-163):
 *
 *   GmshMesh mesh(rank, size, NDOF_PER_NODE);   // (rank, size) instead of MPI_Comm
 *   mesh.setGmshMesh(gmshFile);
 *   unsigned int** localDofMap = mesh.get_localDofMap();
 *   ...
 *   meshMaps.set_map(nelem_owned, localDofMap, nDofsPerElem, nLocalDofs, local2GlobalDofMap,
 *                    startGlobalDof, endGlobalDof, nDofsTotal, mesh.get_rankDofCounts().data());
 *
 * Differences: local node ids are [pre ghosts | owned | post ghosts] with no
 * gap (the reference's get_nLocalDofs counts owned nodes twice,
 * GmshMesh.cpp:541-545), so get_nLocalDofs is the count Maps::set_map checks;
 * prescribed values see each boundary node's own coordinates; failures throw
 * par::amx_error instead of exiting.  Pointers returned by the getters belong
 * to the object and live as long as it does.
 */
#ifndef AMAT_AMD_GMSHMESH_HPP
#define AMAT_AMD_GMSHMESH_HPP

#include <string>
#include <vector>

#include "amat_amd/aMatFree.hpp"

class GmshMesh {
  public:
    GmshMesh(unsigned int rank = 0, unsigned int size = 1, unsigned int nDofPerNode = 1)
        : m_rank(rank), m_size(size), m_ndpn(nDofPerNode)
    {
    }
    ~GmshMesh() { amx_gmsh_destroy(m_h); }
    GmshMesh(const GmshMesh&) = delete;
    GmshMesh& operator=(const GmshMesh&) = delete;

    /* setGmshMesh (examples/src/GmshMesh.cpp:133-331) */
    void setGmshMesh(const std::string& meshFile)
    {
        amx_gmsh_destroy(m_h);
        m_h = nullptr;
        par::detail::check(amx_gmsh_read(meshFile.c_str(), m_rank, m_size, m_ndpn, &m_h), "GmshMesh::setGmshMesh");
        par::detail::check(amx_gmsh_get_info(m_h, &m_info), "GmshMesh::setGmshMesh");
        const unsigned ne = m_info.n_owned_elements, nent = m_info.n_entries, nl = m_info.n_local_nodes;
        m_npe.resize(ne);
        m_ndpe.resize(ne);
        par::detail::check(amx_gmsh_get_element_sizes(m_h, m_npe.data(), m_ndpe.data(), nullptr), "GmshMesh");
        m_lmap.resize(nent);
        m_gmap.resize(nent);
        m_ldof.resize((size_t)nent * m_ndpn);
        m_gdof.resize((size_t)nent * m_ndpn);
        par::detail::check(amx_gmsh_get_maps(m_h, m_lmap.data(), u64(m_gmap.data()), m_ldof.data(), u64(m_gdof.data())),
                           "GmshMesh");
        m_l2g.resize(nl);
        m_l2gdof.resize((size_t)nl * m_ndpn);
        par::detail::check(amx_gmsh_get_local2global(m_h, u64(m_l2g.data()), u64(m_l2gdof.data())), "GmshMesh");
        m_xyz.resize((size_t)nl * 3);
        par::detail::check(amx_gmsh_get_coordinates(m_h, m_xyz.data()), "GmshMesh");
        m_cdofs.resize((size_t)m_info.n_boundary_nodes * m_ndpn);
        par::detail::check(amx_gmsh_get_constraints(m_h, nullptr, u64(m_cdofs.data())), "GmshMesh");
        m_counts.resize(m_size);
        par::detail::check(amx_gmsh_get_rank_node_counts(m_h, m_counts.data()), "GmshMesh");
        for (auto& c : m_counts) c *= m_ndpn;
        rows(m_lmap, m_npe, 1, m_lrows);
        rows(m_gmap, m_npe, 1, m_grows);
        rows(m_ldof, m_npe, m_ndpn, m_ldrows);
        rows(m_gdof, m_npe, m_ndpn, m_gdrows);
    }

    unsigned int get_nOwnedElems() const { return m_info.n_owned_elements; }
    unsigned int get_nOwnedDofs() const { return m_ndpn * m_info.n_owned_nodes; }
    unsigned int* get_nNodesPerElem() { return m_npe.data(); }
    unsigned int** get_localMap() { return m_lrows.data(); }
    unsigned long** get_globalMap() { return m_grows.data(); }
    unsigned long** get_globalDofMap() { return m_gdrows.data(); }
    unsigned int** get_localDofMap() { return m_ldrows.data(); }
    unsigned int* get_nDofsPerElem() { return m_ndpe.data(); }
    unsigned int get_nLocalDofs() const { return m_ndpn * m_info.n_local_nodes; }
    unsigned int get_nLocalNodes() const { return m_info.n_local_nodes; }
    unsigned long* get_local2GlobalMap() { return m_l2g.data(); }
    unsigned long* get_local2GlobalDofMap() { return m_l2gdof.data(); }
    unsigned long get_startGlobalDof() const { return m_info.owned_dof_begin; }
    unsigned long get_endGlobalDof() const { return m_info.owned_dof_end; }
    unsigned long get_nDofsTotal() const { return m_info.n_dofs_total; }
    unsigned int get_n_constraints() const { return (unsigned int)m_cdofs.size(); }
    unsigned long* get_constrained_dofs() { return m_cdofs.data(); }
    /* get_prescribed_values (GmshMesh.cpp:752-783) */
    double* get_prescribed_values(void (*coord2Displ)(const double*, double*))
    {
        m_pvals.assign(m_cdofs.size(), 0.0);
        par::detail::check(amx_gmsh_get_prescribed_values(m_h, coord2Displ, m_pvals.data()),
                           "GmshMesh::get_prescribed_values");
        return m_pvals.data();
    }
    double get_x(unsigned int localNid) const { return m_xyz[3 * (size_t)localNid]; }
    double get_y(unsigned int localNid) const { return m_xyz[3 * (size_t)localNid + 1]; }
    double get_z(unsigned int localNid) const { return m_xyz[3 * (size_t)localNid + 2]; }
    /* owned DoF count of every rank: the rank_dof_counts of Maps::set_map when size > 1 */
    const std::vector<unsigned int>& get_rankDofCounts() const { return m_counts; }
    const amx_gmsh_info& info() const { return m_info; }

  private:
    static uint64_t* u64(unsigned long* p) { return reinterpret_cast<uint64_t*>(p); }
    template <typename T>
    static void rows(std::vector<T>& flat, const std::vector<unsigned int>& npe, unsigned int mult, std::vector<T*>& out)
    {
        out.resize(npe.size());
        size_t k = 0;
        for (size_t e = 0; e < npe.size(); e++) {
            out[e] = flat.data() + k;
            k += (size_t)npe[e] * mult;
        }
    }

    unsigned int m_rank, m_size, m_ndpn;
    amx_gmsh* m_h = nullptr;
    amx_gmsh_info m_info{};
    std::vector<unsigned int> m_npe, m_ndpe, m_lmap, m_ldof, m_counts;
    std::vector<unsigned long> m_gmap, m_gdof, m_l2g, m_l2gdof, m_cdofs;
    std::vector<double> m_xyz, m_pvals;
    std::vector<unsigned int*> m_lrows, m_ldrows;
    std::vector<unsigned long*> m_grows, m_gdrows;
};

#endif /* AMAT_AMD_GMSHMESH_HPP */
