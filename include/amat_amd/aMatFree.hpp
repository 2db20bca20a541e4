/*
 * amat_amd/aMatFree.hpp -- header-only C++ drop-in for aMat's par::Maps,
 * par::aMat and par::aMatFree (paralab/aMat include/maps.hpp, aMat.hpp,
 * aMatFree.hpp) over the C ABI of amat_amd.h.
 *
 * Code written against the reference compiles against this header with the
 * same names and arguments for the matrix-free GPU methods 3/4/5:
 *
 *   par::Maps<double, unsigned long, unsigned int> maps(comm_rank, comm_size);
 *   maps.set_map(nelem, localDofMap, ndofs_per_element, numLocalDofs,
 *                local2GlobalDofMap, start_global_dof, end_global_dof, ndofs_total);
 *   maps.set_bdr_map(constrainedDofs, prescribedValues, n_constraints);
 *   par::aMatFree<double, unsigned long, unsigned int> A(maps, par::BC_METH::BC_IMATRIX);
 *   A.set_matfree_type(par::MATFREE_TYPE::GPU_PURE);
 *   A.set_element_matrix(eid, ke, 0, 0, 1);      // any type with rows()/cols()/operator()(r,c)
 *   A.finalize();
 *   A.matvec(v, u, true);
 *
 * Communicator: with MPI available (BUILD_WITH_PETSC, whose petsc.h brings
 * mpi.h, or AMAT_AMD_WITH_MPI) Maps also takes the reference's MPI_Comm;
 * set_map then does buildScatterMap's collectives (MPI_Allgather of the owned
 * counts, MPI_Alltoall/Alltoallv of the ghost requests, include/maps.hpp:547,
 * 704, 738) and aMatFree registers an MPI halo transport (the Isend/Irecv
 * pairs of include/aMatFree.hpp:1426-1679), so matvec, MatMult_mf,
 * get_diagonal and apply_bc exchange ghosts as the reference does.  Without
 * MPI, Maps takes (rank, size) and a rank's share exchanges through a
 * transport the caller registers (amx_op_set_halo_transport).
 * Differences: failures throw par::amx_error carrying the par::Error code
 * instead of exiting; method 1 reports NOT_IMPLEMENTED.  Method 2 (FREE) runs on the GPU
 * in two forms: set_element_function (built-in element kinds recomputed
 * in-kernel from coordinates at every matvec) or the reference's
 * set_element_matrix_function callback, which is host code: it is evaluated
 * once per element at finalize and its matrices are stored (the operator then
 * runs as method 3, with the callback's values).  The PETSc shell hook
 * (get_matrix with the reference's aMatCTX context and its three callbacks,
 * MatMult_mf, MatGetDiagonal_mf, MatGetDiagonalBlock_mf) and apply_bc(Vec)
 * are compiled only with BUILD_WITH_PETSC; par::solve
 * (examples/include/solve.hpp:31-83) runs against them unchanged,
 * except that block Jacobi (MATOP_GET_DIAGONAL_BLOCK) reports
 * PETSC_ERR_SUP.  dump_mat writes "row col value" lines instead of PETSc's
 * MatView.
 */
#ifndef AMAT_AMD_AMATFREE_HPP
#define AMAT_AMD_AMATFREE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "amat_amd.h"

#ifdef BUILD_WITH_PETSC
#include <petsc.h>
#define AMAT_AMD_HAVE_MPI 1
#elif defined(AMAT_AMD_WITH_MPI)
#include <mpi.h>
#define AMAT_AMD_HAVE_MPI 1
#endif

namespace par {

enum class BC_METH { BC_IMATRIX = AMX_BC_IMATRIX, BC_PENALTY = AMX_BC_PENALTY };

enum class Error {
    SUCCESS = 0, INDEX_OUT_OF_BOUNDS, UNKNOWN_ELEMENT_TYPE, UNKNOWN_ELEMENT_STATUS, NULL_L2G_MAP,
    GHOST_NODE_NOT_FOUND, UNKNOWN_MAT_TYPE, WRONG_COMMUNICATION, UNKNOWN_DOF_TYPE, UNKNOWN_CONSTRAINT,
    UNKNOWN_BC_METH, NOT_IMPLEMENTED, GLOBAL_DOF_ID_NOT_FOUND
};

enum class MATFREE_TYPE { UNDEFINED = 0, QUASI_FREE = 1, FREE = 2, GPU_PURE = 3, GPU_OVER_CPU = 4, GPU_OVER_GPU = 5 };

enum class MATRIX_TYPE { MATRIX_BASED, MATRIX_FREE };

struct amx_error : std::runtime_error {
    int code;
    amx_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace detail {
inline void check(int rc, const char* what)
{
    if (rc != AMX_SUCCESS) {
        char buf[512];
        amx_last_error(buf, sizeof(buf));
        throw amx_error(rc, std::string(what) + ": " + buf);
    }
}
template <typename DT, typename GI, typename LI> constexpr void check_types()
{
    static_assert(std::is_same<DT, double>::value, "the GPU operator stores FP64 element matrices (DT = double)");
    static_assert(sizeof(GI) == 8 && std::is_unsigned<GI>::value, "GI must be a 64-bit unsigned type");
    static_assert(sizeof(LI) == 4 && std::is_unsigned<LI>::value, "LI must be a 32-bit unsigned type");
}
}  // namespace detail

template <typename DT, typename GI, typename LI> class Maps {
  public:
    Maps(unsigned int rank = 0, unsigned int size = 1)
    {
        detail::check_types<DT, GI, LI>();
        detail::check(amx_maps_create(rank, size, &m_h), "Maps");
    }
#ifdef AMAT_AMD_HAVE_MPI
    /* Maps::Maps(MPI_Comm) (include/maps.hpp:328-348) */
    explicit Maps(MPI_Comm comm) : m_comm(comm), m_has_comm(true)
    {
        detail::check_types<DT, GI, LI>();
        int rank = 0, size = 1;
        MPI_Comm_rank(comm, &rank);
        MPI_Comm_size(comm, &size);
        detail::check(amx_maps_create((unsigned)rank, (unsigned)size, &m_h), "Maps");
    }
    /* Maps::get_comm (include/maps.hpp:112-115); maps made from (rank, size)
     * answer MPI_COMM_SELF on one rank, MPI_COMM_WORLD otherwise */
    MPI_Comm get_comm() const { return m_has_comm ? m_comm : (info().size == 1 ? MPI_COMM_SELF : MPI_COMM_WORLD); }
    bool has_comm() const { return m_has_comm; }
#endif
    ~Maps() { amx_maps_destroy(m_h); }
    Maps(const Maps&) = delete;
    Maps& operator=(const Maps&) = delete;

    /* Maps::set_map (include/maps.hpp:398-451) */
    Error set_map(const LI n_elements_on_rank, const LI* const* element_to_rank_map, const LI* dofs_per_element,
                  const LI n_all_dofs_on_rank, const GI* rank_to_global_map, const GI owned_global_dof_range_begin,
                  const GI owned_global_dof_range_end, const GI n_global_dofs, const LI* rank_dof_counts = nullptr)
    {
        std::vector<uint32_t> flat;
        for (LI e = 0; e < n_elements_on_rank; e++)
            flat.insert(flat.end(), element_to_rank_map[e], element_to_rank_map[e] + dofs_per_element[e]);
#ifdef AMAT_AMD_HAVE_MPI
        // buildScatterMap's MPI_Allgather of the owned counts (include/maps.hpp:547-548)
        std::vector<LI> counts;
        if (m_has_comm && rank_dof_counts == nullptr && info().size > 1) {
            unsigned own = (unsigned)(owned_global_dof_range_end - owned_global_dof_range_begin + 1);
            counts.resize(info().size);
            MPI_Allgather(&own, 1, MPI_UNSIGNED, counts.data(), 1, MPI_UNSIGNED, m_comm);
            rank_dof_counts = counts.data();
        }
#endif
        detail::check(amx_maps_set_map(m_h, n_elements_on_rank, flat.data(), dofs_per_element, n_all_dofs_on_rank,
                                       reinterpret_cast<const uint64_t*>(rank_to_global_map),
                                       owned_global_dof_range_begin, owned_global_dof_range_end, n_global_dofs,
                                       rank_dof_counts),
                      "Maps::set_map");
        m_dpe.assign(dofs_per_element, dofs_per_element + n_elements_on_rank);
#ifdef AMAT_AMD_HAVE_MPI
        if (m_has_comm && info().size > 1) exchange_ghost_requests();
#endif
        return Error::SUCCESS;
    }

    /* Maps::set_bdr_map (include/maps.hpp:828-880) */
    Error set_bdr_map(GI* constrainedDofs, DT* prescribedValues, LI numConstraints)
    {
        detail::check(amx_maps_set_bdr_map(m_h, reinterpret_cast<const uint64_t*>(constrainedDofs), prescribedValues,
                                           numConstraints),
                      "Maps::set_bdr_map");
        return Error::SUCCESS;
    }

    amx_maps_info info() const
    {
        amx_maps_info i;
        detail::check(amx_maps_get_info(m_h, &i), "Maps::info");
        return i;
    }
    LI get_NumElems() const { return info().n_elems; }
    LI get_NumDofs() const { return info().n_owned; }
    LI get_NumDofsTotal() const { return info().n_total; }
    LI get_NumPreGhostDofs() const { return info().n_pre_ghost; }
    LI get_NumPostGhostDofs() const { return info().n_post_ghost; }
    LI get_n_owned_constraints() const { return info().n_owned_constraints; }
    const LI* get_DofsPerElem() const { return m_dpe.data(); }
    std::vector<LI> get_LocalMapFlat() const
    {
        std::vector<LI> out(info().n_entries);
        detail::check(amx_maps_get_local_map(m_h, out.data()), "Maps::get_LocalMap");
        return out;
    }
    const amx_maps* handle() const { return m_h; }
    /* send/recv lists (include/maps.hpp:741-785): per-rank counts and offsets */
    void get_exchange_lists(std::vector<LI>& send_counts, std::vector<LI>& send_offsets, std::vector<LI>& recv_counts,
                            std::vector<LI>& recv_offsets) const
    {
        const amx_maps_info i = info();
        send_counts.assign(i.size, 0);
        send_offsets.assign(i.size, 0);
        recv_counts.assign(i.size, 0);
        recv_offsets.assign(i.size, 0);
        std::vector<LI> ids(std::max<uint32_t>(i.n_send, 1));
        detail::check(amx_maps_get_send_lists(m_h, send_counts.data(), send_offsets.data(), ids.data()),
                      "Maps::get_SendDofCounts");
        detail::check(amx_maps_get_recv_lists(m_h, recv_counts.data(), recv_offsets.data()), "Maps::get_RecvDofCounts");
    }

  private:
#ifdef AMAT_AMD_HAVE_MPI
    // buildScatterMap's MPI_Alltoall / MPI_Alltoallv of the ghost global ids
    // (include/maps.hpp:700-739): each rank learns which of its DoFs its peers hold
    void exchange_ghost_requests()
    {
        const amx_maps_info i = info();
        const int P = (int)i.size;
        std::vector<uint32_t> sc(P), rc(P);
        std::vector<uint64_t> gids(std::max<uint32_t>(i.n_pre_ghost + i.n_post_ghost, 1));
        detail::check(amx_maps_get_ghost_requests(m_h, sc.data(), gids.data()), "Maps::set_map");
        MPI_Alltoall(sc.data(), 1, MPI_UINT32_T, rc.data(), 1, MPI_UINT32_T, m_comm);
        std::vector<int> scnt(P), sdsp(P), rcnt(P), rdsp(P);
        int st = 0, rt = 0;
        for (int r = 0; r < P; r++) {
            scnt[r] = (int)sc[r];
            sdsp[r] = st;
            st += scnt[r];
            rcnt[r] = (int)rc[r];
            rdsp[r] = rt;
            rt += rcnt[r];
        }
        std::vector<uint64_t> peer(std::max(rt, 1));
        MPI_Alltoallv(gids.data(), scnt.data(), sdsp.data(), MPI_UINT64_T, peer.data(), rcnt.data(), rdsp.data(),
                      MPI_UINT64_T, m_comm);
        detail::check(amx_maps_set_peer_requests(m_h, rc.data(), peer.data()), "Maps::set_map");
    }
    MPI_Comm m_comm = MPI_COMM_SELF;
    bool m_has_comm = false;
#endif
    amx_maps* m_h = nullptr;
    std::vector<LI> m_dpe;
};

/* par::aMat CRTP facade (include/aMat.hpp:75-242): forwards to Derived */
template <typename Derived, typename DT, typename GI, typename LI> class aMat {
  public:
    typedef DT DTType;
    typedef GI GIType;
    typedef LI LIType;
    // virtual: drivers delete the operator through this base (example07a.cpp:1128)
    virtual ~aMat() = default;
    Error set_matfree_type(MATFREE_TYPE t) { return static_cast<Derived*>(this)->set_matfree_type_impl(t); }
    template <typename MatrixType>
    Error set_element_matrix(LI eid, const MatrixType& e_mat, LI block_i, LI block_j, LI blocks_dim)
    {
        return static_cast<Derived*>(this)->set_element_matrix_impl(eid, e_mat, block_i, block_j, blocks_dim);
    }
    Error finalize() { return static_cast<Derived*>(this)->finalize_impl(); }
    Error matvec(DT* v, const DT* u, bool isGhosted = false)
    {
        return static_cast<Derived*>(this)->matvec_impl(v, u, isGhosted);
    }
    // the rest of the reference facade (include/aMat.hpp:114-196), so that
    // drivers holding a par::aMat<par::aMatFree<...>, ...>* (as
    // examples/src/example07a.cpp:730-745 does) compile unchanged
    void set_element_matrix_function(void (*eMat)(LI, DT*, DT*))
    {
        static_cast<Derived*>(this)->set_element_matrix_function(eMat);
    }
    MATRIX_TYPE get_matrix_type() { return static_cast<Derived*>(this)->get_matrix_type(); }
    Error apply_bc(DT* rhs) { return static_cast<Derived*>(this)->apply_bc(rhs); }
    Error finalize_begin() { return static_cast<Derived*>(this)->finalize_begin(); }
    Error finalize_end() { return static_cast<Derived*>(this)->finalize_end(); }
    Error dump_mat(const char* filename = nullptr) { return static_cast<Derived*>(this)->dump_mat(filename); }
    Error set_num_streams(LI nStreams) { return static_cast<Derived*>(this)->set_num_streams(nStreams); }
    Error get_timer(long double* scatterTime, long double* gatherTime, long double* mvTime, long double* mvTotalTime)
    {
        return static_cast<Derived*>(this)->get_timer(scatterTime, gatherTime, mvTime, mvTotalTime);
    }
#ifdef AMAT_AMD_HAVE_MPI
    /* aMat::get_comm (include/aMat.hpp:119-121) */
    MPI_Comm get_comm() { return static_cast<Derived*>(this)->get_comm(); }
#endif
#ifdef BUILD_WITH_PETSC
    Mat& get_matrix() { return static_cast<Derived*>(this)->get_matrix(); }
    Error apply_bc(Vec rhs) { return static_cast<Derived*>(this)->apply_bc(rhs); }
#endif
};

#ifdef BUILD_WITH_PETSC
template <typename DT, typename GI, typename LI> class aMatFree;
/* context of the shell matrix (include/aMatFree.hpp:357-360): allocated by
 * get_matrix, deleted by its user (par::solve, examples/include/solve.hpp:71-80) */
template <typename DT, typename GI, typename LI> struct aMatCTX {
    aMatFree<DT, GI, LI>* aMatPtr;
};
/* shell callbacks (include/aMatFree.hpp:363-401) */
template <typename DT, typename GI, typename LI> PetscErrorCode aMat_matvec(Mat A, Vec u, Vec v)
{
    aMatCTX<DT, GI, LI>* ctx = nullptr;
    MatShellGetContext(A, &ctx);
    return ctx->aMatPtr->MatMult_mf(A, u, v);
}
template <typename DT, typename GI, typename LI> PetscErrorCode aMat_matgetdiagonal(Mat A, Vec d)
{
    aMatCTX<DT, GI, LI>* ctx = nullptr;
    MatShellGetContext(A, &ctx);
    return ctx->aMatPtr->MatGetDiagonal_mf(A, d);
}
template <typename DT, typename GI, typename LI> PetscErrorCode aMat_matgetdiagonalblock(Mat A, Mat* a)
{
    aMatCTX<DT, GI, LI>* ctx = nullptr;
    MatShellGetContext(A, &ctx);
    return ctx->aMatPtr->MatGetDiagonalBlock_mf(A, a);
}
#endif

template <typename DT, typename GI, typename LI> class aMatFree : public aMat<aMatFree<DT, GI, LI>, DT, GI, LI> {
  public:
    /* aMatFree(Maps&, BC_METH) (include/aMatFree.hpp:405-439) */
    aMatFree(Maps<DT, GI, LI>& mesh_maps, BC_METH bcType = BC_METH::BC_IMATRIX, int device = 0)
        : m_maps(mesh_maps)
    {
        detail::check(amx_op_create(mesh_maps.handle(), (int)bcType, device, &m_h), "aMatFree");
#ifdef AMAT_AMD_HAVE_MPI
        if (mesh_maps.has_comm() && mesh_maps.info().size > 1) {
            m_comm = mesh_maps.get_comm();
            mesh_maps.get_exchange_lists(m_sc, m_so, m_rc, m_ro);
            detail::check(amx_op_set_halo_transport(m_h, &aMatFree::mpi_halo, this, AMX_HALO_HOST), "aMatFree");
        }
#endif
    }
    ~aMatFree()
    {
#ifdef BUILD_WITH_PETSC
        if (m_pMat_made) MatDestroy(&m_pMat);  // as ~aMatFree, include/aMatFree.hpp:523
#endif
        amx_op_destroy(m_h);
    }
    aMatFree(const aMatFree&) = delete;
    aMatFree& operator=(const aMatFree&) = delete;

    Error set_matfree_type_impl(MATFREE_TYPE t)
    {
        m_type = t;
        // FREE without a built-in element kind is resolved at finalize (callback form)
        if (t == MATFREE_TYPE::FREE && !m_kind_set) return Error::SUCCESS;
        detail::check(amx_op_set_matfree_type(m_h, (int)t), "set_matfree_type");
        return Error::SUCCESS;
    }
    /* aMat::set_element_matrix_function (include/aMat.hpp:114-116): eMat(eid,
     * ke, xe) fills the element matrix (row-major) and the node coordinates */
    Error set_element_matrix_function(void (*eMat)(LI, DT*, DT*))
    {
        m_eMat = eMat;
        return Error::SUCCESS;
    }
    /* method 2 on the GPU: K_e = scale[e] * ke(xe[e]) recomputed in-kernel for
     * the built-in kinds (amx_element_kind); xe: n_elems x 3 x nodes */
    Error set_element_function(int kind, const DT* xe, const DT* scale, DT E, DT nu, unsigned ngt)
    {
        detail::check(amx_op_set_element_function(m_h, kind, xe, scale, E, nu, ngt), "set_element_function");
        m_kind_set = true;
        m_type = MATFREE_TYPE::FREE;
        return Error::SUCCESS;
    }
    /* aMatFree::set_num_streams (include/aMatFree.hpp:230-237) */
    Error set_num_streams(LI nStreams)
    {
        detail::check(amx_op_set_num_streams(m_h, nStreams), "set_num_streams");
        return Error::SUCCESS;
    }
    /* aMatFree::get_timer (include/aMatFree.hpp:239-245): seconds accumulated
     * since finalize by the host-array matvec / MatMult_mf calls (u copied in,
     * v copied out, the device application, the whole calls) */
    Error get_timer(long double* scatterTime, long double* gatherTime, long double* mvTime, long double* mvTotalTime)
    {
        double s = 0, g = 0, m = 0, t = 0;
        detail::check(amx_op_get_timer(m_h, &s, &g, &m, &t), "get_timer");
        *scatterTime = s;
        *gatherTime = g;
        *mvTime = m;
        *mvTotalTime = t;
        return Error::SUCCESS;
    }
    /* aMatFree::set_element_matrix (include/aMatFree.hpp:854-880); MatrixType
     * needs rows(), cols() and operator()(r, c) (an Eigen matrix works) */
    template <typename MatrixType>
    Error set_element_matrix_impl(LI eid, const MatrixType& e_mat, LI block_i, LI block_j, LI blocks_dim)
    {
        const LI n = (LI)e_mat.rows();
        if ((LI)e_mat.cols() != n) throw amx_error(AMX_INVALID_ARGUMENT, "set_element_matrix: block is not square");
        m_buf.resize((size_t)n * n);
        for (LI r = 0; r < n; r++)
            for (LI c = 0; c < n; c++) m_buf[(size_t)r * n + c] = e_mat(r, c);
        detail::check(amx_op_set_element_matrix(m_h, eid, m_buf.data(), AMX_ROW_MAJOR, block_i, block_j, blocks_dim),
                      "set_element_matrix");
        return Error::SUCCESS;
    }
    using aMat<aMatFree<DT, GI, LI>, DT, GI, LI>::set_element_matrix;
    /* multi-block overload (include/aMatFree.hpp:141-146) */
    template <typename MatrixType>
    Error set_element_matrix(LI eid, LI* ind_non_zero_block_i, LI* ind_non_zero_block_j,
                             const MatrixType** non_zero_block_mats, LI num_non_zero_blocks)
    {
        for (LI b = 0; b < num_non_zero_blocks; b++)
            set_element_matrix_impl(eid, *non_zero_block_mats[b], ind_non_zero_block_i[b], ind_non_zero_block_j[b],
                                    num_non_zero_blocks);
        return Error::SUCCESS;
    }
    Error finalize_impl()
    {
        if (m_type == MATFREE_TYPE::FREE && !m_kind_set) {
            if (!m_eMat)
                throw amx_error(AMX_INVALID_STATE, "finalize: FREE needs set_element_matrix_function or "
                                                   "set_element_function");
            const LI ne = m_maps.get_NumElems();
            const LI* dpe = m_maps.get_DofsPerElem();
            std::vector<DT> ke, xe;
            for (LI eid = 0; eid < ne; eid++) {
                ke.assign((size_t)dpe[eid] * dpe[eid], DT(0));
                xe.assign((size_t)3 * dpe[eid], DT(0));  // at most 3 coordinates per DoF
                m_eMat(eid, ke.data(), xe.data());
                detail::check(amx_op_set_element_matrix(m_h, eid, ke.data(), AMX_ROW_MAJOR, 0, 0, 1),
                              "set_element_matrix_function");
            }
            detail::check(amx_op_set_matfree_type(m_h, AMX_MATFREE_GPU_PURE), "set_matfree_type");
        }
        detail::check(amx_op_finalize(m_h), "finalize");
        return Error::SUCCESS;
    }
    /* aMatFree::finalize_begin / finalize_end (include/aMatFree.hpp:656-785):
     * the whole set-up (device upload, trace) happens in finalize_begin */
    Error finalize_begin() { return finalize_impl(); }
    Error finalize_end()
    {
        if (!info().finalized) throw amx_error(AMX_INVALID_STATE, "finalize_end: finalize_begin was not called");
        return Error::SUCCESS;
    }
    /* aMatFree::allocate_matrix / update_matrix (include/aMatFree.hpp:574-652)
     * size the host work vectors; here finalize sizes every buffer */
    Error allocate_matrix() { return Error::SUCCESS; }
    Error update_matrix() { return Error::SUCCESS; }
    MATRIX_TYPE get_matrix_type() const { return MATRIX_TYPE::MATRIX_FREE; }
    /* aMatFree::dump_mat (include/aMatFree.hpp:4737-4830): the matrix of
     * matvec (no boundary conditions), probed column by column with unit
     * vectors as the reference does, on one rank.  Written as lines
     * "row col value" (global ids, entries with |value| > 1e-16, by row then
     * column, values to 17 digits) to `filename`, or to stdout when it is
     * NULL, where the reference prints PETSc's MatView of the same entries. */
    Error dump_mat(const char* filename = nullptr)
    {
        const amx_maps_info mi = m_maps.info();
        if (mi.size != 1) throw amx_error(AMX_NOT_IMPLEMENTED, "dump_mat: one rank");
        const LI n = mi.n_total;
        struct Entry {
            uint64_t r, c;
            double v;
            bool operator<(const Entry& o) const { return r != o.r ? r < o.r : c < o.c; }
        };
        std::vector<Entry> entries;
        std::vector<DT> ui(n, DT(0)), vi(n, DT(0));
        for (LI j = 0; j < n; j++) {
            ui[j] = DT(1);
            matvec_impl(vi.data(), ui.data(), true);
            ui[j] = DT(0);
            for (LI r = 0; r < n; r++)
                if (std::fabs(vi[r]) > 1e-16) entries.push_back({mi.owned_begin + r, mi.owned_begin + j, vi[r]});
        }
        std::sort(entries.begin(), entries.end());
        std::FILE* f = filename ? std::fopen(filename, "w") : stdout;
        if (!f) throw amx_error(AMX_INVALID_ARGUMENT, std::string("dump_mat: cannot open ") + filename);
        for (const Entry& e : entries)
            std::fprintf(f, "%llu %llu %.17g\n", (unsigned long long)e.r, (unsigned long long)e.c, e.v);
        if (filename) std::fclose(f);
        return Error::SUCCESS;
    }
    /* aMatFree::matvec (include/aMatFree.hpp:1681-1775), host arrays */
    Error matvec_impl(DT* v, const DT* u, bool isGhosted)
    {
        detail::check(amx_op_matvec(m_h, v, u, isGhosted ? 1 : 0), "matvec");
        return Error::SUCCESS;
    }
    /* MatMult_mf semantics (include/aMatFree.hpp:3893-3941), owned-length host arrays */
    Error matmult_mf(DT* vv, const DT* uu)
    {
        detail::check(amx_op_matmult_mf(m_h, vv, uu), "MatMult_mf");
        return Error::SUCCESS;
    }
    /* MatGetDiagonal_mf (include/aMatFree.hpp:3944-3974), owned length */
    Error get_diagonal(DT* diag)
    {
        detail::check(amx_op_get_diagonal(m_h, diag), "get_diagonal");
        return Error::SUCCESS;
    }
    /* apply_bc -> apply_bc_rhs (include/aMatFree.hpp:957-971, 4536-4700), owned-length array */
    Error apply_bc(DT* rhs)
    {
        detail::check(amx_op_apply_bc_rhs(m_h, rhs), "apply_bc");
        return Error::SUCCESS;
    }
    /* device forms: ghosted-length device arrays on a HIP stream */
    Error matvec_dev(DT* v, const DT* u, void* stream = nullptr)
    {
        detail::check(amx_op_matvec_dev(m_h, v, u, stream), "matvec_dev");
        return Error::SUCCESS;
    }
    Error matmult_dev(DT* v, const DT* u, void* stream = nullptr)
    {
        detail::check(amx_op_matmult_dev(m_h, v, u, stream), "matmult_dev");
        return Error::SUCCESS;
    }
    amx_op_info info() const
    {
        amx_op_info i;
        detail::check(amx_op_get_info(m_h, &i), "info");
        return i;
    }
    const Maps<DT, GI, LI>& get_maps() { return m_maps; }
    amx_op* handle() { return m_h; }

#ifdef AMAT_AMD_HAVE_MPI
    /* aMat::get_comm (include/aMat.hpp:119-121): the communicator of the maps */
    MPI_Comm get_comm() { return m_maps.get_comm(); }
#endif
#ifdef BUILD_WITH_PETSC
    /* aMatFree::apply_bc(Vec) (include/aMatFree.hpp:957-971) */
    Error apply_bc(Vec rhs)
    {
        PetscScalar* r;
        VecGetArray(rhs, &r);
        const int rc = amx_op_apply_bc_rhs(m_h, r);
        VecRestoreArray(rhs, &r);
        detail::check(rc, "apply_bc");
        return Error::SUCCESS;
    }
    /* aMatFree::get_matrix (include/aMatFree.hpp:535-570): a MATSHELL over the
     * owned rows whose context is a heap aMatCTX pointing back here (its user
     * deletes it, as par::solve does), with the reference's three operations */
    Mat& get_matrix()
    {
        const LI n = m_maps.get_NumDofs();
        aMatCTX<DT, GI, LI>* ctx = new aMatCTX<DT, GI, LI>();
        ctx->aMatPtr = this;
        if (m_pMat_made) MatDestroy(&m_pMat);
        MatCreateShell(get_comm(), n, n, PETSC_DETERMINE, PETSC_DETERMINE, ctx, &m_pMat);
        m_pMat_made = true;
        MatShellSetOperation(m_pMat, MATOP_MULT, (void (*)(void))aMat_matvec<DT, GI, LI>);
        MatShellSetOperation(m_pMat, MATOP_GET_DIAGONAL, (void (*)(void))aMat_matgetdiagonal<DT, GI, LI>);
        MatShellSetOperation(m_pMat, MATOP_GET_DIAGONAL_BLOCK, (void (*)(void))aMat_matgetdiagonalblock<DT, GI, LI>);
        return m_pMat;
    }
    /* aMatFree::MatMult_mf (include/aMatFree.hpp:3893-3941) on PETSc vectors */
    PetscErrorCode MatMult_mf(Mat, Vec u, Vec v)
    {
        const PetscScalar* uu;
        PetscScalar* vv;
        VecGetArrayRead(u, &uu);
        VecGetArray(v, &vv);
        const int rc = amx_op_matmult_mf(m_h, vv, uu);
        VecRestoreArrayRead(u, &uu);
        VecRestoreArray(v, &vv);
        return rc == AMX_SUCCESS ? 0 : PETSC_ERR_LIB;
    }
    /* aMatFree::MatGetDiagonal_mf (include/aMatFree.hpp:3944-3974) */
    PetscErrorCode MatGetDiagonal_mf(Mat, Vec d)
    {
        PetscScalar* dd;
        VecGetArray(d, &dd);
        const int rc = amx_op_get_diagonal(m_h, dd);
        VecRestoreArray(d, &dd);
        return rc == AMX_SUCCESS ? 0 : PETSC_ERR_LIB;
    }
    /* aMatFree::MatGetDiagonalBlock_mf (include/aMatFree.hpp:3977-4031) feeds
     * PETSc's block Jacobi, which this operator does not provide */
    PetscErrorCode MatGetDiagonalBlock_mf(Mat, Mat*) { return PETSC_ERR_SUP; }

  private:
    Mat m_pMat = nullptr;
    bool m_pMat_made = false;
#endif
#ifdef AMAT_AMD_HAVE_MPI
    // The MPI exchange of ghost_receive_begin/end and ghost_send_begin/end
    // (include/aMatFree.hpp:1426-1679) over host buffers: peer i's part of
    // the owner-side buffer sits at send offset i, of the ghost-side buffer
    // at recv offset i.
    static int mpi_halo(void* ctx, int direction, const double* send, double* recv, void*)
    {
        aMatFree* self = static_cast<aMatFree*>(ctx);
        const int P = (int)self->m_sc.size();
        const bool to_ghost = direction == AMX_HALO_OWNER_TO_GHOST;
        const int tag = to_ghost ? 7141 : 7142;
        std::vector<MPI_Request> req;
        req.reserve(2 * (size_t)P);
        for (int i = 0; i < P; i++) {
            const LI n_in = to_ghost ? self->m_rc[i] : self->m_sc[i];
            const LI at_in = to_ghost ? self->m_ro[i] : self->m_so[i];
            if (n_in) {
                req.emplace_back();
                MPI_Irecv(recv + at_in, (int)n_in, MPI_DOUBLE, i, tag, self->m_comm, &req.back());
            }
        }
        for (int i = 0; i < P; i++) {
            const LI n_out = to_ghost ? self->m_sc[i] : self->m_rc[i];
            const LI at_out = to_ghost ? self->m_so[i] : self->m_ro[i];
            if (n_out) {
                req.emplace_back();
                MPI_Isend(const_cast<double*>(send) + at_out, (int)n_out, MPI_DOUBLE, i, tag, self->m_comm,
                          &req.back());
            }
        }
        const int rc = req.empty() ? MPI_SUCCESS : MPI_Waitall((int)req.size(), req.data(), MPI_STATUSES_IGNORE);
        return rc == MPI_SUCCESS ? 0 : 1;
    }
    MPI_Comm m_comm = MPI_COMM_SELF;
    std::vector<LI> m_sc, m_so, m_rc, m_ro;
#endif

  private:
    Maps<DT, GI, LI>& m_maps;
    amx_op* m_h = nullptr;
    std::vector<double> m_buf;
    MATFREE_TYPE m_type = MATFREE_TYPE::GPU_PURE;
    bool m_kind_set = false;
    void (*m_eMat)(LI, DT*, DT*) = nullptr;
};

}  // namespace par

#endif /* AMAT_AMD_AMATFREE_HPP */
