/*
 * amx_oracle.c -- CPU restatement of aMat's assembly-free matvec path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product never
 * does.  See amx_oracle.h for what pins each function.
 *
 * Citations are to the paralab/aMat source tree.
 */
#include "amx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------ helpers */

static int cmp_u64(const void* a, const void* b)
{
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

/* std::sort + std::unique + erase on a u64 array; returns the new length. */
static uint32_t sort_unique_u64(uint64_t* a, uint32_t n)
{
    if (n == 0) return 0;
    qsort(a, n, sizeof(uint64_t), cmp_u64);
    uint32_t w = 1;
    for (uint32_t i = 1; i < n; i++)
        if (a[i] != a[w - 1]) a[w++] = a[i];
    return w;
}

/* std::lower_bound on a sorted u64 array */
static uint32_t lower_bound_u64(const uint64_t* a, uint32_t n, uint64_t key)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* m_ulvLocalDofScan: exclusive prefix of the Allgather'ed owned counts
 * (include/maps.hpp:550-557) */
static void dof_scan(uint32_t size, const uint32_t* counts, uint64_t* scan)
{
    scan[0] = 0;
    for (uint32_t p = 1; p < size; p++) scan[p] = scan[p - 1] + counts[p - 1];
}

static uint64_t total_entries(uint32_t n_elem, const uint32_t* dpe)
{
    uint64_t s = 0;
    for (uint32_t e = 0; e < n_elem; e++) s += dpe[e];
    return s;
}

/* ------------------------------------------------------------ maps */

int orc_build_scatter_map(uint32_t rank, uint32_t size, const uint32_t* rank_dof_counts,
                          uint32_t n_elem, const uint32_t* dofs_per_elem,
                          const uint32_t* elem_to_rank_map, uint32_t n_all_dofs,
                          const uint64_t* rank_to_global,
                          uint64_t owned_begin, uint64_t owned_end, uint64_t n_global,
                          uint64_t* global_map, uint32_t* local_map,
                          uint64_t* local2global, uint32_t* localid2rank,
                          uint32_t* n_pre, uint32_t* n_post,
                          uint32_t* ghost_req_counts, uint64_t* ghost_req_gids)
{
    (void)owned_begin;
    (void)n_global;
    const uint64_t n_owned = owned_end - owned_begin + 1; /* maps.hpp:411 */
    const uint64_t total = total_entries(n_elem, dofs_per_elem);
    if (rank_to_global == NULL || elem_to_rank_map == NULL) return ORC_NULL_L2G_MAP;

    /* global map m_ulpMap[e][i] = rank_to_global[element_to_rank_map[e][i]] (:426-434) */
    for (uint64_t k = 0; k < total; k++) global_map[k] = rank_to_global[elem_to_rank_map[k]];

    uint64_t* scan = (uint64_t*)malloc(sizeof(uint64_t) * size);
    dof_scan(size, rank_dof_counts, scan);
    const uint64_t lo = scan[rank], hi = scan[rank] + n_owned;

    /* pre/post ghost candidates (:563-588) */
    uint64_t* pre = (uint64_t*)malloc(sizeof(uint64_t) * (total + 1));
    uint64_t* post = (uint64_t*)malloc(sizeof(uint64_t) * (total + 1));
    uint32_t npre = 0, npost = 0;
    for (uint64_t k = 0; k < total; k++) {
        uint64_t g = global_map[k];
        if (g < lo) pre[npre++] = g;
        else if (g >= hi) post[npost++] = g;
    }
    npre = sort_unique_u64(pre, npre);   /* :591-596 */
    npost = sort_unique_u64(post, npost);

    int err = ORC_SUCCESS;
    const uint64_t n_total = n_owned + npre + npost; /* :615 */
    if (n_total != n_all_dofs) { err = ORC_INDEX_OUT_OF_BOUNDS; goto done; }

    /* owners of pre ghosts (:623-640) */
    uint32_t* pre_owner = (uint32_t*)malloc(sizeof(uint32_t) * (npre + 1));
    uint32_t* post_owner = (uint32_t*)malloc(sizeof(uint32_t) * (npost + 1));
    {
        uint32_t p = 0;
        for (uint32_t g = 0; g < npre; g++) {
            uint64_t gid = pre[g];
            while (p < rank && !(gid >= scan[p] && gid < scan[p] + rank_dof_counts[p])) p++;
            if (p >= size || !(gid >= scan[p] && gid < scan[p] + rank_dof_counts[p])) {
                err = ORC_GHOST_NODE_NOT_FOUND;
                break;
            }
            pre_owner[g] = p;
        }
        /* owners of post ghosts (:642-660) */
        p = rank;
        for (uint32_t g = 0; err == ORC_SUCCESS && g < npost; g++) {
            uint64_t gid = post[g];
            while (p < size && !(gid >= scan[p] && gid < scan[p] + rank_dof_counts[p])) p++;
            if (p >= size) { err = ORC_GHOST_NODE_NOT_FOUND; break; }
            post_owner[g] = p;
        }
    }
    if (err == ORC_SUCCESS) {
        /* m_localid2rank (:662-677) */
        for (uint64_t i = 0; i < n_total; i++) localid2rank[i] = rank;
        for (uint32_t g = 0; g < npre; g++) localid2rank[g] = pre_owner[g];
        for (uint32_t g = 0; g < npost; g++) localid2rank[npre + n_owned + g] = post_owner[g];

        /* per-owner request counts and sendBuf = pre then post (:684-729) */
        for (uint32_t p = 0; p < size; p++) ghost_req_counts[p] = 0;
        for (uint32_t g = 0; g < npre; g++) ghost_req_counts[pre_owner[g]]++;
        for (uint32_t g = 0; g < npost; g++) ghost_req_counts[post_owner[g]]++;
        for (uint32_t g = 0; g < npre; g++) ghost_req_gids[g] = pre[g];
        for (uint32_t g = 0; g < npost; g++) ghost_req_gids[npre + g] = post[g];

        /* local map and local-to-global (:791-817) */
        for (uint64_t k = 0; k < total; k++) {
            uint64_t gid = global_map[k];
            uint32_t lid;
            if (gid >= lo && gid < hi) lid = (uint32_t)(gid - lo + npre);
            else if (gid < lo) lid = lower_bound_u64(pre, npre, gid);
            else lid = (uint32_t)(npre + n_owned) + lower_bound_u64(post, npost, gid);
            local_map[k] = lid;
            local2global[lid] = gid;
        }
        *n_pre = npre;
        *n_post = npost;
    }
    free(pre_owner);
    free(post_owner);
done:
    free(pre);
    free(post);
    free(scan);
    return err;
}

int orc_set_send_ids(uint32_t rank, uint32_t size, const uint32_t* rank_dof_counts,
                     uint32_t n_pre, uint32_t n_recv, const uint64_t* recv_gids,
                     uint32_t* send_dof_ids)
{
    uint64_t* scan = (uint64_t*)malloc(sizeof(uint64_t) * size);
    dof_scan(size, rank_dof_counts, scan);
    int err = ORC_SUCCESS;
    for (uint32_t i = 0; i < n_recv; i++) {
        uint64_t gid = recv_gids[i];
        if (gid < scan[rank] || gid >= scan[rank] + rank_dof_counts[rank]) err = ORC_GHOST_NODE_NOT_FOUND;
        send_dof_ids[i] = (uint32_t)(n_pre + (gid - scan[rank])); /* :763 */
    }
    free(scan);
    return err;
}

int orc_set_bdr_map(uint32_t rank, uint32_t size, const uint32_t* rank_dof_counts,
                    uint32_t n_elem, const uint32_t* dofs_per_elem, const uint64_t* global_map,
                    const uint64_t* constrained, const double* values, uint32_t n_constraints,
                    uint64_t* owned_cdofs, double* owned_cvals, uint32_t* n_owned_c,
                    uint32_t* bdr_map, double* presval_map,
                    uint64_t* owned_free, uint32_t* n_owned_free)
{
    uint64_t* scan = (uint64_t*)malloc(sizeof(uint64_t) * size);
    dof_scan(size, rank_dof_counts, scan);
    const uint64_t lo = scan[rank], hi = scan[rank] + rank_dof_counts[rank];

    /* owned constrained dofs in input order (:833-839) */
    uint32_t nc = 0;
    for (uint32_t i = 0; i < n_constraints; i++) {
        if (constrained[i] >= lo && constrained[i] < hi) {
            owned_cdofs[nc] = constrained[i];
            owned_cvals[nc] = values[i];
            nc++;
        }
    }
    /* elemental masks by linear search, first match wins (:849-869) */
    uint64_t k = 0;
    uint32_t nf = 0;
    for (uint32_t e = 0; e < n_elem; e++) {
        for (uint32_t i = 0; i < dofs_per_elem[e]; i++, k++) {
            uint64_t gid = global_map[k];
            uint32_t idx;
            for (idx = 0; idx < n_constraints; idx++) {
                if (gid == constrained[idx]) {
                    bdr_map[k] = 1;
                    presval_map[k] = values[idx];
                    break;
                }
            }
            if (idx == n_constraints) {
                bdr_map[k] = 0;
                presval_map[k] = -1E16;
                if (gid >= lo && gid < hi) owned_free[nf++] = gid;
            }
        }
    }
    nf = sort_unique_u64(owned_free, nf); /* :871-874 */
    *n_owned_c = nc;
    *n_owned_free = nf;
    free(scan);
    return ORC_SUCCESS;
}

int orc_identify_independent(uint32_t rank, uint32_t size, const uint32_t* rank_dof_counts,
                             uint32_t n_elem, const uint32_t* dofs_per_elem,
                             const uint64_t* global_map,
                             uint32_t* indep, uint32_t* n_indep, uint32_t* dep, uint32_t* n_dep)
{
    uint64_t* scan = (uint64_t*)malloc(sizeof(uint64_t) * size);
    dof_scan(size, rank_dof_counts, scan);
    const uint64_t lo = scan[rank], hi = scan[rank] + rank_dof_counts[rank];
    uint32_t ni = 0, nd = 0;
    uint64_t k = 0;
    for (uint32_t e = 0; e < n_elem; e++) {
        uint32_t did;
        for (did = 0; did < dofs_per_elem[e]; did++) {
            uint64_t gid = global_map[k + did];
            if (gid < lo || gid >= hi) { dep[nd++] = e; break; }
        }
        if (did == dofs_per_elem[e]) indep[ni++] = e;
        k += dofs_per_elem[e];
    }
    *n_indep = ni;
    *n_dep = nd;
    free(scan);
    return ORC_SUCCESS;
}

/* ------------------------------------------------------------ matvec */

static void elem_offsets(uint32_t n_elem, const uint32_t* dpe, uint64_t* moff, uint64_t* koff)
{
    uint64_t m = 0, kk = 0;
    for (uint32_t e = 0; e < n_elem; e++) {
        moff[e] = m;
        koff[e] = kk;
        m += dpe[e];
        kk += (uint64_t)dpe[e] * dpe[e];
    }
}

/* one element: gather (:1858-1862), GEMV (:1994-1999 or :1965-1973), atomic
 * scatter (:2004-2008) */
static inline void elem_apply(uint32_t n, const uint32_t* map, const double* Ke, int layout,
                              const double* u, double* v, double* ue, double* ve)
{
    for (uint32_t c = 0; c < n; c++) {
        ue[c] = u[map[c]];
        ve[c] = 0.0;
    }
    if (layout == 0) {
        for (uint32_t r = 0; r < n; r++)
            for (uint32_t c = 0; c < n; c++) ve[r] += Ke[r * n + c] * ue[c];
    } else {
        for (uint32_t c = 0; c < n; c++) {
            const double alpha = ue[c];
            const double* x = &Ke[c * n];
#pragma omp simd
            for (uint32_t r = 0; r < n; r++) ve[r] += alpha * x[r];
        }
    }
    for (uint32_t r = 0; r < n; r++) {
#pragma omp atomic
        v[map[r]] += ve[r];
    }
}

void orc_matvec_ghosted_omp(uint32_t n_elem, const uint32_t* dofs_per_elem,
                            const uint32_t* local_map, const double* K, int layout,
                            const uint32_t* indep, uint32_t n_indep,
                            const uint32_t* dep, uint32_t n_dep,
                            const double* u, double* v, uint32_t n_total, int nthreads)
{
    uint64_t* moff = (uint64_t*)malloc(sizeof(uint64_t) * (n_elem + 1));
    uint64_t* koff = (uint64_t*)malloc(sizeof(uint64_t) * (n_elem + 1));
    elem_offsets(n_elem, dofs_per_elem, moff, koff);
    uint32_t nmax = 0;
    for (uint32_t e = 0; e < n_elem; e++)
        if (dofs_per_elem[e] > nmax) nmax = dofs_per_elem[e];

    for (uint32_t i = 0; i < n_total; i++) v[i] = 0.0; /* :1789-1792 */
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    int nt = 1;
    (void)nthreads;
#endif
    /* independent elements (:1799-2013), then dependent elements (:2019-2231) */
    for (int phase = 0; phase < 2; phase++) {
        const uint32_t* list = phase == 0 ? indep : dep;
        const uint32_t cnt = phase == 0 ? n_indep : n_dep;
#pragma omp parallel num_threads(nt)
        {
            double* ue = (double*)malloc(sizeof(double) * (nmax + 1));
            double* ve = (double*)malloc(sizeof(double) * (nmax + 1));
#pragma omp for schedule(static)
            for (uint32_t i = 0; i < cnt; i++) {
                uint32_t e = list[i];
                elem_apply(dofs_per_elem[e], local_map + moff[e], K + koff[e], layout, u, v, ue, ve);
            }
            free(ue);
            free(ve);
        }
    }
    free(moff);
    free(koff);
}

void orc_matvec_blocks(uint32_t n_elem, const uint32_t* dofs_per_elem,
                       const uint32_t* blocks_dim, const uint8_t* block_present,
                       const uint32_t* local_map, const double* K,
                       const double* u, double* v, uint32_t n_total)
{
    for (uint32_t i = 0; i < n_total; i++) v[i] = 0.0;
    uint64_t moff = 0, koff = 0, boff = 0;
    double ue[4096], ve[4096];
    for (uint32_t e = 0; e < n_elem; e++) {
        const uint32_t bd = blocks_dim[e];
        const uint32_t nb = dofs_per_elem[e] / bd; /* :1837 */
        const uint32_t* map = local_map + moff;
        for (uint32_t bi = 0; bi < bd; bi++) {
            const uint32_t row_off = bi * nb;
            for (uint32_t bj = 0; bj < bd; bj++) {
                const uint32_t col_off = bj * nb;
                if (!block_present[boff + bi * bd + bj]) continue; /* :1852 */
                const double* Kb = K + koff;
                koff += (uint64_t)nb * nb;
                for (uint32_t c = 0; c < nb; c++) {
                    ue[c] = u[map[col_off + c]];
                    ve[c] = 0.0;
                }
                for (uint32_t r = 0; r < nb; r++)
                    for (uint32_t c = 0; c < nb; c++) ve[r] += Kb[r * nb + c] * ue[c];
                for (uint32_t r = 0; r < nb; r++) v[map[row_off + r]] += ve[r];
            }
        }
        boff += (uint64_t)bd * bd;
        moff += dofs_per_elem[e];
    }
}

void orc_local_to_ghost(uint32_t n_total, uint32_t n_pre, uint32_t n_owned,
                        double* g, const double* local)
{
    for (uint32_t i = 0; i < n_total; i++)
        g[i] = (i >= n_pre && i < n_pre + n_owned) ? local[i - n_pre] : 0.0;
}

void orc_ghost_to_local(uint32_t n_pre, uint32_t n_owned, double* local, const double* g)
{
    for (uint32_t i = 0; i < n_owned; i++) local[i] = g[i + n_pre];
}

void orc_matmult_mf(uint32_t n_elem, const uint32_t* dofs_per_elem,
                    const uint32_t* local_map, const double* K,
                    uint32_t n_total, uint32_t n_pre, uint32_t n_owned,
                    const uint64_t* owned_cdofs, uint32_t n_owned_c, uint64_t dof_scan_rank,
                    const double* uu, double* vv, int nthreads)
{
    double* uug = (double*)malloc(sizeof(double) * n_total);
    double* vvg = (double*)malloc(sizeof(double) * n_total);
    double* Uc = (double*)malloc(sizeof(double) * (n_owned_c + 1));
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n_elem + 1));
    for (uint32_t e = 0; e < n_elem; e++) order[e] = e;

    orc_local_to_ghost(n_total, n_pre, n_owned, uug, uu); /* :3913 */
    for (uint32_t i = 0; i < n_owned_c; i++) {            /* :3916-3922 */
        uint64_t lid = owned_cdofs[i] - dof_scan_rank + n_pre;
        Uc[i] = uug[lid];
        uug[lid] = 0.0;
    }
    orc_matvec_ghosted_omp(n_elem, dofs_per_elem, local_map, K, 0, order, n_elem, NULL, 0,
                           uug, vvg, n_total, nthreads); /* :3926 */
    for (uint32_t i = 0; i < n_owned_c; i++) {            /* :3929-3933 */
        uint64_t lid = owned_cdofs[i] - dof_scan_rank + n_pre;
        vvg[lid] = Uc[i];
    }
    orc_ghost_to_local(n_pre, n_owned, vv, vvg); /* :3936 */
    free(uug);
    free(vvg);
    free(Uc);
    free(order);
}

typedef struct {
    uint32_t row, col;
    uint64_t seq;
    double val;
} coo_t;

static int cmp_coo(const void* a, const void* b)
{
    const coo_t* x = (const coo_t*)a;
    const coo_t* y = (const coo_t*)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return (x->seq > y->seq) - (x->seq < y->seq);
}

void orc_matmult_penalty(uint32_t n_elem, const uint32_t* dofs_per_elem,
                         const uint32_t* local_map, const double* K,
                         const uint32_t* bdr_map, double penalty,
                         uint32_t n_total, const double* u, double* v)
{
    uint64_t nnz_in = 0;
    for (uint32_t e = 0; e < n_elem; e++) nnz_in += (uint64_t)dofs_per_elem[e] * dofs_per_elem[e];
    coo_t* coo = (coo_t*)malloc(sizeof(coo_t) * (nnz_in + 1));
    /* ADD_VALUES assembly in element, row, column order
     * (aMatBased::petsc_set_element_matrix, include/aMatBased.hpp:212-259) */
    uint64_t k = 0, moff = 0, koff = 0;
    for (uint32_t e = 0; e < n_elem; e++) {
        uint32_t n = dofs_per_elem[e];
        for (uint32_t r = 0; r < n; r++)
            for (uint32_t c = 0; c < n; c++, k++) {
                coo[k].row = local_map[moff + r];
                coo[k].col = local_map[moff + c];
                coo[k].seq = k;
                coo[k].val = K[koff + (uint64_t)r * n + c];
            }
        moff += n;
        koff += (uint64_t)n * n;
    }
    qsort(coo, nnz_in, sizeof(coo_t), cmp_coo);
    /* compress duplicates: sum in insertion order */
    uint64_t* rowptr = (uint64_t*)calloc(n_total + 1, sizeof(uint64_t));
    uint32_t* cols = (uint32_t*)malloc(sizeof(uint32_t) * (nnz_in + 1));
    double* vals = (double*)malloc(sizeof(double) * (nnz_in + 1));
    uint64_t nnz = 0;
    for (uint64_t i = 0; i < nnz_in;) {
        uint64_t j = i;
        double s = 0.0;
        while (j < nnz_in && coo[j].row == coo[i].row && coo[j].col == coo[i].col) s += coo[j++].val;
        cols[nnz] = coo[i].col;
        vals[nnz] = s;
        rowptr[coo[i].row + 1]++;
        nnz++;
        i = j;
    }
    for (uint32_t r = 0; r < n_total; r++) rowptr[r + 1] += rowptr[r];
    /* penalty branch of apply_bc_mat (include/aMatBased.hpp:332-347) */
    moff = 0;
    for (uint32_t e = 0; e < n_elem; e++) {
        uint32_t n = dofs_per_elem[e];
        for (uint32_t r = 0; r < n; r++) {
            if (bdr_map[moff + r] != 1) continue;
            uint32_t row = local_map[moff + r];
            for (uint32_t c = 0; c < n; c++) {
                if (local_map[moff + c] != row) continue;
                for (uint64_t q = rowptr[row]; q < rowptr[row + 1]; q++)
                    if (cols[q] == row) vals[q] = penalty;
            }
        }
        moff += n;
    }
    for (uint32_t r = 0; r < n_total; r++) {
        double s = 0.0;
        for (uint64_t q = rowptr[r]; q < rowptr[r + 1]; q++) s += vals[q] * u[cols[q]];
        v[r] = s;
    }
    free(coo);
    free(rowptr);
    free(cols);
    free(vals);
}

double orc_trace(uint32_t n_elem, const uint32_t* dofs_per_elem, const double* K)
{
    double t = 0.0;
    uint64_t koff = 0;
    for (uint32_t e = 0; e < n_elem; e++) {
        uint32_t n = dofs_per_elem[e];
        for (uint32_t r = 0; r < n; r++) t += K[koff + (uint64_t)r * n + r];
        koff += (uint64_t)n * n;
    }
    return t;
}

void orc_get_diagonal(uint32_t n_elem, const uint32_t* dofs_per_elem,
                      const uint32_t* local_map, const double* K,
                      const uint32_t* bdr_map, int bc_meth, double penalty_override,
                      uint32_t n_total, double* diag)
{
    for (uint32_t i = 0; i < n_total; i++) diag[i] = 0.0;
    uint64_t moff = 0, koff = 0;
    for (uint32_t e = 0; e < n_elem; e++) { /* :1016-1050 */
        uint32_t n = dofs_per_elem[e];
        for (uint32_t r = 0; r < n; r++) diag[local_map[moff + r]] += K[koff + (uint64_t)r * n + r];
        moff += n;
        koff += (uint64_t)n * n;
    }
    double pval = penalty_override > 0.0 ? penalty_override
                                         : ORC_PENALTY_FACTOR * orc_trace(n_elem, dofs_per_elem, K);
    if (bdr_map == NULL) return;
    moff = 0;
    for (uint32_t e = 0; e < n_elem; e++) { /* :4318-4333 */
        uint32_t n = dofs_per_elem[e];
        for (uint32_t r = 0; r < n; r++)
            if (bdr_map[moff + r] == 1) diag[local_map[moff + r]] = bc_meth == 0 ? 1.0 : pval;
        moff += n;
    }
}

void orc_build_ew_maps(uint32_t n_elem, const uint32_t* dofs_per_elem,
                       const uint32_t* local_map, uint32_t n_total,
                       uint32_t* offsets, uint32_t* positions)
{
    uint64_t total = total_entries(n_elem, dofs_per_elem);
    for (uint32_t d = 0; d <= n_total; d++) offsets[d] = 0;
    for (uint64_t k = 0; k < total; k++) offsets[local_map[k] + 1]++;
    for (uint32_t d = 0; d < n_total; d++) offsets[d + 1] += offsets[d];
    uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * (n_total + 1));
    memcpy(fill, offsets, sizeof(uint32_t) * (n_total + 1));
    /* push_back in element order (:378-396) */
    for (uint64_t k = 0; k < total; k++) positions[fill[local_map[k]]++] = (uint32_t)k;
    free(fill);
}

void orc_gather_ew(uint32_t n_total, const uint32_t* offsets, const uint32_t* positions,
                   const double* ve, double* v)
{
    for (uint32_t d = 0; d < n_total; d++)
        for (uint32_t q = offsets[d]; q < offsets[d + 1]; q++) v[d] += ve[positions[q]];
}

void orc_element_products(uint32_t n_elem, const uint32_t* dofs_per_elem,
                          const uint32_t* local_map, const double* K,
                          const double* u, double* ve)
{
    uint64_t moff = 0, koff = 0;
    for (uint32_t e = 0; e < n_elem; e++) {
        uint32_t n = dofs_per_elem[e];
        for (uint32_t r = 0; r < n; r++) {
            double s = 0.0;
            for (uint32_t c = 0; c < n; c++) s += K[koff + (uint64_t)r * n + c] * u[local_map[moff + c]];
            ve[moff + r] = s;
        }
        moff += n;
        koff += (uint64_t)n * n;
    }
}

/* ------------------------------------------------------------ solver */

static double dot(uint32_t n, const double* a, const double* b)
{
    double s = 0.0;
    for (uint32_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int orc_pcg(int op, uint32_t n_elem, const uint32_t* dofs_per_elem,
            const uint32_t* local_map, const double* K, const uint32_t* bdr_map,
            const uint64_t* owned_cdofs, uint32_t n_owned_c, double penalty,
            uint32_t n, const double* b, double* x, double rtol, int max_it,
            double* final_relres)
{
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    double* dinv = (double*)malloc(sizeof(double) * n);
    orc_get_diagonal(n_elem, dofs_per_elem, local_map, K, bdr_map, op == 0 ? 0 : 1,
                     op == 0 ? 0.0 : penalty, n, dinv);
    for (uint32_t i = 0; i < n; i++) dinv[i] = 1.0 / dinv[i];

    for (uint32_t i = 0; i < n; i++) {
        x[i] = 0.0;
        r[i] = b[i];
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
    }
    const double bnorm = sqrt(dot(n, b, b));
    double rz = dot(n, r, z);
    int it = 0;
    double relres = bnorm > 0 ? sqrt(dot(n, r, r)) / bnorm : 0.0;
    while (it < max_it && relres > rtol) {
        if (op == 0)
            orc_matmult_mf(n_elem, dofs_per_elem, local_map, K, n, 0, n, owned_cdofs, n_owned_c, 0,
                           p, q, 1);
        else
            orc_matmult_penalty(n_elem, dofs_per_elem, local_map, K, bdr_map, penalty, n, p, q);
        const double alpha = rz / dot(n, p, q);
        for (uint32_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        it++;
        relres = sqrt(dot(n, r, r)) / bnorm;
        if (relres <= rtol) break;
        for (uint32_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
        const double rz_new = dot(n, r, z);
        const double beta = rz_new / rz;
        rz = rz_new;
        for (uint32_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    if (final_relres) *final_relres = relres;
    free(r);
    free(z);
    free(p);
    free(q);
    free(dinv);
    return it;
}
