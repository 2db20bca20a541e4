// amx_maps.cpp -- DoF maps of one rank: the par::Maps interface
// (include/maps.hpp of paralab/aMat) on plain arrays.
//
// The MPI collectives of the reference's buildScatterMap become inputs and
// outputs of the C ABI: the Allgather'ed owned counts are passed in, and the
// ghost-id Alltoallv is left to the caller's transport (get_ghost_requests /
// set_peer_requests).  Outputs are bit-identical to the reference's (checked
// against fixtures written by the reference maps.hpp run under MPICH).
#include <algorithm>
#include <unordered_map>

#include "amx_internal.hpp"

using amx::fail;
using amx::guarded;

namespace {

// Owner of a global id: the rank whose owned range [scan, scan + count)
// contains it, or -1.  Ranges are contiguous and ascending by rank
// (include/maps.hpp:550-557), so a binary search over the scan finds the same
// owner as the reference's linear walk (:623-660).
int64_t owner_of(const amx_maps& m, uint64_t gid)
{
    auto it = std::upper_bound(m.dof_scan.begin(), m.dof_scan.end(), gid);
    if (it == m.dof_scan.begin()) return -1;
    int64_t p = (it - m.dof_scan.begin()) - 1;
    // empty ranks share their scan value with the next rank: step back to a
    // rank that actually holds gid
    while (p >= 0 && gid >= m.dof_scan[p] + m.rank_dof_counts[p]) {
        if (p == 0 || m.dof_scan[p - 1] != m.dof_scan[p]) return -1;
        p--;
    }
    return p;
}

void build_scatter_map(amx_maps& m, const uint32_t* e2r, const uint64_t* r2g, uint32_t n_all)
{
    const uint64_t total = m.eoff.back();
    m.global_map.resize(total);
    for (uint64_t k = 0; k < total; k++) {
        AMX_REQUIRE(e2r[k] < n_all, AMX_INDEX_OUT_OF_BOUNDS,
                    "element_to_rank_map entry " + std::to_string(k) + " exceeds n_all_dofs_on_rank");
        m.global_map[k] = r2g[e2r[k]];
    }
    const uint64_t lo = m.dof_scan[m.rank], hi = lo + m.n_owned;

    // ghosts below / above the owned range, sorted and unique (:563-596)
    m.pre_gids.clear();
    m.post_gids.clear();
    for (uint64_t g : m.global_map) {
        if (g < lo) m.pre_gids.push_back(g);
        else if (g >= hi) m.post_gids.push_back(g);
    }
    for (auto* v : {&m.pre_gids, &m.post_gids}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
    }
    m.n_pre = (uint32_t)m.pre_gids.size();
    m.n_post = (uint32_t)m.post_gids.size();
    const uint64_t n_total = (uint64_t)m.n_owned + m.n_pre + m.n_post;
    AMX_REQUIRE(n_total == n_all, AMX_INVALID_ARGUMENT,
                "n_all_dofs_on_rank (" + std::to_string(n_all) + ") differs from owned + ghost DoFs (" +
                    std::to_string(n_total) + ")");
    m.n_total = (uint32_t)n_total;

    // owners of the ghosts (:618-660) and the per-owner request counts
    m.ghost_owner.resize(m.n_pre + m.n_post);
    m.recv_counts.assign(m.size, 0);
    for (uint32_t i = 0; i < m.n_pre + m.n_post; i++) {
        const uint64_t g = i < m.n_pre ? m.pre_gids[i] : m.post_gids[i - m.n_pre];
        const int64_t p = owner_of(m, g);
        if (p < 0 || (i < m.n_pre && p >= (int64_t)m.rank) || (i >= m.n_pre && p <= (int64_t)m.rank))
            fail(AMX_GHOST_NODE_NOT_FOUND, "ghost global id " + std::to_string(g) + " is owned by no rank");
        m.ghost_owner[i] = (uint32_t)p;
        m.recv_counts[p]++;
    }
    m.recv_offsets.assign(m.size, 0);
    for (uint32_t p = 1; p < m.size; p++) m.recv_offsets[p] = m.recv_offsets[p - 1] + m.recv_counts[p - 1];

    // owner of every local id (:662-677)
    m.localid2rank.assign(m.n_total, m.rank);
    for (uint32_t i = 0; i < m.n_pre; i++) m.localid2rank[i] = m.ghost_owner[i];
    for (uint32_t i = 0; i < m.n_post; i++) m.localid2rank[m.n_pre + m.n_owned + i] = m.ghost_owner[m.n_pre + i];

    // local numbering [pre | owned | post] (:787-817)
    m.local_map.resize(total);
    m.local2global.assign(m.n_total, 0);
    for (uint32_t i = 0; i < m.n_owned; i++) m.local2global[m.n_pre + i] = lo + i;
    for (uint64_t k = 0; k < total; k++) {
        const uint64_t g = m.global_map[k];
        uint32_t lid;
        if (g >= lo && g < hi) {
            lid = (uint32_t)(g - lo) + m.n_pre;
        } else if (g < lo) {
            lid = (uint32_t)(std::lower_bound(m.pre_gids.begin(), m.pre_gids.end(), g) - m.pre_gids.begin());
        } else {
            lid = m.n_pre + m.n_owned +
                  (uint32_t)(std::lower_bound(m.post_gids.begin(), m.post_gids.end(), g) - m.post_gids.begin());
        }
        m.local_map[k] = lid;
        m.local2global[lid] = g;
    }

    // independent / dependent elements (identifyIndependentElements, :918-949)
    m.independent.clear();
    m.dependent.clear();
    for (uint32_t e = 0; e < m.n_elem; e++) {
        bool dep = false;
        for (uint64_t k = m.eoff[e]; k < m.eoff[e + 1]; k++)
            if (m.global_map[k] < lo || m.global_map[k] >= hi) { dep = true; break; }
        (dep ? m.dependent : m.independent).push_back(e);
    }

    m.has_peer_lists = false;
    if (m.size == 1) {
        m.send_counts.assign(1, 0);
        m.send_offsets.assign(1, 0);
        m.send_ids.clear();
        m.send_rank_ids.clear();
        m.recv_rank_ids.clear();
        m.has_peer_lists = true;
    }
}

}  // namespace

extern "C" {

int amx_maps_create(uint32_t rank, uint32_t size, amx_maps** out)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(out != nullptr, AMX_INVALID_ARGUMENT, "out is NULL");
        AMX_REQUIRE(size >= 1 && rank < size, AMX_INVALID_ARGUMENT, "rank must be < size");
        auto* m = new amx_maps();
        m->rank = rank;
        m->size = size;
        *out = m;
    });
}

int amx_maps_destroy(amx_maps* maps)
{
    delete maps;
    return AMX_SUCCESS;
}

int amx_maps_set_map(amx_maps* maps, uint32_t n_elements_on_rank, const uint32_t* element_to_rank_map,
                     const uint32_t* dofs_per_element, uint32_t n_all_dofs_on_rank,
                     const uint64_t* rank_to_global_map, uint64_t owned_begin, uint64_t owned_end,
                     uint64_t n_global_dofs, const uint32_t* rank_dof_counts)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps != nullptr, AMX_INVALID_ARGUMENT, "maps is NULL");
        amx_maps& m = *maps;
        if (n_elements_on_rank > 0) {
            AMX_REQUIRE(element_to_rank_map && rank_to_global_map, AMX_NULL_L2G_MAP,
                        "element_to_rank_map / rank_to_global_map is NULL");
            AMX_REQUIRE(dofs_per_element != nullptr, AMX_INVALID_ARGUMENT, "dofs_per_element is NULL");
        }
        AMX_REQUIRE(owned_end + 1 >= owned_begin, AMX_INVALID_ARGUMENT, "owned range is negative");
        m.n_elem = n_elements_on_rank;
        m.n_owned = (uint32_t)(owned_end - owned_begin + 1);  // :411
        m.owned_begin = owned_begin;
        m.owned_end = owned_end;
        m.dpe.assign(dofs_per_element, dofs_per_element + n_elements_on_rank);
        m.eoff.assign(m.n_elem + 1, 0);
        m.max_dpe = 0;
        for (uint32_t e = 0; e < m.n_elem; e++) {
            AMX_REQUIRE(m.dpe[e] > 0, AMX_INVALID_ARGUMENT, "element " + std::to_string(e) + " has no DoFs");
            m.eoff[e + 1] = m.eoff[e] + m.dpe[e];
            m.max_dpe = std::max(m.max_dpe, m.dpe[e]);
        }
        if (m.size == 1) {
            m.rank_dof_counts.assign(1, m.n_owned);
        } else {
            AMX_REQUIRE(rank_dof_counts != nullptr, AMX_INVALID_ARGUMENT, "rank_dof_counts is NULL for size > 1");
            m.rank_dof_counts.assign(rank_dof_counts, rank_dof_counts + m.size);
            AMX_REQUIRE(m.rank_dof_counts[m.rank] == m.n_owned, AMX_INVALID_ARGUMENT,
                        "rank_dof_counts[rank] differs from the owned range length");
        }
        m.dof_scan.assign(m.size, 0);
        for (uint32_t p = 1; p < m.size; p++) m.dof_scan[p] = m.dof_scan[p - 1] + m.rank_dof_counts[p - 1];
        const uint64_t n_glob = m.dof_scan[m.size - 1] + m.rank_dof_counts[m.size - 1];
        AMX_REQUIRE(n_glob == n_global_dofs, AMX_INVALID_ARGUMENT,
                    "n_global_dofs differs from the sum of rank_dof_counts");
        AMX_REQUIRE(m.dof_scan[m.rank] == owned_begin, AMX_INVALID_ARGUMENT,
                    "owned range does not start at this rank's scan offset");
        m.n_global = n_global_dofs;
        build_scatter_map(m, element_to_rank_map, rank_to_global_map, n_all_dofs_on_rank);
        m.has_map = true;
        m.has_bdr = false;
    });
}

int amx_maps_get_ghost_requests(const amx_maps* maps, uint32_t* counts, uint64_t* gids)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_map, AMX_INVALID_STATE, "set_map has not run");
        for (uint32_t p = 0; p < maps->size; p++) counts[p] = maps->recv_counts[p];
        // sendBuf = pre ghosts then post ghosts (:723-729); both are sorted and
        // their owners ascend, so the buffer is grouped by owner rank
        std::copy(maps->pre_gids.begin(), maps->pre_gids.end(), gids);
        std::copy(maps->post_gids.begin(), maps->post_gids.end(), gids + maps->n_pre);
    });
}

int amx_maps_set_peer_requests(amx_maps* maps, const uint32_t* counts, const uint64_t* gids)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_map, AMX_INVALID_STATE, "set_map has not run");
        amx_maps& m = *maps;
        m.send_counts.assign(counts, counts + m.size);
        m.send_offsets.assign(m.size, 0);
        for (uint32_t p = 1; p < m.size; p++) m.send_offsets[p] = m.send_offsets[p - 1] + m.send_counts[p - 1];
        const uint64_t n = (uint64_t)m.send_offsets[m.size - 1] + m.send_counts[m.size - 1];
        const uint64_t lo = m.dof_scan[m.rank], hi = lo + m.n_owned;
        m.send_ids.resize(n);
        for (uint64_t i = 0; i < n; i++) {
            // :741-768; the reference prints and continues on a foreign id, this returns the error
            AMX_REQUIRE(gids[i] >= lo && gids[i] < hi, AMX_GHOST_NODE_NOT_FOUND,
                        "requested id " + std::to_string(gids[i]) + " is not owned by this rank");
            m.send_ids[i] = m.n_pre + (uint32_t)(gids[i] - lo);
        }
        m.send_rank_ids.clear();
        m.recv_rank_ids.clear();
        for (uint32_t p = 0; p < m.size; p++) {  // :779-785
            if (m.send_counts[p] > 0) m.send_rank_ids.push_back(p);
            if (m.recv_counts[p] > 0) m.recv_rank_ids.push_back(p);
        }
        m.has_peer_lists = true;
    });
}

int amx_maps_set_bdr_map(amx_maps* maps, const uint64_t* constrained_dofs, const double* prescribed_values,
                         uint32_t n_constraints)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_map, AMX_INVALID_STATE, "set_map has not run");
        AMX_REQUIRE(n_constraints == 0 || (constrained_dofs && prescribed_values), AMX_INVALID_ARGUMENT,
                    "constraint arrays are NULL");
        amx_maps& m = *maps;
        const uint64_t lo = m.dof_scan[m.rank], hi = lo + m.n_owned;
        m.owned_cdofs.clear();
        m.owned_cvals.clear();
        for (uint32_t i = 0; i < n_constraints; i++) {  // :833-839
            if (constrained_dofs[i] >= lo && constrained_dofs[i] < hi) {
                m.owned_cdofs.push_back(constrained_dofs[i]);
                m.owned_cvals.push_back(prescribed_values[i]);
            }
        }
        // first index of every constrained id: the reference scans the list
        // per element DoF and stops at the first match (:853-859)
        const uint64_t total = m.eoff.back();
        m.bdr_map.assign(total, 0);
        m.presval_map.assign(total, -1E16);  // :862
        std::vector<uint8_t> free_owned(m.n_owned, 0);
        auto mark = [&](uint64_t k, int64_t idx) {
            if (idx >= 0) {
                m.bdr_map[k] = 1;
                m.presval_map[k] = prescribed_values[idx];
            } else if (m.global_map[k] >= lo && m.global_map[k] < hi) {
                free_owned[m.global_map[k] - lo] = 1;
            }
        };
        if (m.n_global <= (uint64_t(1) << 28)) {
            std::vector<int32_t> first(m.n_global, -1);
            for (uint32_t i = n_constraints; i-- > 0;)
                if (constrained_dofs[i] < m.n_global) first[constrained_dofs[i]] = (int32_t)i;
            for (uint64_t k = 0; k < total; k++) {
                const uint64_t g = m.global_map[k];
                mark(k, g < m.n_global ? first[g] : -1);
            }
        } else {
            std::unordered_map<uint64_t, int64_t> first;
            first.reserve(n_constraints * 2 + 1);
            for (uint32_t i = 0; i < n_constraints; i++) first.emplace(constrained_dofs[i], (int64_t)i);
            for (uint64_t k = 0; k < total; k++) {
                auto it = first.find(m.global_map[k]);
                mark(k, it == first.end() ? -1 : it->second);
            }
        }
        // owned free DoFs sorted and unique (:871-874)
        m.owned_free.clear();
        for (uint32_t i = 0; i < m.n_owned; i++)
            if (free_owned[i]) m.owned_free.push_back(lo + i);
        m.has_bdr = true;
    });
}

int amx_maps_get_info(const amx_maps* maps, amx_maps_info* info)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && info, AMX_INVALID_ARGUMENT, "NULL argument");
        const amx_maps& m = *maps;
        std::memset(info, 0, sizeof(*info));
        info->rank = m.rank;
        info->size = m.size;
        info->n_elems = m.n_elem;
        info->n_owned = m.n_owned;
        info->n_pre_ghost = m.n_pre;
        info->n_post_ghost = m.n_post;
        info->n_total = m.n_total;
        info->n_entries = m.eoff.empty() ? 0 : (uint32_t)m.eoff.back();
        info->max_dofs_per_elem = m.max_dpe;
        info->n_independent = (uint32_t)m.independent.size();
        info->n_dependent = (uint32_t)m.dependent.size();
        info->n_owned_constraints = (uint32_t)m.owned_cdofs.size();
        info->n_owned_free = (uint32_t)m.owned_free.size();
        info->n_send = (uint32_t)m.send_ids.size();
        info->n_recv = m.n_pre + m.n_post;
        info->has_bdr_map = m.has_bdr;
        info->has_peer_lists = m.has_peer_lists;
        info->owned_begin = m.owned_begin;
        info->owned_end = m.owned_end;
        info->n_global = m.n_global;
    });
}

#define AMX_GETTER(name, type, field, need)                                                  \
    int name(const amx_maps* maps, type* out)                                                \
    {                                                                                        \
        return guarded(__func__, [&] {                                                       \
            AMX_REQUIRE(maps && maps->need, AMX_INVALID_STATE, #need " is not set");         \
            std::copy(maps->field.begin(), maps->field.end(), out);                          \
        });                                                                                  \
    }
AMX_GETTER(amx_maps_get_local_map, uint32_t, local_map, has_map)
AMX_GETTER(amx_maps_get_global_map, uint64_t, global_map, has_map)
AMX_GETTER(amx_maps_get_local2global, uint64_t, local2global, has_map)
AMX_GETTER(amx_maps_get_localid2rank, uint32_t, localid2rank, has_map)
AMX_GETTER(amx_maps_get_owned_free, uint64_t, owned_free, has_bdr)
AMX_GETTER(amx_maps_get_dof_scan, uint64_t, dof_scan, has_map)
#undef AMX_GETTER

int amx_maps_get_bdr_map(const amx_maps* maps, uint32_t* mask, double* presval)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_bdr, AMX_INVALID_STATE, "set_bdr_map has not run");
        if (mask) std::copy(maps->bdr_map.begin(), maps->bdr_map.end(), mask);
        if (presval) std::copy(maps->presval_map.begin(), maps->presval_map.end(), presval);
    });
}

int amx_maps_get_owned_constraints(const amx_maps* maps, uint64_t* dofs, double* values)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_bdr, AMX_INVALID_STATE, "set_bdr_map has not run");
        if (dofs) std::copy(maps->owned_cdofs.begin(), maps->owned_cdofs.end(), dofs);
        if (values) std::copy(maps->owned_cvals.begin(), maps->owned_cvals.end(), values);
    });
}

int amx_maps_get_element_lists(const amx_maps* maps, uint32_t* independent, uint32_t* dependent)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_map, AMX_INVALID_STATE, "set_map has not run");
        if (independent) std::copy(maps->independent.begin(), maps->independent.end(), independent);
        if (dependent) std::copy(maps->dependent.begin(), maps->dependent.end(), dependent);
    });
}

int amx_maps_get_send_lists(const amx_maps* maps, uint32_t* counts, uint32_t* offsets, uint32_t* ids)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_peer_lists, AMX_INVALID_STATE, "peer requests have not been set");
        if (counts) std::copy(maps->send_counts.begin(), maps->send_counts.end(), counts);
        if (offsets) std::copy(maps->send_offsets.begin(), maps->send_offsets.end(), offsets);
        if (ids) std::copy(maps->send_ids.begin(), maps->send_ids.end(), ids);
    });
}

int amx_maps_get_recv_lists(const amx_maps* maps, uint32_t* counts, uint32_t* offsets)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(maps && maps->has_map, AMX_INVALID_STATE, "set_map has not run");
        if (counts) std::copy(maps->recv_counts.begin(), maps->recv_counts.end(), counts);
        if (offsets) std::copy(maps->recv_offsets.begin(), maps->recv_offsets.end(), offsets);
    });
}

}  // extern "C"
