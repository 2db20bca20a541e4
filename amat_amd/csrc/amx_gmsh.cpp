// amx_gmsh.cpp -- ASCII Gmsh 2.x meshes as one rank's inputs of par::Maps:
// the GmshMesh helper of paralab/aMat (examples/include/GmshMesh.hpp,
// examples/src/GmshMesh.cpp:133-804), on the host.
//
// Every rank parses the whole file (GmshMesh.hpp:8-10).  A volume element
// belongs to the partition in its tags; a node belongs to the partition of the
// first volume element, in file order, that names it (GmshMesh.cpp:379-466);
// each partition's nodes get consecutive global ids in ascending Gmsh id
// order, partitions in order (:480-495).  Boundary nodes are the nodes of the
// surface elements of the rank's partition (:687-713).
//
// Where this differs from the reference, on purpose:
// * local node ids are [pre ghosts | owned | post ghosts] with no gap.  The
//   reference's build_localMap files every node at or above the rank's offset,
//   owned ones included, as a post ghost (:541-545), so its post-ghost local
//   ids and get_nLocalNodes exceed these by the owned count, and its
//   get_nLocalDofs disagrees with the owned + ghost count Maps::buildScatterMap
//   asserts (include/maps.hpp:606-607);
// * elements are classed by type; the reference takes the last n_volume
//   elements of the file as the volume elements (:300-307), which is the same
//   for files in Gmsh's order (points, lines, surfaces, volumes);
// * tags are read by their count; the reference reads 3 tag fields on one rank
//   and 5 on more (:196-208), so a partitioned file also reads on one rank
//   here;
// * node labels need not be 1..N in order (the reference takes label - 1 as
//   the node's position, :335-345);
// * 20-node hexahedra (type 17, nodes permuted to the examples' hex20 order)
//   and 8-node quadrangles (16) are read; the reference stops on them (:294);
// * prescribed values are evaluated at each boundary node's coordinates; the
//   reference steps through the coordinates by the DoFs-per-node count (:765),
//   the right stride only at 3 DoFs per node.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <numeric>

#include "amx_internal.hpp"

using amx::fail;
using amx::guarded;

namespace {

enum Cls : uint8_t { NONE = 0, POINT, LINE, SURFACE, VOLUME };

struct TypeInfo {
    uint8_t nodes = 0;
    Cls cls = NONE;
};

// Gmsh element types read here (GmshMesh.hpp:28-38 lists the numbering)
TypeInfo type_info(int64_t t)
{
    switch (t) {
    case 15: return {1, POINT};
    case 1: return {2, LINE};
    case 8: return {3, LINE};
    case 2: return {3, SURFACE};
    case 3: return {4, SURFACE};
    case 9: return {6, SURFACE};
    case 10: return {9, SURFACE};
    case 16: return {8, SURFACE};
    case 4: return {4, VOLUME};
    case 5: return {8, VOLUME};
    case 11: return {10, VOLUME};
    case 12: return {27, VOLUME};
    case 17: return {20, VOLUME};
    default: return {};
    }
}

// aMat node a of an element is Gmsh node perm[a].
// tet10: Gmsh orders the mid-edge nodes (0,1)(1,2)(2,0)(3,0)(3,2)(3,1), the
// examples (0,1)(1,2)(2,0)(0,3)(1,3)(2,3) (shapeFunc.hpp:178-199), so nodes 8
// and 9 trade places (GmshMesh.cpp:364-377).
const uint8_t kTet10Perm[10] = {0, 1, 2, 3, 4, 5, 6, 7, 9, 8};
// hex20: Gmsh's mid-edge nodes 8..19 lie on the corner pairs (0,1)(0,3)(0,4)
// (1,2)(1,5)(2,3)(2,6)(3,7)(4,5)(4,7)(5,6)(6,7); the examples' on
// (0,1)(1,2)(2,3)(3,0) (4,5)(5,6)(6,7)(7,4) (0,4)(1,5)(2,6)(3,7)
// (shapeFunc.hpp:200-269).
const uint8_t kHex20Perm[20] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 13, 9, 16, 18, 19, 17, 10, 12, 14, 15};

// Cursor over the file's text (NUL-terminated, so strto* stop at the end).
struct Reader {
    const char* p;
    const char* end;
    const std::string& path;

    [[noreturn]] void bad(const std::string& what) const
    {
        fail(AMX_INVALID_ARGUMENT, path + ": " + what);
    }
    // next line without its end-of-line characters; false at the end
    bool line(std::string& out)
    {
        if (p >= end) return false;
        const char* q = p;
        while (q < end && *q != '\n') q++;
        const char* e = q;
        while (e > p && (e[-1] == '\r' || e[-1] == ' ' || e[-1] == '\t')) e--;
        out.assign(p, e);
        p = q < end ? q + 1 : q;
        return true;
    }
    // moves past the line `tag`; false if there is none
    bool seek(const char* tag)
    {
        std::string s;
        while (line(s))
            if (s == tag) return true;
        return false;
    }
    int64_t integer(const char* what)
    {
        char* e;
        errno = 0;
        const long long v = std::strtoll(p, &e, 10);
        if (e == p || errno != 0) bad(std::string("expected an integer (") + what + ")");
        p = e;
        return v;
    }
    double real(const char* what)
    {
        char* e;
        const double v = std::strtod(p, &e);
        if (e == p) bad(std::string("expected a number (") + what + ")");
        p = e;
        return v;
    }
};

}  // namespace

struct amx_gmsh {
    uint32_t rank = 0, size = 1, dpn = 1;
    uint32_t n_parts = 1;
    uint64_t n_file_nodes = 0, n_claimed = 0;
    uint64_t n_vol = 0, n_surf = 0, n_line = 0, n_point = 0;
    std::vector<double> coords;          // [node position][3], positions in ascending label order
    std::vector<uint64_t> gid;           // node position -> global node id (UINT64_MAX: no volume element)
    std::vector<uint64_t> g2pos;         // global node id -> node position
    std::vector<uint32_t> part_counts;   // owned nodes of every partition (size entries)
    uint64_t offset = 0;                 // first global id of this rank
    // owned elements, file order
    std::vector<uint32_t> npe, types;
    std::vector<uint64_t> eoff;
    std::vector<uint64_t> gmap;          // global node per entry
    std::vector<uint32_t> lmap;          // local node per entry
    std::vector<uint64_t> l2g;           // local node -> global node
    uint32_t n_pre = 0, n_owned = 0, n_post = 0, max_npe = 0;
    std::vector<uint64_t> bdr;           // boundary global node ids, ascending
};

namespace {

void parse(amx_gmsh& m, const std::string& path)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    AMX_REQUIRE(f != nullptr, AMX_INVALID_ARGUMENT, "cannot open " + path);
    std::string text;
    std::fseek(f, 0, SEEK_END);
    const long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (len > 0) {
        text.resize((size_t)len);
        const size_t got = std::fread(&text[0], 1, (size_t)len, f);
        text.resize(got);
    }
    std::fclose(f);
    Reader rd{text.c_str(), text.c_str() + text.size(), path};

    // $MeshFormat: version 2.x, ASCII
    {
        Reader fmt = rd;
        if (fmt.seek("$MeshFormat")) {
            const double version = fmt.real("format version");
            const int64_t file_type = fmt.integer("file type");
            if (version < 2.0 || version >= 3.0)
                rd.bad("Gmsh format " + std::to_string(version) + " is not 2.x (save as version 2 ASCII)");
            if (file_type != 0) rd.bad("binary Gmsh files are not read (save as ASCII)");
        }
    }

    // $Nodes (GmshMesh.cpp:141-158)
    if (!rd.seek("$Nodes")) rd.bad("no $Nodes section");
    const int64_t nn = rd.integer("node count");
    if (nn < 0) rd.bad("negative node count");
    m.n_file_nodes = (uint64_t)nn;
    std::vector<uint64_t> label(m.n_file_nodes);
    std::vector<double> xyz(3 * m.n_file_nodes);
    bool sequential = true;
    for (uint64_t i = 0; i < m.n_file_nodes; i++) {
        const int64_t l = rd.integer("node label");
        if (l < 1) rd.bad("node label " + std::to_string(l) + " is below 1");
        label[i] = (uint64_t)l;
        sequential = sequential && label[i] == i + 1;
        for (int c = 0; c < 3; c++) xyz[3 * i + c] = rd.real("node coordinate");
    }
    // node positions in ascending label order; labels -> positions
    std::vector<uint64_t> sorted_labels;
    if (sequential) {
        m.coords.swap(xyz);
    } else {
        std::vector<uint64_t> order(m.n_file_nodes);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return label[a] < label[b]; });
        sorted_labels.resize(m.n_file_nodes);
        m.coords.resize(3 * m.n_file_nodes);
        for (uint64_t i = 0; i < m.n_file_nodes; i++) {
            sorted_labels[i] = label[order[i]];
            if (i > 0 && sorted_labels[i] == sorted_labels[i - 1])
                rd.bad("node label " + std::to_string(sorted_labels[i]) + " appears twice");
            for (int c = 0; c < 3; c++) m.coords[3 * i + c] = xyz[3 * order[i] + c];
        }
    }
    auto position = [&](int64_t l) -> uint64_t {
        if (sequential) {
            if (l < 1 || (uint64_t)l > m.n_file_nodes) rd.bad("element names unknown node " + std::to_string(l));
            return (uint64_t)l - 1;
        }
        auto it = std::lower_bound(sorted_labels.begin(), sorted_labels.end(), (uint64_t)l);
        if (l < 1 || it == sorted_labels.end() || *it != (uint64_t)l)
            rd.bad("element names unknown node " + std::to_string(l));
        return (uint64_t)(it - sorted_labels.begin());
    };

    // $Elements (GmshMesh.cpp:160-300)
    if (!rd.seek("$Elements")) rd.bad("no $Elements section");
    const int64_t ne = rd.integer("element count");
    if (ne < 0) rd.bad("negative element count");
    // volume elements of every partition, file order
    std::vector<uint32_t> vtype, vpart;
    std::vector<uint64_t> voff{0}, vnodes;
    // surface elements of this rank's partition
    std::vector<uint64_t> snodes;
    std::vector<int64_t> tags;
    uint64_t pos[27];
    for (int64_t e = 0; e < ne; e++) {
        rd.integer("element label");
        const int64_t t = rd.integer("element type");
        const TypeInfo ti = type_info(t);
        if (ti.cls == NONE)
            amx::fail(AMX_UNKNOWN_ELEMENT_TYPE,
                      path + ": element " + std::to_string(e) + " has unsupported Gmsh type " + std::to_string(t));
        const int64_t nt = rd.integer("tag count");
        if (nt < 0 || nt > 64) rd.bad("element " + std::to_string(e) + " has " + std::to_string(nt) + " tags");
        tags.resize((size_t)nt);
        for (int64_t k = 0; k < nt; k++) tags[k] = rd.integer("tag");
        for (int k = 0; k < ti.nodes; k++) pos[k] = position(rd.integer("element node"));
        // partition (1-based): tags are physical, elementary, partition
        // count, owner, negated ghost partitions
        int64_t part = 1;
        if (m.size > 1) {
            if (nt < 4 || tags[2] < 1)
                rd.bad("element " + std::to_string(e) + " has no partition tag (partition the mesh for " +
                       std::to_string(m.size) + " ranks)");
            part = tags[3];
            if (part < 1 || part > (int64_t)m.size)
                rd.bad("element " + std::to_string(e) + " is in partition " + std::to_string(part) + ", not in 1.." +
                       std::to_string(m.size));
        }
        if (nt >= 4 && tags[2] >= 1) m.n_parts = std::max<uint32_t>(m.n_parts, (uint32_t)std::max<int64_t>(tags[3], 1));
        switch (ti.cls) {
        case POINT: m.n_point++; break;
        case LINE: m.n_line++; break;
        case SURFACE:
            m.n_surf++;
            if (part - 1 == (int64_t)m.rank) snodes.insert(snodes.end(), pos, pos + ti.nodes);
            break;
        case VOLUME: {
            m.n_vol++;
            vtype.push_back((uint32_t)t);
            vpart.push_back((uint32_t)(part - 1));
            const uint8_t* perm = t == 11 ? kTet10Perm : t == 17 ? kHex20Perm : nullptr;
            for (int a = 0; a < ti.nodes; a++) vnodes.push_back(pos[perm ? perm[a] : a]);
            voff.push_back(vnodes.size());
            break;
        }
        default: break;
        }
    }
    if (!rd.seek("$EndElements")) rd.bad("no $EndElements line");
    AMX_REQUIRE(m.n_vol > 0, AMX_INVALID_ARGUMENT, path + ": no volume elements");

    // owners: partition of the first volume element naming the node (:398-418)
    std::vector<int32_t> owner(m.n_file_nodes, -1);
    m.part_counts.assign(m.size, 0);
    for (uint64_t e = 0; e < m.n_vol; e++)
        for (uint64_t k = voff[e]; k < voff[e + 1]; k++)
            if (owner[vnodes[k]] < 0) {
                owner[vnodes[k]] = (int32_t)vpart[e];
                m.part_counts[vpart[e]]++;
            }
    // global ids: partitions in order, ascending Gmsh id within one (:480-495)
    std::vector<uint64_t> next(m.size, 0);
    for (uint32_t p = 1; p < m.size; p++) next[p] = next[p - 1] + m.part_counts[p - 1];
    m.offset = next[m.rank];
    m.gid.assign(m.n_file_nodes, UINT64_MAX);
    for (uint64_t i = 0; i < m.n_file_nodes; i++)
        if (owner[i] >= 0) m.gid[i] = next[owner[i]]++;
    m.n_claimed = 0;
    for (uint32_t p = 0; p < m.size; p++) m.n_claimed += m.part_counts[p];
    m.g2pos.assign(m.n_claimed, 0);
    for (uint64_t i = 0; i < m.n_file_nodes; i++)
        if (m.gid[i] != UINT64_MAX) m.g2pos[m.gid[i]] = i;

    // this rank's elements and their global maps (:469-528)
    m.eoff.assign(1, 0);
    for (uint64_t e = 0; e < m.n_vol; e++) {
        if (vpart[e] != m.rank) continue;
        const uint32_t n = (uint32_t)(voff[e + 1] - voff[e]);
        m.npe.push_back(n);
        m.types.push_back(vtype[e]);
        m.max_npe = std::max(m.max_npe, n);
        for (uint64_t k = voff[e]; k < voff[e + 1]; k++) m.gmap.push_back(m.gid[vnodes[k]]);
        m.eoff.push_back(m.gmap.size());
    }
    AMX_REQUIRE(m.gmap.size() < (1ull << 32), AMX_INVALID_ARGUMENT, path + ": more than 2^32 element nodes on a rank");

    // local numbering [pre | owned | post] (the order of :533-594, without the gap)
    const uint64_t lo = m.offset, hi = lo + m.part_counts[m.rank];
    std::vector<uint64_t> pre, post;
    for (uint64_t g : m.gmap) {
        if (g < lo) pre.push_back(g);
        else if (g >= hi) post.push_back(g);
    }
    for (auto* v : {&pre, &post}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
    }
    m.n_pre = (uint32_t)pre.size();
    m.n_owned = m.part_counts[m.rank];
    m.n_post = (uint32_t)post.size();
    m.lmap.resize(m.gmap.size());
    for (size_t k = 0; k < m.gmap.size(); k++) {
        const uint64_t g = m.gmap[k];
        if (g >= lo && g < hi) m.lmap[k] = (uint32_t)(g - lo) + m.n_pre;
        else if (g < lo) m.lmap[k] = (uint32_t)(std::lower_bound(pre.begin(), pre.end(), g) - pre.begin());
        else m.lmap[k] = m.n_pre + m.n_owned + (uint32_t)(std::lower_bound(post.begin(), post.end(), g) - post.begin());
    }
    m.l2g.clear();
    m.l2g.insert(m.l2g.end(), pre.begin(), pre.end());
    for (uint64_t g = lo; g < hi; g++) m.l2g.push_back(g);
    m.l2g.insert(m.l2g.end(), post.begin(), post.end());

    // boundary nodes of this rank's surface elements, sorted and unique (:687-713)
    m.bdr.clear();
    m.bdr.reserve(snodes.size());
    for (uint64_t p : snodes) {
        if (m.gid[p] == UINT64_MAX)
            rd.bad("a surface element names node " + std::to_string(sequential ? p + 1 : sorted_labels[p]) +
                   ", which no volume element has");
        m.bdr.push_back(m.gid[p]);
    }
    std::sort(m.bdr.begin(), m.bdr.end());
    m.bdr.erase(std::unique(m.bdr.begin(), m.bdr.end()), m.bdr.end());
}

}  // namespace

extern "C" {

int amx_gmsh_read(const char* path, uint32_t rank, uint32_t size, uint32_t dofs_per_node, amx_gmsh** out)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(out != nullptr && path != nullptr, AMX_INVALID_ARGUMENT, "path or out is NULL");
        AMX_REQUIRE(size >= 1 && rank < size, AMX_INVALID_ARGUMENT, "rank must be < size");
        AMX_REQUIRE(dofs_per_node >= 1 && dofs_per_node <= 16, AMX_INVALID_ARGUMENT,
                    "dofs_per_node must be in 1..16");
        auto* m = new amx_gmsh();
        m->rank = rank;
        m->size = size;
        m->dpn = dofs_per_node;
        try {
            parse(*m, path);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int amx_gmsh_destroy(amx_gmsh* mesh)
{
    delete mesh;
    return AMX_SUCCESS;
}

int amx_gmsh_get_info(const amx_gmsh* mesh, amx_gmsh_info* info)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh && info, AMX_INVALID_ARGUMENT, "mesh or info is NULL");
        const amx_gmsh& m = *mesh;
        amx_gmsh_info i{};
        i.rank = m.rank;
        i.size = m.size;
        i.dofs_per_node = m.dpn;
        i.n_owned_elements = (uint32_t)m.npe.size();
        i.n_owned_nodes = m.n_owned;
        i.n_pre_ghost_nodes = m.n_pre;
        i.n_post_ghost_nodes = m.n_post;
        i.n_local_nodes = m.n_pre + m.n_owned + m.n_post;
        i.n_entries = (uint32_t)m.gmap.size();
        i.max_nodes_per_element = m.max_npe;
        i.n_boundary_nodes = (uint32_t)m.bdr.size();
        i.n_partitions = m.n_parts;
        i.n_nodes = m.n_claimed;
        i.n_file_nodes = m.n_file_nodes;
        i.n_volume_elements = m.n_vol;
        i.n_surface_elements = m.n_surf;
        i.n_line_elements = m.n_line;
        i.n_point_elements = m.n_point;
        i.node_offset = m.offset;
        i.n_dofs_total = m.n_claimed * m.dpn;
        // an empty rank has begin = end + 1 (Maps::set_map's owned count is end - begin + 1)
        i.owned_dof_begin = m.offset * m.dpn;
        i.owned_dof_end = (m.offset + m.n_owned) * m.dpn - 1;
        *info = i;
    });
}

int amx_gmsh_get_element_sizes(const amx_gmsh* mesh, uint32_t* nodes_per_element, uint32_t* dofs_per_element,
                               uint32_t* gmsh_types)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh != nullptr, AMX_INVALID_ARGUMENT, "mesh is NULL");
        for (size_t e = 0; e < mesh->npe.size(); e++) {
            if (nodes_per_element) nodes_per_element[e] = mesh->npe[e];
            if (dofs_per_element) dofs_per_element[e] = mesh->npe[e] * mesh->dpn;
            if (gmsh_types) gmsh_types[e] = mesh->types[e];
        }
    });
}

int amx_gmsh_get_maps(const amx_gmsh* mesh, uint32_t* local_map, uint64_t* global_map, uint32_t* local_dof_map,
                      uint64_t* global_dof_map)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh != nullptr, AMX_INVALID_ARGUMENT, "mesh is NULL");
        const amx_gmsh& m = *mesh;
        const uint32_t d = m.dpn;
        for (size_t k = 0; k < m.gmap.size(); k++) {
            if (local_map) local_map[k] = m.lmap[k];
            if (global_map) global_map[k] = m.gmap[k];
            for (uint32_t c = 0; c < d; c++) {
                if (local_dof_map) local_dof_map[k * d + c] = m.lmap[k] * d + c;
                if (global_dof_map) global_dof_map[k * d + c] = m.gmap[k] * d + c;
            }
        }
    });
}

int amx_gmsh_get_local2global(const amx_gmsh* mesh, uint64_t* nodes, uint64_t* dofs)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh != nullptr, AMX_INVALID_ARGUMENT, "mesh is NULL");
        const uint32_t d = mesh->dpn;
        for (size_t i = 0; i < mesh->l2g.size(); i++) {
            if (nodes) nodes[i] = mesh->l2g[i];
            if (dofs)
                for (uint32_t c = 0; c < d; c++) dofs[i * d + c] = mesh->l2g[i] * d + c;
        }
    });
}

int amx_gmsh_get_rank_node_counts(const amx_gmsh* mesh, uint32_t* counts)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh && counts, AMX_INVALID_ARGUMENT, "mesh or counts is NULL");
        std::copy(mesh->part_counts.begin(), mesh->part_counts.end(), counts);
    });
}

int amx_gmsh_get_coordinates(const amx_gmsh* mesh, double* xyz)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh && xyz, AMX_INVALID_ARGUMENT, "mesh or xyz is NULL");
        for (size_t i = 0; i < mesh->l2g.size(); i++) {
            const uint64_t p = mesh->g2pos[mesh->l2g[i]];
            for (int c = 0; c < 3; c++) xyz[3 * i + c] = mesh->coords[3 * p + c];
        }
    });
}

int amx_gmsh_get_constraints(const amx_gmsh* mesh, uint64_t* boundary_nodes, uint64_t* constrained_dofs)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh != nullptr, AMX_INVALID_ARGUMENT, "mesh is NULL");
        const uint32_t d = mesh->dpn;
        for (size_t i = 0; i < mesh->bdr.size(); i++) {
            if (boundary_nodes) boundary_nodes[i] = mesh->bdr[i];
            if (constrained_dofs)
                for (uint32_t c = 0; c < d; c++) constrained_dofs[i * d + c] = mesh->bdr[i] * d + c;
        }
    });
}

int amx_gmsh_get_prescribed_values(const amx_gmsh* mesh, void (*displacement)(const double*, double*),
                                   double* values)
{
    return guarded(__func__, [&] {
        AMX_REQUIRE(mesh && displacement && values, AMX_INVALID_ARGUMENT, "mesh, displacement or values is NULL");
        const uint32_t d = mesh->dpn;
        for (size_t i = 0; i < mesh->bdr.size(); i++) {
            const uint64_t p = mesh->g2pos[mesh->bdr[i]];
            displacement(&mesh->coords[3 * p], values + i * d);
        }
    });
}

}  // extern "C"
