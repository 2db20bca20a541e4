"""Gmsh ingest (amx_gmsh_*, amat_amd.gmsh) against the reference's own
GmshMesh (examples/src/GmshMesh.cpp) run on the same files.

The fixtures gmsh_<case>.npz hold an ASCII Gmsh 2.2 file with partition tags
and, per rank, the getters of the reference GmshMesh compiled from its sources
and run under MPICH (tests/golden/make_golden.py).  Global maps, coordinates,
owned ranges and constraints must agree exactly.  Local ids agree on the pre
ghosts and owned nodes; the reference files owned nodes as post ghosts too
(GmshMesh.cpp:541-545), so its post-ghost ids and local counts exceed these by
the owned count, which the tests state.
"""
from __future__ import annotations

import numpy as np
import pytest

from amat_amd import AmatError, Maps, gmsh, mesh
from conftest import GOLDEN
from test_maps import alltoallv

CASES = ["hex8_p1", "hex8_p2_slab", "hex8_p3_scatter", "tet10_p1", "tet10_p2_scatter", "tet10_p4_blocks"]

# Gmsh 2 node orders (Gmsh reference manual, "Node ordering"): mid-edge node
# -> the two corners it lies between
GMSH_TET10_EDGES = {4: (0, 1), 5: (1, 2), 6: (2, 0), 7: (3, 0), 8: (3, 2), 9: (3, 1)}
GMSH_HEX20_EDGES = {8: (0, 1), 9: (0, 3), 10: (0, 4), 11: (1, 2), 12: (1, 5), 13: (2, 3), 14: (2, 6), 15: (3, 7),
                    16: (4, 5), 17: (4, 7), 18: (5, 6), 19: (6, 7)}


def _fixture(case, tmp_path):
    z = np.load(GOLDEN / f"gmsh_{case}.npz")
    path = tmp_path / f"{case}.msh"
    path.write_bytes(z["msh"].tobytes())
    return z, path, int(z["size"][0]), int(z["dofs_per_node"][0])


def _read_elements(path):
    """(type, tags, node labels) of every element and {label: xyz}: a plain
    reading of the file the test wrote."""
    lines = open(path).read().split("\n")
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    xyz = {}
    for ln in lines[i + 2:i + 2 + n]:
        f = ln.split()
        xyz[int(f[0])] = np.array([float(v) for v in f[1:4]])
    i = lines.index("$Elements")
    els = []
    for ln in lines[i + 2:i + 2 + int(lines[i + 1])]:
        f = [int(v) for v in ln.split()]
        nt = f[2]
        els.append((f[1], f[3:3 + nt], f[3 + nt:]))
    return els, xyz


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_reference_gmshmesh(case, tmp_path):
    z, path, P, d = _fixture(case, tmp_path)
    for r in range(P):
        g = gmsh.GmshMesh(path, r, P, d)
        sc = z[f"r{r}_scalars"].astype(np.int64)
        ne, n_owned_dofs, ref_nlocal_dofs, ref_nlocal_nodes, nc, start, end, ntot = sc
        assert g.get_nOwnedElems() == ne and g.get_nOwnedDofs() == n_owned_dofs
        assert (g.get_startGlobalDof(), g.get_endGlobalDof(), g.get_nDofsTotal()) == (start, end, ntot)
        assert g.get_n_constraints() == nc
        np.testing.assert_array_equal(g.get_nNodesPerElem(), z[f"r{r}_npe"])
        np.testing.assert_array_equal(g.get_nDofsPerElem(), z[f"r{r}_ndpe"])
        np.testing.assert_array_equal(np.asarray(g.get_globalMap()).reshape(-1), z[f"r{r}_global_map"])
        np.testing.assert_array_equal(np.asarray(g.get_globalDofMap()).reshape(-1), z[f"r{r}_global_dof_map"])
        lmap = np.asarray(g.get_localMap()).reshape(-1)
        ldof = np.asarray(g.get_localDofMap()).reshape(-1)
        np.testing.assert_array_equal(g.get_local2GlobalMap()[lmap], z[f"r{r}_l2g_at_entries"])
        np.testing.assert_array_equal(g.get_local2GlobalDofMap()[ldof], z[f"r{r}_l2gdof_at_entries"])
        # coordinates of every element node, bit for bit
        np.testing.assert_array_equal(g.element_coordinates().reshape(-1), z[f"r{r}_xyz_at_entries"])
        np.testing.assert_array_equal(g.get_constrained_dofs(), z[f"r{r}_constrained_dofs"])
        # local ids: equal below the post ghosts, shifted by the owned count above
        inf = g.info
        n_pre, n_own = inf.n_pre_ghost_nodes, inf.n_owned_nodes
        shift = np.where(lmap >= n_pre + n_own, n_own, 0)
        np.testing.assert_array_equal(lmap + shift, z[f"r{r}_local_map"])
        dshift = np.where(ldof >= (n_pre + n_own) * d, n_own * d, 0)
        np.testing.assert_array_equal(ldof + dshift, z[f"r{r}_local_dof_map"])
        assert ref_nlocal_nodes == g.get_nLocalNodes() + n_own
        assert ref_nlocal_dofs == g.get_nLocalDofs() + n_own * d
        # prescribed values: displacement(x) = x + 10 y + 100 z + 1000 j
        f = lambda x: [x[0] + 10 * x[1] + 100 * x[2] + 1000.0 * j for j in range(d)]  # noqa: E731
        pv = g.get_prescribed_values(f)
        if d == 3:
            np.testing.assert_array_equal(pv, z[f"r{r}_prescribed"])
        else:
            # the reference reads the coordinates at gmsh id * dofs-per-node
            # (GmshMesh.cpp:765); here they are each boundary node's own
            bn = g.get_boundary_nodes().astype(np.int64)
            xyz = g.coordinates()[np.searchsorted(g.get_local2GlobalMap(), bn)] if len(bn) else np.zeros((0, 3))
            own = np.array([f(x)[0] for x in xyz])
            np.testing.assert_array_equal(pv, own)


@pytest.mark.parametrize("case", CASES)
def test_maps_accept_reader_output(case, tmp_path):
    """Maps::set_map on the reader's maps: its local numbering is the reader's,
    and the ghost exchange lists close (every request has an owner)."""
    z, path, P, d = _fixture(case, tmp_path)
    meshes = [gmsh.GmshMesh(path, r, P, d) for r in range(P)]
    maps = []
    for g in meshes:
        m = Maps(g.rank, g.size)
        m.set_map(*g.set_map_args())
        np.testing.assert_array_equal(m.get_LocalMap().reshape(-1), np.asarray(g.get_localDofMap()).reshape(-1))
        np.testing.assert_array_equal(m.get_Map().reshape(-1), np.asarray(g.get_globalDofMap()).reshape(-1))
        maps.append(m)
    if P > 1:
        for m, (c, gids) in zip(maps, alltoallv([m.get_ghost_requests() for m in maps])):
            m.set_peer_requests(c, gids)
    for g, m in zip(meshes, maps):
        m.set_bdr_map(g.get_constrained_dofs(), np.zeros(g.get_n_constraints()))
    # the reference's local count (owned nodes counted twice) is refused
    g = meshes[-1]
    args = list(g.set_map_args())
    args[3] = int(z[f"r{P - 1}_scalars"][2])
    with pytest.raises(AmatError) as e:
        Maps(g.rank, g.size).set_map(*args)
    assert e.value.name == "INVALID_ARGUMENT"


def test_tet10_nodes_follow_gmsh_order(tmp_path):
    """The written tet10 elements put each mid-edge node at the midpoint of
    the corners the Gmsh manual names, and the reader returns the examples'
    order (mid-edge nodes 8, 9 on edges (1,3), (2,3))."""
    m = mesh.tet10(2, seed=9)
    p = tmp_path / "t.msh"
    gmsh.write_msh(p, m)
    els, xyz = _read_elements(p)
    vols = [e for e in els if e[0] == 11]
    assert len(vols) == m.n_elem
    for _, _, nd in vols:
        for k, (a, b) in GMSH_TET10_EDGES.items():
            np.testing.assert_allclose(xyz[nd[k]], 0.5 * (xyz[nd[a]] + xyz[nd[b]]), rtol=0, atol=1e-15)
    g = gmsh.GmshMesh(p)
    np.testing.assert_array_equal(np.asarray(g.get_globalMap()), m.elem_nodes)
    np.testing.assert_array_equal(g.element_coordinates(), m.coords[m.elem_nodes])


def test_hex20_nodes_follow_gmsh_order(tmp_path):
    """20-node hexahedra (Gmsh type 17, which the reference reader stops on):
    written in Gmsh's order, read back in the examples' order
    (examples/include/shapeFunc.hpp:200-269), quadrangle faces of 8 nodes."""
    m = mesh.hex20(2, 2, 3)
    p = tmp_path / "h.msh"
    parts = (np.arange(m.n_elem) >= 6).astype(int)
    gmsh.write_msh(p, m, parts=parts)
    els, xyz = _read_elements(p)
    vols = [e for e in els if e[0] == 17]
    assert len(vols) == m.n_elem and sum(e[0] == 16 for e in els) == 2 * (4 + 6 + 6)
    for _, _, nd in vols:
        for k, (a, b) in GMSH_HEX20_EDGES.items():
            np.testing.assert_allclose(xyz[nd[k]], 0.5 * (xyz[nd[a]] + xyz[nd[b]]), rtol=0, atol=1e-15)
    g = gmsh.GmshMesh(p, 0, 1, 3)
    np.testing.assert_array_equal(np.asarray(g.get_globalMap()), m.elem_nodes)
    np.testing.assert_array_equal(np.asarray(g.get_globalDofMap()), m.dof_map())
    bn = mesh.boundary_nodes(m)
    np.testing.assert_array_equal(g.get_boundary_nodes(), bn)
    # two ranks: each rank's elements keep their geometry
    for r in range(2):
        gr = gmsh.GmshMesh(p, r, 2, 3)
        np.testing.assert_array_equal(gr.element_coordinates(), m.coords[m.elem_nodes[parts == r]])


def test_node_labels_need_not_be_sequential(tmp_path):
    """Labels in any order and with gaps: nodes are numbered by label."""
    m = mesh.tet10(2, seed=4)
    parts = np.arange(m.n_elem) * 3 // m.n_elem
    a, b = tmp_path / "a.msh", tmp_path / "b.msh"
    gmsh.write_msh(a, m, parts=parts)
    rng = np.random.default_rng(3)
    gmsh.write_msh(b, m, parts=parts, labels=7 + 5 * np.arange(m.n_nodes), node_order=rng.permutation(m.n_nodes))
    for r in range(3):
        ga, gb = gmsh.GmshMesh(a, r, 3, 2), gmsh.GmshMesh(b, r, 3, 2)
        for fn in ("get_globalDofMap", "get_localDofMap", "get_local2GlobalDofMap", "get_constrained_dofs",
                   "coordinates"):
            np.testing.assert_array_equal(np.asarray(getattr(ga, fn)()), np.asarray(getattr(gb, fn)()))
    # labels that reorder the nodes: the geometry of every element is kept
    c = tmp_path / "c.msh"
    gmsh.write_msh(c, m, labels=1 + rng.permutation(m.n_nodes))
    gc = gmsh.GmshMesh(c)
    np.testing.assert_array_equal(gc.element_coordinates(), m.coords[m.elem_nodes])


def test_partitioned_file_reads_on_one_rank(tmp_path):
    m = mesh.hex8(2, 3, 2)
    a, b = tmp_path / "a.msh", tmp_path / "b.msh"
    gmsh.write_msh(a, m)
    gmsh.write_msh(b, m, parts=np.arange(m.n_elem) % 2)
    ga, gb = gmsh.GmshMesh(a, 0, 1, 3), gmsh.GmshMesh(b, 0, 1, 3)
    np.testing.assert_array_equal(np.asarray(ga.get_globalDofMap()), np.asarray(gb.get_globalDofMap()))
    assert ga.info.n_partitions == 1 and gb.info.n_partitions == 2
    assert (gb.info.n_point_elements, gb.info.n_line_elements) == (1, 1)


def test_ownership_invariants(tmp_path):
    """Scattered partitions: every node has one owner, the owned ranges tile
    [0, N), a node belongs to the partition of the first element naming it,
    and every rank's elements keep their geometry."""
    m = mesh.tet10(3, seed=2)
    P = 4
    rng = np.random.default_rng(11)
    parts = rng.integers(0, P, m.n_elem)
    parts[:P] = np.arange(P)
    p = tmp_path / "s.msh"
    gmsh.write_msh(p, m, parts=parts)
    gs = [gmsh.GmshMesh(p, r, P, 1) for r in range(P)]
    counts = gs[0].rank_dof_counts()
    assert counts.sum() == m.n_nodes
    first = np.full(m.n_nodes, -1)
    for e in range(m.n_elem):
        nd = m.elem_nodes[e]
        first[nd[first[nd] < 0]] = parts[e]
    np.testing.assert_array_equal(counts, np.bincount(first, minlength=P))
    begin = 0
    for r, g in enumerate(gs):
        np.testing.assert_array_equal(g.rank_dof_counts(), counts)
        assert (g.get_startGlobalDof(), g.get_endGlobalDof()) == (begin, begin + counts[r] - 1)
        begin += counts[r]
        np.testing.assert_array_equal(g.element_coordinates(), m.coords[m.elem_nodes[parts == r]])
        # owned global ids are the partition's nodes in ascending Gmsh id
        own = g.get_local2GlobalMap()[g.info.n_pre_ghost_nodes:g.info.n_pre_ghost_nodes + g.info.n_owned_nodes]
        xyz = g.coordinates()[g.info.n_pre_ghost_nodes:g.info.n_pre_ghost_nodes + g.info.n_owned_nodes]
        np.testing.assert_array_equal(xyz, m.coords[np.nonzero(first == r)[0]])
        assert (np.diff(own.astype(np.int64)) == 1).all()


def _write(tmp_path, body, fmt="2.2 0 8"):
    p = tmp_path / f"bad{len(list(tmp_path.iterdir()))}.msh"
    p.write_text(f"$MeshFormat\n{fmt}\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n"
                 f"$EndNodes\n$Elements\n{body}$EndElements\n")
    return p


def test_reader_errors(tmp_path):
    tet = "1\n1 4 2 1 1 1 2 3 4\n"
    assert gmsh.GmshMesh(_write(tmp_path, tet)).info.n_owned_elements == 1
    cases = [
        (_write(tmp_path, "1\n1 6 2 1 1 1 2 3 4 1 2\n"), 1, "UNKNOWN_ELEMENT_TYPE"),   # prism
        (_write(tmp_path, tet), 2, "INVALID_ARGUMENT"),                                 # no partition tag
        (_write(tmp_path, "1\n1 4 4 1 1 1 3 1 2 3 4\n"), 2, "INVALID_ARGUMENT"),        # partition 3 of 2
        (_write(tmp_path, tet, fmt="4.1 0 8"), 1, "INVALID_ARGUMENT"),
        (_write(tmp_path, tet, fmt="2.2 1 8"), 1, "INVALID_ARGUMENT"),                  # binary
        (_write(tmp_path, "1\n1 4 2 1 1 1 2 3 9\n"), 1, "INVALID_ARGUMENT"),            # unknown node
        (_write(tmp_path, "1\n1 2 2 1 1 1 2 3\n"), 1, "INVALID_ARGUMENT"),              # no volume
        (_write(tmp_path, "2\n1 2 2 1 1 1 2 3\n"), 1, "INVALID_ARGUMENT"),              # truncated
        (tmp_path / "missing.msh", 1, "INVALID_ARGUMENT"),
    ]
    for path, size, name in cases:
        with pytest.raises(AmatError) as e:
            gmsh.GmshMesh(path, 0, size, 1)
        assert e.value.name == name, (path, str(e.value))
