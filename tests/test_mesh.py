"""Mesh numbering against the examples' closed forms, tet10 mesh validity,
z-slab partitions, and the counter-based RNG."""
import numpy as np
import pytest

from amat_amd import mesh


def example07a_local_map(Nex, Ney, Nez):
    """The hex20 element-to-node formula of examples/src/example07a.cpp:357-390
    (one rank), restated with numpy."""
    out = np.zeros((Nex * Ney * Nez, 20), dtype=np.int64)
    S = (2 * Nex + 1) * (2 * Ney + 1) - Nex * Ney + (Nex + 1) * (Ney + 1)
    P = (2 * Nex + 1) * (2 * Ney + 1) - Nex * Ney
    for k in range(Nez):
        for j in range(Ney):
            for i in range(Nex):
                e = Nex * Ney * k + Nex * j + i
                m = out[e]
                m[0] = 2 * i + j * (3 * Nex + 2) + k * S
                m[1] = m[0] + 2
                m[3] = m[0] + (3 * Nex + 2)
                m[2] = m[3] + 2
                m[4] = m[0] + S
                m[5] = m[4] + 2
                m[7] = m[4] + (3 * Nex + 2)
                m[6] = m[7] + 2
                m[8] = m[0] + 1
                m[10] = m[3] + 1
                m[11] = m[0] + (2 * Nex + 1) - i
                m[9] = m[11] + 1
                m[12] = m[4] + 1
                m[14] = m[7] + 1
                m[15] = m[4] + (2 * Nex + 1) - i
                m[13] = m[15] + 1
                m[16] = m[0] + (P - i) - j * (2 * Nex + 1)
                m[17] = m[16] + 1
                m[19] = m[16] + (Nex + 1)
                m[18] = m[19] + 1
    return out


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (4, 2, 3)])
def test_hex20_numbering_matches_example07a(dims):
    m = mesh.hex20(*dims)
    np.testing.assert_array_equal(m.elem_nodes, example07a_local_map(*dims))
    nx, ny, nz = dims
    assert m.n_nodes == (nz + 1) * ((2 * nx + 1) * (2 * ny + 1) - nx * ny) + nz * (nx + 1) * (ny + 1)


def test_hex20_node_count_of_c2():
    nx = 64
    assert (nx + 1) * ((2 * nx + 1) ** 2 - nx * nx) + nx * (nx + 1) ** 2 == 1_085_825


def test_hex8_numbering_matches_example03a():
    m = mesh.hex8(3, 4, 2)
    Nex, Ney = 3, 4
    for (i, j, k) in [(0, 0, 0), (2, 3, 1), (1, 2, 0)]:
        e = Nex * Ney * k + Nex * j + i
        g0 = i + j * (Nex + 1) + k * (Nex + 1) * (Ney + 1)   # example03a.cpp:244-252
        exp = [g0, g0 + 1, g0 + (Nex + 1) + 1, g0 + (Nex + 1)]
        exp += [x + (Nex + 1) * (Ney + 1) for x in exp]
        np.testing.assert_array_equal(m.elem_nodes[e], exp)


def test_tet10_mesh_is_valid():
    n = 4
    m = mesh.tet10(n, seed=1, jitter=0.2)
    assert m.n_elem == 6 * n ** 3 and m.n_nodes == (2 * n + 1) ** 3
    v = m.coords[m.elem_nodes[:, :4]]
    det = np.linalg.det(np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=1))
    assert (det > 0).all()
    np.testing.assert_allclose(det.sum() / 6, 1.0, rtol=1e-12)
    # every node is used; mid-edge nodes are edge midpoints
    assert len(np.unique(m.elem_nodes)) == m.n_nodes
    c = m.coords[m.elem_nodes]
    for b, (i, j) in enumerate([(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]):
        np.testing.assert_allclose(c[:, 4 + b], 0.5 * (c[:, i] + c[:, j]), atol=1e-15)
    # the boundary stays on the cube: (2n+1)^3 - (2n-1)^3 boundary nodes
    assert len(mesh.boundary_nodes(m)) == (2 * n + 1) ** 3 - (2 * n - 1) ** 3
    # conforming: every interior face is shared by exactly two tets
    faces = np.sort(np.concatenate([m.elem_nodes[:, f] for f in ([0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3])]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2 and (cnt == 1).sum() == 2 * 6 * n * n


@pytest.mark.parametrize("kind,P", [("hex8", 3), ("hex20", 2), ("tet10", 2)])
def test_zslab_partition_ranges(kind, P):
    m = {"hex8": lambda: mesh.hex8(3, 3, 5), "hex20": lambda: mesh.hex20(2, 2, 4),
         "tet10": lambda: mesh.tet10(2)}[kind]()
    ranks = mesh.zslab_partition(m, P)
    assert sum(r.n_owned for r in ranks) == m.n_dofs
    assert sum(len(r.elements) for r in ranks) == m.n_elem
    for r in ranks:
        assert r.owned_begin == int(np.sum(r.rank_dof_counts[:r.rank]))
        g = r.rank_to_global_map[r.element_to_rank_map]
        np.testing.assert_array_equal(g, m.dof_map()[r.elements])


def test_uniform_is_counter_based():
    a = mesh.uniform(42, 3, 1000, -1, 1)
    b = mesh.uniform(42, 3, 300, -1, 1, start=500)
    np.testing.assert_array_equal(a[500:800], b)
    assert a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.1
    assert not np.array_equal(a, mesh.uniform(42, 4, 1000, -1, 1))
