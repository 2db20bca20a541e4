"""Structured meshes of the examples and the seeded inputs of the workloads.

* hex8 (trilinear) on an nx x ny x nz grid, numbered as
  examples/src/example03a.cpp:240-255 (node i + j(nx+1) + k(nx+1)(ny+1),
  element i + nx j + nx ny k, corners counter-clockwise bottom then top);
* hex20 (serendipity) numbered as examples/src/example07a.cpp:320-409: nodes
  of the (2nx+1)(2ny+1)(2nz+1) half-step grid minus face and body centres, in
  lexicographic order, 3 DoFs per node, DoF = 3 node + component;
* tet10: every hex of an n^3 grid split into the six Kuhn tetrahedra around its
  main diagonal, interior vertices jittered by U[-j h, j h] per coordinate,
  mid-edge nodes at the midpoints of the (jittered) straight edges.  All
  half-step grid points are then nodes, numbered lexicographically.

Z-slab partitions follow examples/src/example03a.cpp:190-211: rank r gets
element layers [emin, emax] and owns the node planes above its bottom plane
(rank 0 also owns the bottom plane), so every rank's owned global ids form one
contiguous range.

Random inputs come from a counter-based SplitMix64 stream so that any
sub-range can be regenerated on its own (the bench's CPU sample uses this).
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on uint64 values (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64, copy=True)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, lo: float = 0.0, hi: float = 1.0, start: int = 0) -> np.ndarray:
    """x_i = lo + (hi - lo) * u_i, u_i in [0, 1) from 53 bits of
    SplitMix64(seed * G + stream * H + (i + 1) * G), i = start .. start + n - 1."""
    with np.errstate(over="ignore"):
        g = np.uint64(0x9E3779B97F4A7C15)
        h = np.uint64(0xD1B54A32D192ED03)
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        x = np.uint64(seed) * g + np.uint64(stream) * h + i * g
        r = splitmix64(x)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


@dataclasses.dataclass
class Mesh:
    kind: str                  # "hex8" | "hex20" | "tet10"
    dims: tuple                # (nx, ny, nz)
    L: tuple                   # domain lengths
    elem_nodes: np.ndarray     # (E, nodes per element) global node ids
    coords: np.ndarray         # (n_nodes, 3)
    dofs_per_node: int
    node_plane: np.ndarray     # half-step (or grid) z index of every node, for partitioning
    plane_stride: int          # 2 for hex20/tet10 half-step planes, 1 for hex8

    @property
    def n_elem(self) -> int:
        return self.elem_nodes.shape[0]

    @property
    def n_nodes(self) -> int:
        return self.coords.shape[0]

    @property
    def n_dofs(self) -> int:
        return self.n_nodes * self.dofs_per_node

    def dof_map(self) -> np.ndarray:
        """(E, nodes * dpn) global DoF ids, DoF = dpn * node + component."""
        d = self.dofs_per_node
        if d == 1:
            return self.elem_nodes.astype(np.uint64)
        base = self.elem_nodes.astype(np.uint64)[:, :, None] * np.uint64(d)
        return (base + np.arange(d, dtype=np.uint64)[None, None, :]).reshape(self.n_elem, -1)


def hex8(nx: int, ny: int, nz: int, L=(1.0, 1.0, 1.0)) -> Mesh:
    L = tuple(float(v) for v in (L if hasattr(L, "__len__") else (L, L, L)))
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (i + j * sx + k * sy).reshape(-1).astype(np.int64)
    corners = [0, 1, sx + 1, sx, sy, sy + 1, sy + sx + 1, sy + sx]
    elem = np.stack([n0 + c for c in corners], axis=1).astype(np.uint32)
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([ii.reshape(-1) * (L[0] / nx), jj.reshape(-1) * (L[1] / ny), kk.reshape(-1) * (L[2] / nz)],
                      axis=1)
    return Mesh("hex8", (nx, ny, nz), L, elem, coords, 1, kk.reshape(-1).astype(np.int64), 1)


# half-step offsets of the hex20 nodes (examples/include/shapeFunc.hpp:200-269):
# corners, then mid-edges 8..11 bottom, 12..15 top, 16..19 vertical
_HEX20_OFFS = [(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 2), (2, 0, 2), (2, 2, 2), (0, 2, 2),
               (1, 0, 0), (2, 1, 0), (1, 2, 0), (0, 1, 0), (1, 0, 2), (2, 1, 2), (1, 2, 2), (0, 1, 2),
               (0, 0, 1), (2, 0, 1), (2, 2, 1), (0, 2, 1)]


def hex20(nx: int, ny: int, nz: int, L=(1.0, 1.0, 1.0), dofs_per_node: int = 3) -> Mesh:
    L = tuple(float(v) for v in (L if hasattr(L, "__len__") else (L, L, L)))
    fx, fy, fz = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    K, J, I = np.meshgrid(np.arange(fz), np.arange(fy), np.arange(fx), indexing="ij")
    odd = (I % 2) + (J % 2) + (K % 2)
    keep = (odd <= 1).reshape(-1)             # no face or body centres
    ids = np.full(fx * fy * fz, -1, dtype=np.int64)
    ids[keep] = np.arange(int(keep.sum()))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f0 = (2 * i + 2 * j * fx + 2 * k * fx * fy).reshape(-1)
    elem = np.stack([ids[f0 + a + b * fx + c * fx * fy] for (a, b, c) in _HEX20_OFFS], axis=1)
    assert (elem >= 0).all()
    h = (L[0] / (2 * nx), L[1] / (2 * ny), L[2] / (2 * nz))
    coords = np.stack([I.reshape(-1)[keep] * h[0], J.reshape(-1)[keep] * h[1], K.reshape(-1)[keep] * h[2]], axis=1)
    return Mesh("hex20", (nx, ny, nz), L, elem.astype(np.uint32), coords, dofs_per_node,
                K.reshape(-1)[keep].astype(np.int64), 2)


# the six Kuhn tetrahedra of the unit cube: a path 000 -> 111 adding unit
# vectors in the order of a permutation; odd permutations swap vertices 1, 2
# so that every Jacobian is positive
def _kuhn_tets():
    tets = []
    for perm in itertools.permutations(range(3)):
        v = [np.zeros(3, dtype=int)]
        for ax in perm:
            w = v[-1].copy()
            w[ax] = 1
            v.append(w)
        vol = np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]]))
        if vol < 0:
            v[1], v[2] = v[2], v[1]
        tets.append([tuple(x) for x in v])
    return tets


_KUHN = _kuhn_tets()
# tet10 mid-edge nodes 4..9 (examples/include/shapeFunc.hpp:178-199)
_TET_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]


def tet10(n: int, seed: int = 0, jitter: float = 0.2, L: float = 1.0) -> Mesh:
    f = 2 * n + 1
    h = L / n
    K, J, I = np.meshgrid(np.arange(f), np.arange(f), np.arange(f), indexing="ij")
    I, J, K = I.reshape(-1), J.reshape(-1), K.reshape(-1)
    fid = lambda a, b, c: a + f * b + f * f * c  # noqa: E731
    coords = np.stack([I, J, K], axis=1).astype(np.float64) * (h / 2)
    # jitter interior vertices (even half-step indices strictly inside)
    vert = (I % 2 == 0) & (J % 2 == 0) & (K % 2 == 0)
    inner = vert & (I > 0) & (I < 2 * n) & (J > 0) & (J < 2 * n) & (K > 0) & (K < 2 * n)
    vid = np.nonzero(inner)[0]
    if jitter > 0 and len(vid):
        d = uniform(seed, 7, 3 * len(vid), -jitter * h, jitter * h).reshape(-1, 3)
        coords[vid] += d
    # mid-edge points: average of the two vertices of their (Kuhn) edge
    o = np.stack([I % 2, J % 2, K % 2], axis=1)
    mid = np.nonzero(o.sum(axis=1) > 0)[0]
    lo_ = fid(I[mid] - o[mid, 0], J[mid] - o[mid, 1], K[mid] - o[mid, 2])
    hi_ = fid(I[mid] + o[mid, 0], J[mid] + o[mid, 1], K[mid] + o[mid, 2])
    coords[mid] = 0.5 * (coords[lo_] + coords[hi_])
    # elements: hex-major, 6 tets per hex
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = np.stack([2 * i.reshape(-1), 2 * j.reshape(-1), 2 * k.reshape(-1)], axis=1)
    elems = []
    for tv in _KUHN:
        v = [base + 2 * np.array(t) for t in tv]
        nodes = [fid(p[:, 0], p[:, 1], p[:, 2]) for p in v]
        for (a, b) in _TET_EDGES:
            m = (v[a] + v[b]) // 2
            nodes.append(fid(m[:, 0], m[:, 1], m[:, 2]))
        elems.append(np.stack(nodes, axis=1))
    elem = np.stack(elems, axis=1).reshape(-1, 10).astype(np.uint32)
    return Mesh("tet10", (n, n, n), (L, L, L), elem, coords, 1, K.astype(np.int64), 2)


def boundary_nodes(mesh: Mesh, faces=("x0", "x1", "y0", "y1", "z0", "z1"), tol=1e-12) -> np.ndarray:
    """Nodes on the requested faces of the box (coordinate predicates of
    examples/src/example03a.cpp:351-370)."""
    c = mesh.coords
    m = np.zeros(mesh.n_nodes, dtype=bool)
    for fc in faces:
        ax = "xyz".index(fc[0])
        val = 0.0 if fc[1] == "0" else mesh.L[ax]
        m |= np.abs(c[:, ax] - val) < tol * max(1.0, mesh.L[ax])
    return np.nonzero(m)[0]


@dataclasses.dataclass
class RankMaps:
    """Inputs of Maps::set_map for one rank, built the way the examples do."""
    rank: int
    size: int
    elements: np.ndarray           # global element ids of this rank (ascending)
    element_to_rank_map: np.ndarray  # (E_r, n) caller-local DoF ids
    rank_to_global_map: np.ndarray   # local DoF -> global DoF
    owned_begin: int
    owned_end: int                 # inclusive
    n_global: int
    rank_dof_counts: np.ndarray
    n_pre: int
    n_owned: int
    n_post: int

    @property
    def n_all_dofs(self) -> int:
        return self.n_pre + self.n_owned + self.n_post


def zslab_partition(mesh: Mesh, size: int) -> list[RankMaps]:
    """Split the element layers over `size` ranks (example03a.cpp:190-211) and
    number each rank's DoFs [pre ghosts | owned | post ghosts]."""
    nz = mesh.dims[2]
    if size > nz:
        raise ValueError("more ranks than element layers")
    nzmin, rem = divmod(nz, size)
    st = mesh.plane_stride
    # element layer of every element: structured meshes are layer-major
    per_layer = mesh.n_elem // nz
    layer = np.arange(mesh.n_elem) // per_layer
    slabs = []
    for r in range(size):
        nel = nzmin + (1 if r < rem else 0)
        emin = r * nzmin + min(r, rem)
        slabs.append((emin, emin + nel - 1))
    # node owner by plane: rank 0 owns [0, st*(emax+1)], rank r > 0 owns
    # (st*emin, st*(emax+1)]
    owner = np.zeros(mesh.n_nodes, dtype=np.int64)
    for r, (emin, emax) in enumerate(slabs):
        lo = 0 if r == 0 else st * emin + 1
        owner[(mesh.node_plane >= lo) & (mesh.node_plane <= st * (emax + 1))] = r
    d = mesh.dofs_per_node
    dof_owner = np.repeat(owner, d)
    counts = np.bincount(dof_owner, minlength=size).astype(np.uint32)
    # owned ids contiguous per rank because numbering is plane-major
    scan = np.concatenate([[0], np.cumsum(counts.astype(np.int64))[:-1]])
    gmap_all = mesh.dof_map()
    out = []
    for r, (emin, emax) in enumerate(slabs):
        els = np.nonzero((layer >= emin) & (layer <= emax))[0]
        g = gmap_all[els].astype(np.int64)
        b, e = int(scan[r]), int(scan[r] + counts[r])
        pre = np.unique(g[g < b])
        post = np.unique(g[g >= e])
        local = np.empty_like(g)
        own = (g >= b) & (g < e)
        local[own] = g[own] - b + len(pre)
        local[g < b] = np.searchsorted(pre, g[g < b])
        local[g >= e] = len(pre) + (e - b) + np.searchsorted(post, g[g >= e])
        r2g = np.concatenate([pre, np.arange(b, e), post]).astype(np.uint64)
        out.append(RankMaps(r, size, els.astype(np.uint32), local.astype(np.uint32), r2g, b, e - 1,
                            int(counts.sum()), counts, len(pre), e - b, len(post)))
    return out


def constraints_for(mesh: Mesh, elements: np.ndarray, nodes: np.ndarray, components=None, value: float = 0.0):
    """Constrained global DoFs that appear in `elements` (the per-rank lists of
    examples/src/example03a.cpp:372-398: sorted, unique)."""
    d = mesh.dofs_per_node
    comps = range(d) if components is None else components
    on = np.zeros(mesh.n_nodes, dtype=bool)
    on[nodes] = True
    used = np.unique(mesh.elem_nodes[elements].reshape(-1))
    cn = used[on[used]].astype(np.uint64)
    dofs = np.sort(np.concatenate([cn * np.uint64(d) + np.uint64(c) for c in comps])) if len(cn) else np.zeros(0, np.uint64)
    return dofs.astype(np.uint64), np.full(len(dofs), value)
