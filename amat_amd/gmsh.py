"""Gmsh meshes: the GmshMesh helper of paralab/aMat's examples
(examples/include/GmshMesh.hpp, examples/src/GmshMesh.cpp) over the C ABI
(amx_gmsh_*), and a writer of ASCII Gmsh 2.2 files for the structured meshes
of amat_amd.mesh.

    mesh = GmshMesh("cube_4p.msh", rank, size, ndof_per_node=1)
    maps = Maps(rank, size)
    maps.set_map(*mesh.set_map_args())
    maps.set_bdr_map(mesh.get_constrained_dofs(), mesh.get_prescribed_values(fn))

is the set-up of examples/src/example05a.cpp:135-163.  Every rank reads the
whole file; the partition of each element comes from its Gmsh tags.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _capi
from ._capi import call
from .mesh import Mesh

# Gmsh element types (GmshMesh.hpp:28-38)
GMSH_TYPE = {"hex8": 5, "tet10": 11, "hex20": 17, "tet4": 4, "hex27": 12}
NODES_OF_TYPE = {15: 1, 1: 2, 8: 3, 2: 3, 3: 4, 9: 6, 10: 9, 16: 8, 4: 4, 5: 8, 11: 10, 12: 27, 17: 20}

# aMat node a of an element is Gmsh node PERM[kind][a] (tet10: Gmsh's mid-edge
# nodes 8, 9 lie on edges (2,3), (1,3); hex20: Gmsh's mid-edge nodes 8..19 on
# (0,1)(0,3)(0,4)(1,2)(1,5)(2,3)(2,6)(3,7)(4,5)(4,7)(5,6)(6,7))
PERM = {
    "hex8": np.arange(8),
    "tet10": np.array([0, 1, 2, 3, 4, 5, 6, 7, 9, 8]),
    "hex20": np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 13, 9, 16, 18, 19, 17, 10, 12, 14, 15]),
}

# faces of the aMat elements as (corner nodes, nodes in Gmsh face order, Gmsh face type)
_TET10_EDGE = {(0, 1): 4, (1, 2): 5, (0, 2): 6, (0, 3): 7, (1, 3): 8, (2, 3): 9}
_HEX20_EDGE = {(0, 1): 8, (1, 2): 9, (2, 3): 10, (0, 3): 11, (4, 5): 12, (5, 6): 13, (6, 7): 14, (4, 7): 15,
               (0, 4): 16, (1, 5): 17, (2, 6): 18, (3, 7): 19}
_HEX_FACES = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
_TET_FACES = [(0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)]


def _face_nodes(kind: str):
    """[(corners, face node list in Gmsh order, Gmsh type)] of one element."""
    out = []
    if kind == "tet10":
        for f in _TET_FACES:
            mids = [_TET10_EDGE[tuple(sorted((f[i], f[(i + 1) % 3])))] for i in range(3)]
            out.append((f, list(f) + mids, 9))
    elif kind == "hex20":
        for f in _HEX_FACES:
            mids = [_HEX20_EDGE[tuple(sorted((f[i], f[(i + 1) % 4])))] for i in range(4)]
            out.append((f, list(f) + mids, 16))
    elif kind == "hex8":
        out = [(f, list(f), 3) for f in _HEX_FACES]
    else:
        raise ValueError(f"no face table for {kind}")
    return out


def boundary_faces(mesh: Mesh):
    """Faces that belong to one element only: (element, Gmsh type, (F, k) node ids)."""
    faces = _face_nodes(mesh.kind)
    E = mesh.n_elem
    keys = np.concatenate([np.sort(mesh.elem_nodes[:, list(c)], axis=1) for c, _, _ in faces])
    elem = np.tile(np.arange(E), len(faces))
    which = np.repeat(np.arange(len(faces)), E)
    _, inv, cnt = np.unique(keys, axis=0, return_inverse=True, return_counts=True)
    once = np.nonzero(cnt[inv.reshape(-1)] == 1)[0]
    order = np.lexsort((which[once], elem[once]))   # element order, then face order
    once = once[order]
    nodes = np.stack([mesh.elem_nodes[elem[i], faces[which[i]][1]] for i in once]) if len(once) else None
    return elem[once], faces[0][2], nodes


def write_msh(path, mesh: Mesh, parts=None, boundary: bool = True, extras: bool = True, labels=None,
              node_order=None) -> None:
    """Writes `mesh` as an ASCII Gmsh 2.2 file.

    parts: 0-based partition of every element (None: no partition tags).
    boundary: write the boundary faces as surface elements (the constrained
    nodes GmshMesh reads).  extras: lead with a point and a line element, as
    Gmsh files do.  labels: node labels (default 1..N); node_order: the order
    in which $Nodes lists the nodes.  Elements follow Gmsh's order: points,
    lines, surfaces, volumes."""
    n = mesh.n_nodes
    lab = np.arange(1, n + 1, dtype=np.int64) if labels is None else np.asarray(labels, dtype=np.int64)
    vtype = GMSH_TYPE[mesh.kind]
    gm = np.empty_like(mesh.elem_nodes)
    gm[:, PERM[mesh.kind]] = mesh.elem_nodes

    def tags(p):
        return "2 1 1" if parts is None else f"4 1 1 1 {int(p) + 1}"

    lines = []
    if extras:
        p0 = 0 if parts is None else parts[0]
        lines.append((15, tags(p0), [gm[0, 0]]))
        lines.append((1, tags(p0), [gm[0, 0], gm[0, 1]]))
    if boundary:
        fe, ftype, fnodes = boundary_faces(mesh)
        for e, nd in zip(fe, fnodes if fnodes is not None else []):
            lines.append((ftype, tags(0 if parts is None else parts[e]), nd))
    for e in range(mesh.n_elem):
        lines.append((vtype, tags(0 if parts is None else parts[e]), gm[e]))
    order = np.arange(n) if node_order is None else np.asarray(node_order)
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % n)
        c = mesh.coords
        f.write("".join("%d %.17g %.17g %.17g\n" % (lab[i], c[i, 0], c[i, 1], c[i, 2]) for i in order))
        f.write("$EndNodes\n$Elements\n%d\n" % len(lines))
        f.write("".join("%d %d %s %s\n" % (k + 1, t, tg, " ".join(str(int(lab[v])) for v in nd))
                        for k, (t, tg, nd) in enumerate(lines)))
        f.write("$EndElements\n")


class GmshMesh:
    """GmshMesh (examples/include/GmshMesh.hpp:40-226) for one rank.

    Local node ids are [pre ghosts | owned | post ghosts]; the element maps are
    2-D arrays when every element has the same node count, else lists of rows
    (the forms Maps.set_map takes)."""

    def __init__(self, path, rank: int = 0, size: int = 1, ndof_per_node: int = 1):
        h = C.c_void_p()
        call("amx_gmsh_read", os.fsencode(os.fspath(path)), rank, size, ndof_per_node, C.byref(h))
        self._h = h
        inf = _capi.GmshInfo()
        call("amx_gmsh_get_info", h, C.byref(inf))
        self.info = inf
        self.rank, self.size, self.ndof = rank, size, ndof_per_node
        ne = inf.n_owned_elements
        self._npe = np.zeros(max(ne, 1), np.uint32)
        self._types = np.zeros(max(ne, 1), np.uint32)
        call("amx_gmsh_get_element_sizes", h, self._ptr(self._npe), None, self._ptr(self._types))
        self._npe, self._types = self._npe[:ne], self._types[:ne]
        self._eoff = np.concatenate([[0], np.cumsum(self._npe, dtype=np.int64)])
        nent = inf.n_entries
        self._lmap = np.zeros(max(nent, 1), np.uint32)
        self._gmap = np.zeros(max(nent, 1), np.uint64)
        call("amx_gmsh_get_maps", h, self._ptr(self._lmap), self._ptr(self._gmap), None, None)
        self._lmap, self._gmap = self._lmap[:nent], self._gmap[:nent]
        nl = inf.n_local_nodes
        self._l2g = np.zeros(max(nl, 1), np.uint64)
        call("amx_gmsh_get_local2global", h, self._ptr(self._l2g), None)
        self._l2g = self._l2g[:nl]
        self._xyz = np.zeros((max(nl, 1), 3))
        call("amx_gmsh_get_coordinates", h, self._ptr(self._xyz))
        self._xyz = self._xyz[:nl]
        self._counts = np.zeros(size, np.uint32)
        call("amx_gmsh_get_rank_node_counts", h, self._ptr(self._counts))
        nb = inf.n_boundary_nodes
        self._bdr = np.zeros(max(nb, 1), np.uint64)
        call("amx_gmsh_get_constraints", h, self._ptr(self._bdr), None)
        self._bdr = self._bdr[:nb]

    @staticmethod
    def _ptr(a):
        return a.ctypes.data_as(C.c_void_p)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().amx_gmsh_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def _rows(self, flat):
        if len(self._npe) and (self._npe == self._npe[0]).all():
            return flat.reshape(len(self._npe), -1)
        k = self._eoff * (len(flat) // max(int(self._eoff[-1]), 1))
        return [flat[k[e]:k[e + 1]] for e in range(len(self._npe))]

    def _dof(self, node_ids):
        d = np.uint64(self.ndof)
        return (node_ids.astype(np.uint64)[:, None] * d + np.arange(self.ndof, dtype=np.uint64)).reshape(-1)

    # -- getters of GmshMesh.hpp:150-220 --------------------------------
    def get_nOwnedElems(self) -> int:
        return int(self.info.n_owned_elements)

    def get_nOwnedDofs(self) -> int:
        return int(self.info.n_owned_nodes) * self.ndof

    def get_nNodesPerElem(self) -> np.ndarray:
        return self._npe.copy()

    def get_nDofsPerElem(self) -> np.ndarray:
        return self._npe * np.uint32(self.ndof)

    def get_element_types(self) -> np.ndarray:
        """Gmsh type of every owned element."""
        return self._types.copy()

    def get_localMap(self):
        return self._rows(self._lmap.copy())

    def get_globalMap(self):
        return self._rows(self._gmap.copy())

    def get_localDofMap(self):
        return self._rows(self._dof(self._lmap).astype(np.uint32))

    def get_globalDofMap(self):
        return self._rows(self._dof(self._gmap))

    def get_nLocalNodes(self) -> int:
        return int(self.info.n_local_nodes)

    def get_nLocalDofs(self) -> int:
        return int(self.info.n_local_nodes) * self.ndof

    def get_local2GlobalMap(self) -> np.ndarray:
        return self._l2g.copy()

    def get_local2GlobalDofMap(self) -> np.ndarray:
        return self._dof(self._l2g)

    def get_startGlobalDof(self) -> int:
        return int(self.info.owned_dof_begin)

    def get_endGlobalDof(self) -> int:
        return int(self.info.owned_dof_end)

    def get_nDofsTotal(self) -> int:
        return int(self.info.n_dofs_total)

    def get_n_constraints(self) -> int:
        return int(self.info.n_boundary_nodes) * self.ndof

    def get_constrained_dofs(self) -> np.ndarray:
        return self._dof(self._bdr)

    def get_boundary_nodes(self) -> np.ndarray:
        return self._bdr.copy()

    def get_prescribed_values(self, coord2displ) -> np.ndarray:
        """coord2displ(xyz) -> ndof_per_node values at a boundary node
        (GmshMesh.cpp:752-783)."""
        d = self.ndof
        err = []

        def cb(xyz, out):
            try:
                v = np.asarray(coord2displ(np.array([xyz[0], xyz[1], xyz[2]])), dtype=np.float64).reshape(-1)
                for i in range(d):
                    out[i] = float(v[i])
            except Exception as e:  # an exception cannot cross the C frame
                err.append(e)

        fn = _capi.DISPLACEMENT_FN(cb)
        out = np.zeros(max(len(self._bdr) * d, 1))
        call("amx_gmsh_get_prescribed_values", self._h, fn, self._ptr(out))
        if err:
            raise err[0]
        return out[:len(self._bdr) * d]

    def get_x(self, localNid: int) -> float:
        return float(self._xyz[localNid, 0])

    def get_y(self, localNid: int) -> float:
        return float(self._xyz[localNid, 1])

    def get_z(self, localNid: int) -> float:
        return float(self._xyz[localNid, 2])

    # -- conveniences -------------------------------------------------------
    def coordinates(self) -> np.ndarray:
        """(n_local_nodes, 3) coordinates of the local nodes."""
        return self._xyz.copy()

    def element_coordinates(self) -> np.ndarray:
        """(E, nodes, 3) node coordinates of every owned element (one element size)."""
        return self._xyz[np.asarray(self.get_localMap(), dtype=np.int64)]

    def rank_dof_counts(self) -> np.ndarray:
        """Owned DoF count of every rank (the Allgather'ed counts of Maps::set_map)."""
        return self._counts * np.uint32(self.ndof)

    def set_map_args(self) -> tuple:
        """The arguments of Maps.set_map for this rank."""
        return (self.get_nOwnedElems(), self.get_localDofMap(), self.get_nDofsPerElem(), self.get_nLocalDofs(),
                self.get_local2GlobalDofMap(), self.get_startGlobalDof(), self.get_endGlobalDof(),
                self.get_nDofsTotal(), self.rank_dof_counts() if self.size > 1 else None)
