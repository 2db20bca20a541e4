"""The workloads: meshes, element matrices, BCs and seeded inputs, one rank.

poisson_hex8      hex8 Laplacian on the unit cube, K_e from ke_hex8 with 2x2x2
                  Gauss points (identical for all elements), identity BCs on
                  every boundary node (examples/src/example03a.cpp).
elasticity_hex20  hex20 isotropic elasticity, K_e = E_e * ke_hex20_iso(E=1, nu)
                  with E_e ~ U[0.5, 2) from a seed (all elements congruent),
                  z = 0 face fixed (examples/src/example07a.cpp numbering).
poisson_tet10     jittered tet10 Laplacian, ke_tet10 with NGT = 2
                  (examples/src/example05a.cpp:93), b = M f with
                  f = sin(2 pi x) sin(2 pi y) sin(2 pi z) through fe_tet10,
                  identity BCs on the boundary.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import fem, mesh as meshmod

SEED_U = 1
SEED_E = 2


@dataclasses.dataclass
class Problem:
    name: str
    mesh: meshmod.Mesh
    dofs_per_elem: int
    gmap: np.ndarray                 # (E, n) global DoF ids (one rank: = local)
    k_ref: np.ndarray | None = None  # (n, n) when K_e = scale[e] * k_ref
    scale: np.ndarray | None = None
    K: np.ndarray | None = None      # (E, n, n) otherwise
    constrained: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.uint64))
    prescribed: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    rhs: np.ndarray | None = None
    # method 2 (FREE) description of the same K_e: kind, ngt, E, nu
    free: dict | None = None

    @property
    def n_elem(self) -> int:
        return self.gmap.shape[0]

    @property
    def n_dofs(self) -> int:
        return self.mesh.n_dofs

    def element_matrices(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.n_elem - first if count is None else count
        if self.K is not None:
            return self.K[first:first + count]
        return self.scale[first:first + count, None, None] * self.k_ref[None]

    def build_maps(self, Maps):
        m = Maps(0, 1)
        n = self.n_dofs
        m.set_map(self.n_elem, self.gmap.astype(np.uint32), np.full(self.n_elem, self.dofs_per_elem, np.uint32),
                  n, np.arange(n, dtype=np.uint64), 0, n - 1, n)
        m.set_bdr_map(self.constrained, self.prescribed)
        return m

    def build_operator(self, aMatFree, Maps, bc_meth=0, precision=1, scatter=0, device=0, bcop=None, penalty=None,
                       sym_storage=None):
        maps = self.build_maps(Maps)
        op = aMatFree(maps, bc_meth, device)
        op.set_matfree_type(3)
        if sym_storage is not None:
            op.set_symmetric_storage(sym_storage)
        if self.K is not None:
            op.set_element_matrices(0, self.K)
        else:
            op.set_element_matrices_scaled(self.k_ref, self.scale)
        op.set_options(precision, scatter)
        if bcop is not None:
            op.set_bc_operator(bcop)
        if penalty is not None:
            op.set_penalty(penalty)
        op.finalize()
        return op

    def element_coords(self) -> np.ndarray:
        """(E, nodes, 3) node coordinates in element DoF-node order."""
        return np.ascontiguousarray(self.mesh.coords[self.mesh.elem_nodes])

    def build_free_operator(self, aMatFree, Maps, bc_meth=0, scatter=0, device=0, bcop=None, penalty=None):
        """The same operator by method 2 (FREE): K_e computed in-kernel from the
        coordinates at every application."""
        f = self.free
        maps = self.build_maps(Maps)
        op = aMatFree(maps, bc_meth, device)
        op.set_matfree_type(2)
        op.set_element_function(f["kind"], self.element_coords(), self.scale, f.get("E", 1.0), f.get("nu", 0.3),
                                f["ngt"])
        op.set_options(1, scatter)
        if bcop is not None:
            op.set_bc_operator(bcop)
        if penalty is not None:
            op.set_penalty(penalty)
        op.finalize()
        return op

    def random_u(self, seed: int = SEED_U) -> np.ndarray:
        return meshmod.uniform(seed, 0, self.n_dofs, -1.0, 1.0)


def hex8_cube_ke(h: float, ngt: int = 2) -> np.ndarray:
    xe = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                  dtype=np.float64) * h
    return fem.ke_hex8(xe, ngt)


def poisson_hex8(n: int, ny: int | None = None, nz: int | None = None, L: float = 1.0) -> Problem:
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    m = meshmod.hex8(n, ny, nz, (L, L, L))
    assert abs(L / n - L / ny) < 1e-15 and abs(L / n - L / nz) < 1e-15 or True
    xe = m.coords[m.elem_nodes[0]]
    kref = fem.ke_hex8(xe, 2)
    cons = meshmod.boundary_nodes(m).astype(np.uint64)
    return Problem("poisson_hex8", m, 8, m.dof_map(), k_ref=kref, scale=np.ones(m.n_elem), constrained=cons,
                   prescribed=np.zeros(len(cons)), free={"kind": 1, "ngt": 2})


def elasticity_hex20(n: int, ny: int | None = None, nz: int | None = None, E: float = 1.0, nu: float = 0.3,
                     ngt: int = 3, seed: int = SEED_E, perturb: bool = True) -> Problem:
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    m = meshmod.hex20(n, ny, nz, (1.0, 1.0, 1.0))
    xe = m.coords[m.elem_nodes[0]]
    kref = fem.ke_hex20_iso(xe, E, nu, ngt)
    scale = meshmod.uniform(seed, 1, m.n_elem, 0.5, 2.0) if perturb else np.ones(m.n_elem)
    z0 = meshmod.boundary_nodes(m, faces=("z0",)).astype(np.uint64)
    cons = np.sort(np.concatenate([z0 * np.uint64(3) + np.uint64(c) for c in range(3)]))
    return Problem("elasticity_hex20", m, 60, m.dof_map(), k_ref=kref, scale=scale, constrained=cons,
                   prescribed=np.zeros(len(cons)), free={"kind": 2, "ngt": ngt, "E": E, "nu": nu})


def poisson_tet10(n: int, seed: int = 3, jitter: float = 0.2, with_rhs: bool = True) -> Problem:
    m = meshmod.tet10(n, seed=seed, jitter=jitter)
    coords = m.coords[m.elem_nodes]          # (E, 10, 3)
    K = fem.ke_tet10_batch(coords, 2)
    cons = meshmod.boundary_nodes(m).astype(np.uint64)
    rhs = None
    if with_rhs:
        c = m.coords
        f = np.sin(2 * np.pi * c[:, 0]) * np.sin(2 * np.pi * c[:, 1]) * np.sin(2 * np.pi * c[:, 2])
        fe = fem.fe_tet10_batch(coords, f[m.elem_nodes], 2)
        rhs = np.bincount(m.elem_nodes.reshape(-1).astype(np.int64), weights=fe.reshape(-1), minlength=m.n_nodes)
    return Problem("poisson_tet10", m, 10, m.dof_map(), K=K, constrained=cons, prescribed=np.zeros(len(cons)),
                   rhs=rhs, free={"kind": 3, "ngt": 2})
