"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run in the build container only (it needs the reference tree and MPICH):

    python tests/golden/make_golden.py

* fem_ref.npz  -- element matrices/vectors from the reference's
  examples/src/ke_matrix.cpp and fe_vector.cpp (compiled by `make -C oracle
  ref` into oracle/_ref/libamat_ref_fem.so) on fixed element geometries, and
  Gauss-Legendre points of integration<double>.
* maps_<case>_p<P>.npz -- inputs of par::Maps::set_map / set_bdr_map for every
  rank of a z-slab partition (built by amat_amd.mesh) and the outputs of the
  reference include/maps.hpp run under `mpiexec -n P` (oracle/_ref/ref_maps_driver).
* gmsh_<case>.npz -- an ASCII Gmsh 2.2 file of a structured mesh with partition
  tags (written by amat_amd.gmsh.write_msh) and, for every rank, the getters of
  the reference examples/src/GmshMesh.cpp run on it under `mpiexec -n P`
  (oracle/_ref/ref_gmsh_driver).

The fixtures are data only: inputs and the reference's outputs.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from amat_amd import mesh as meshmod  # noqa: E402

OUT = ROOT / "tests" / "golden"
REF = ROOT / "oracle" / "_ref"
MPIEXEC = "/opt/conda/bin/mpiexec"


def fem_fixtures() -> None:
    L = C.CDLL(str(REF / "libamat_ref_fem.so"))
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    out = {}
    for n in range(1, 6):
        g = np.zeros(2 * n)
        L.ref_gauss(C.c_uint(n), p(g))
        out[f"gauss_{n}"] = g
    cube = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    rng = np.random.default_rng(20240611)
    hexes = {"cube32": cube / 32.0, "cube1": cube, "distorted": cube + rng.uniform(-0.15, 0.15, cube.shape)}
    for name, xe in hexes.items():
        x = np.ascontiguousarray(xe)
        for ngt in (2, 3):
            ke = np.zeros(64)
            L.ref_ke_hex8(p(ke), p(x), C.c_uint(ngt))
            out[f"hex8_{name}_ngt{ngt}_xe"] = x
            out[f"hex8_{name}_ngt{ngt}_ke"] = ke.reshape(8, 8)
        ke = np.zeros(64)
        L.ref_ke_hex8_eig(p(ke), p(x), C.c_uint(2))
        out[f"hex8eig_{name}_ke"] = ke.reshape(8, 8)
    # hex20 geometries: the unit cube of the 64^3 elasticity workload (h = 1/64) and a distorted one
    m = meshmod.hex20(1, 1, 1)
    base = m.coords[m.elem_nodes[0]]
    h20 = {"cube64": base / 64.0, "cube1": base,
           "distorted": base + rng.uniform(-0.08, 0.08, base.shape)}
    for name, xe in h20.items():
        x = np.ascontiguousarray(xe)
        for (E, nu, ngt) in ((1.0, 0.3, 3), (1.0e6, 0.3, 2), (2.5, 0.25, 3)):
            ke = np.zeros(3600)
            L.ref_ke_hex20_iso(p(ke), p(x), C.c_double(E), C.c_double(nu), C.c_uint(ngt))
            key = f"hex20_{name}_E{E:g}_nu{nu:g}_ngt{ngt}"
            out[key + "_xe"] = x
            out[key + "_ke"] = ke.reshape(60, 60)
    # tet10: elements of a jittered mesh
    t = meshmod.tet10(2, seed=5, jitter=0.2)
    for idx in (0, 7, 23, 40):
        x = np.ascontiguousarray(t.coords[t.elem_nodes[idx]])
        be = np.ascontiguousarray(rng.uniform(-1, 1, 10))
        for ngt in (2, 3):
            ke = np.zeros(100)
            L.ref_ke_tet10(p(ke), p(x), C.c_uint(ngt))
            fe = np.zeros(10)
            L.ref_fe_tet10(p(fe), p(x), p(be), C.c_uint(ngt))
            out[f"tet10_{idx}_ngt{ngt}_xe"] = x
            out[f"tet10_{idx}_ngt{ngt}_be"] = be
            out[f"tet10_{idx}_ngt{ngt}_ke"] = ke.reshape(10, 10)
            out[f"tet10_{idx}_ngt{ngt}_fe"] = fe
    np.savez_compressed(OUT / "fem_ref.npz", **out)
    print("wrote fem_ref.npz", len(out), "arrays")


def _read_vec(f, dtype):
    n = int(np.frombuffer(f.read(8), dtype=np.uint64)[0])
    return np.frombuffer(f.read(n * np.dtype(dtype).itemsize), dtype=dtype).copy()


def maps_case(name: str, m: meshmod.Mesh, P: int, cons_nodes, comps=None, extra=None) -> None:
    ranks = meshmod.zslab_partition(m, P)
    with tempfile.TemporaryDirectory() as td:
        tdp = pathlib.Path(td)
        rec = {}
        for r in ranks:
            cd, cv = meshmod.constraints_for(m, r.elements, cons_nodes, comps)
            cv = cv + np.arange(len(cv)) * 0.25  # distinct prescribed values
            if extra is not None:  # duplicates (first wins), foreign ids
                xd, xv = extra(r, cd)
                cd, cv = np.concatenate([xd, cd]).astype(np.uint64), np.concatenate([xv, cv])
            e2r = r.element_to_rank_map.astype(np.uint32)
            with open(tdp / f"in_r{r.rank}.bin", "wb") as f:
                f.write(np.array([e2r.shape[0], r.n_all_dofs, len(cd)], np.uint32).tobytes())
                f.write(np.array([r.owned_begin, r.owned_end, r.n_global], np.uint64).tobytes())
                f.write(np.full(e2r.shape[0], e2r.shape[1], np.uint32).tobytes())
                f.write(e2r.tobytes())
                f.write(r.rank_to_global_map.astype(np.uint64).tobytes())
                f.write(cd.astype(np.uint64).tobytes())
                f.write(cv.astype(np.float64).tobytes())
            pre = f"r{r.rank}_"
            rec[pre + "e2r"] = e2r
            rec[pre + "r2g"] = r.rank_to_global_map
            rec[pre + "range"] = np.array([r.owned_begin, r.owned_end, r.n_global], np.uint64)
            rec[pre + "cd"] = cd
            rec[pre + "cv"] = cv
            rec[pre + "elements"] = r.elements
        rec["rank_dof_counts"] = ranks[0].rank_dof_counts
        cmd = [MPIEXEC, "-launcher", "fork", "-n", str(P), str(REF / "ref_maps_driver"), str(tdp)]
        subprocess.run(cmd, check=True, capture_output=True, timeout=300)
        names = ["local_map", "global_map", "local2global", "localid2rank", "send_counts", "send_offsets",
                 "send_ids", "send_rank_ids", "recv_counts", "recv_offsets", "recv_rank_ids", "independent",
                 "dependent", "bdr_map", "presval", "owned_cdofs", "owned_cvals", "owned_free", "dof_scan"]
        types = [np.uint32, np.uint64, np.uint64, np.uint32, np.uint32, np.uint32, np.uint32, np.uint32, np.uint32,
                 np.uint32, np.uint32, np.uint32, np.uint32, np.uint32, np.float64, np.uint64, np.float64, np.uint64,
                 np.uint64]
        for r in ranks:
            with open(tdp / f"out_r{r.rank}.bin", "rb") as f:
                counts = np.frombuffer(f.read(16), dtype=np.uint32).copy()
                rec[f"r{r.rank}_counts"] = counts
                for nm, ty in zip(names, types):
                    rec[f"r{r.rank}_ref_{nm}"] = _read_vec(f, ty)
    rec["size"] = np.array([P])
    np.savez_compressed(OUT / f"maps_{name}_p{P}.npz", **rec)
    print(f"wrote maps_{name}_p{P}.npz")


def maps_fixtures() -> None:
    h8 = meshmod.hex8(3, 3, 4)
    bnd8 = meshmod.boundary_nodes(h8)

    def extra8(r, cd):
        # a duplicate of an existing constraint with another value (the
        # reference keeps the first occurrence) and an id no element touches
        if len(cd) == 0:
            return np.zeros(0, np.uint64), np.zeros(0)
        return np.array([cd[len(cd) // 2], 10 ** 6], np.uint64), np.array([-7.5, 3.0])

    for P in (1, 2, 3):
        maps_case("hex8", h8, P, bnd8, extra=extra8)
    h20 = meshmod.hex20(2, 2, 3)
    z0 = meshmod.boundary_nodes(h20, faces=("z0",))
    x1 = meshmod.boundary_nodes(h20, faces=("x1",))
    for P in (1, 2, 3):
        maps_case("hex20", h20, P, np.union1d(z0, x1[:3]))
    t10 = meshmod.tet10(2, seed=5)
    for P in (1, 2):
        maps_case("tet10", t10, P, meshmod.boundary_nodes(t10))


def gmsh_case(name: str, m: meshmod.Mesh, P: int, dpn: int, parts) -> None:
    from amat_amd import gmsh as gmshmod
    with tempfile.TemporaryDirectory() as td:
        tdp = pathlib.Path(td)
        path = tdp / "mesh.msh"
        gmshmod.write_msh(path, m, parts=parts)
        cmd = [MPIEXEC, "-launcher", "fork", "-n", str(P), str(REF / "ref_gmsh_driver"), str(path), str(dpn),
               str(tdp)]
        subprocess.run(cmd, check=True, capture_output=True, timeout=300)
        rec = {"msh": np.frombuffer(path.read_bytes(), dtype=np.uint8), "size": np.array([P]),
               "dofs_per_node": np.array([dpn])}
        names = ["scalars", "npe", "ndpe", "local_map", "global_map", "local_dof_map", "global_dof_map",
                 "l2g_at_entries", "l2gdof_at_entries", "xyz_at_entries", "constrained_dofs", "prescribed"]
        types = [np.uint64, np.uint32, np.uint32, np.uint32, np.uint64, np.uint32, np.uint64, np.uint64, np.uint64,
                 np.float64, np.uint64, np.float64]
        for r in range(P):
            with open(tdp / f"gmsh_r{r}.bin", "rb") as f:
                for nm, ty in zip(names, types):
                    rec[f"r{r}_{nm}"] = _read_vec(f, ty)
    np.savez_compressed(OUT / f"gmsh_{name}.npz", **rec)
    print(f"wrote gmsh_{name}.npz")


def gmsh_fixtures() -> None:
    rng = np.random.default_rng(20241202)

    def scatter(E, P):
        parts = rng.integers(0, P, E)
        parts[:P] = np.arange(P)          # no empty partition
        return rng.permutation(parts)

    h8 = meshmod.hex8(3, 3, 4)
    layer = np.arange(h8.n_elem) // 9
    gmsh_case("hex8_p1", h8, 1, 3, None)
    gmsh_case("hex8_p2_slab", h8, 2, 3, (layer >= 2).astype(int))
    gmsh_case("hex8_p3_scatter", h8, 3, 3, scatter(h8.n_elem, 3))
    t10 = meshmod.tet10(2, seed=5)
    gmsh_case("tet10_p1", t10, 1, 1, None)
    gmsh_case("tet10_p2_scatter", t10, 2, 1, scatter(t10.n_elem, 2))
    gmsh_case("tet10_p4_blocks", t10, 4, 1, np.arange(t10.n_elem) * 4 // t10.n_elem)


def main() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    if shutil.which(MPIEXEC) is None and not pathlib.Path(MPIEXEC).exists():
        raise SystemExit("MPICH mpiexec not found")
    fem_fixtures()
    maps_fixtures()
    gmsh_fixtures()


if __name__ == "__main__":
    main()
