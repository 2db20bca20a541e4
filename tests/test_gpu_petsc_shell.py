"""The PETSc shell-matrix path of the C++ drop-in, driven as the reference's
par::solve drives it (examples/include/solve.hpp:31-83): get_comm,
get_matrix (a MATSHELL whose context is a heap aMatCTX), MATOP_MULT and
MATOP_GET_DIAGONAL called by the Krylov loop, MATOP_GET_DIAGONAL_BLOCK
answering PETSC_ERR_SUP (block Jacobi is not provided), then
MatShellGetContext + delete.  tests/cpp/petsc_shell_example is built against
the test double tests/cpp/petsc_stub/petsc.h (PETSc is absent) and real MPICH;
on two ranks the Maps run set_map's collectives on MPI_COMM_WORLD and the
operator's halo exchanges are MPI messages.

Bar: one rank -- MatMult through the shell is bit-exact against the
oracle's MatMult_mf, the diagonal and the owned-length matvec bit-exact too;
two ranks -- 1e-14 relative L2 (a rank adds its ghosts' contributions after
its own elements); the CG solution (residual 1e-10) within 1e-8 of scipy's direct solve of
the same system.
"""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as orc
from amat_amd import fem, mesh
from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu

EXE = ROOT / "tests" / "cpp" / "petsc_shell_example"
MPIEXEC = pathlib.Path("/opt/conda/bin/mpiexec")
DIMS = (6, 5, 8)


def reference(nx, ny, nz):
    """The same problem on one rank: maps, K_e = E_e K_ref, boundary."""
    m = mesh.hex8(nx, ny, nz)
    gmap = m.dof_map()
    xe = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                  dtype=np.float64) * np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    kref = fem.ke_hex8(xe, 2)
    scale = 0.5 + 0.25 * (np.arange(m.n_elem) % 7)
    K = scale[:, None, None] * kref[None]
    cons = mesh.boundary_nodes(m).astype(np.uint64)
    return m, gmap, K, cons


def run(outdir, ranks):
    cmd = [str(EXE), str(outdir), *map(str, DIMS)]
    if ranks > 1:
        cmd = [str(MPIEXEC), "-launcher", "fork", "-n", str(ranks), *cmd]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=180)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = sorted(ln.split() for ln in r.stdout.splitlines() if ln.startswith("rank "))
    assert len(lines) == ranks
    parts = {}
    for ln in lines:
        kv = dict(zip(ln[0::2], ln[1::2]))
        rk = int(kv["rank"])
        assert kv["bad"] == "0" and kv["bjacobi_rc"] == "56"   # PETSC_ERR_SUP of the stub
        parts[rk] = (int(kv["owned_begin"]), int(kv["n_owned"]), int(kv["cg_iterations"]))
    out = {}
    for name in ("u", "v", "d", "x", "kv"):
        out[name] = np.concatenate([np.fromfile(outdir / f"{name}.r{rk}.bin", dtype=np.float64) for rk in sorted(parts)])
    offs = [parts[rk][0] for rk in sorted(parts)]
    assert offs == sorted(offs)
    return out, parts


@pytest.mark.parametrize("ranks", [1, 2])
def test_par_solve_sequence_on_the_shell(tmp_path, ranks):
    assert EXE.exists(), "build with python -m amat_amd.build"
    if ranks > 1 and not MPIEXEC.exists():
        pytest.skip("MPICH mpiexec is not installed")
    out, parts = run(tmp_path, ranks)
    m, gmap, K, cons = reference(*DIMS)
    n = m.n_dofs
    g = np.arange(n, dtype=np.int64)
    np.testing.assert_array_equal(out["u"], ((g * 7919) % 2001 - 1000) / 1000.0)
    u = out["u"]
    want_v = orc.matmult_mf(gmap, K, u, cons)
    bmaps = orc.set_bdr_map(0, 1, [n], gmap, cons, np.zeros(len(cons)))
    want_d = orc.get_diagonal(gmap, K, bmaps["bdr_map"], 0, n)
    want_kv = orc.matvec(gmap, K, u)
    if ranks == 1:
        np.testing.assert_array_equal(out["v"], want_v)
        np.testing.assert_array_equal(out["d"], want_d)
        np.testing.assert_array_equal(out["kv"], want_kv)
    else:
        assert rel_l2(out["v"], want_v) <= 1e-14
        assert rel_l2(out["d"], want_d) <= 1e-14
        assert rel_l2(out["kv"], want_kv) <= 1e-14
    # the system the loop solved: MatMult_mf's matrix (identity rows), b = 1 on free rows
    rows, cols, vals = [], [], []
    c = cons.astype(np.int64)
    for e in range(gmap.shape[0]):
        rr, cc = np.meshgrid(gmap[e], gmap[e], indexing="ij")
        rows.append(rr.ravel()); cols.append(cc.ravel()); vals.append(K[e].ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tolil()
    A[c, :] = 0.0
    A[:, c] = 0.0
    A[c, c] = 1.0
    b = np.ones(n)
    b[c] = 0.0
    x = spla.spsolve(A.tocsc(), b)
    assert rel_l2(out["x"], x) <= 1e-8   # residual 1e-10 times the condition number
    assert all(p[2] == parts[0][2] for p in parts.values())   # one iteration count on every rank
