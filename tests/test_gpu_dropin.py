"""The header-only C++ drop-in (include/amat_amd/aMatFree.hpp) used the way
examples/src/example03a.cpp uses the reference (and example07a.cpp sets up
method 2, example05a.cpp reads a Gmsh mesh through
include/amat_amd/GmshMesh.hpp): a compiled C++ program runs the GPU operator
and checks it against the oracle."""
import subprocess

import pytest

from amat_amd import gmsh, mesh
from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_dropin_example(tmp_path):
    exe = ROOT / "tests" / "cpp" / "dropin_example"
    assert exe.exists(), "build with python -m amat_amd.build"
    msh = tmp_path / "tet10.msh"
    gmsh.write_msh(msh, mesh.tet10(3, seed=8))
    dump = tmp_path / "matrix.out"
    r = subprocess.run([str(exe), str(msh), str(dump)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("maxdiff_matvec=0.000e+00 maxdiff_mf=0.000e+00") == 3
    assert "dropin gmsh_tet10_162 maxdiff_matvec=0.000e+00 maxdiff_mf=0.000e+00 constraints=218" in r.stdout
    assert "error-path code=11" in r.stdout
    assert "timers=ok" in r.stdout and "timers=bad" not in r.stdout
    # example07a's driver pattern through the par::aMat base: dump_mat, get_timer
    line = next(ln for ln in r.stdout.splitlines() if ln.startswith("dropin base_pointer_4x5x6"))
    assert "symmetric=ok" in line and "base=ok" in line
    assert float(line.split("dump_apply_rel=")[1].split()[0]) <= 1e-13
    assert dump.exists()
    # method 2: the callback form is stored (bit-exact), the in-kernel form within 1e-13 of max |v|
    assert "dropin free_callback_7x6x5 maxdiff_matvec=0.000e+00" in r.stdout
    rel = float(r.stdout.split("free_inkernel_7x6x5 relmax=")[1].split()[0])
    assert rel <= 1e-13
