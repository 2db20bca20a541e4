"""The BUILD_WITH_PETSC block of the C++ drop-in header compiles (CPU): the
reference's shell-matrix interface -- aMatCTX, aMat_matvec /
aMat_matgetdiagonal / aMat_matgetdiagonalblock, get_matrix, get_comm,
MatMult_mf, MatGetDiagonal_mf, MatGetDiagonalBlock_mf, apply_bc(Vec) -- and
the MPI set-up and halo transport, against the test double
tests/cpp/petsc_stub/petsc.h and MPICH's mpi.h."""
import pathlib
import subprocess

import pytest

from conftest import ROOT

MPI_INC = pathlib.Path("/opt/conda/include")


@pytest.mark.skipif(not (MPI_INC / "mpi.h").exists(), reason="MPICH headers are not installed")
def test_petsc_block_compiles():
    src = ROOT / "tests" / "cpp" / "petsc_shell_example.cpp"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-DBUILD_WITH_PETSC",
                        f"-I{ROOT / 'include'}", f"-I{ROOT / 'tests' / 'cpp' / 'petsc_stub'}", f"-I{MPI_INC}",
                        str(src)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not (MPI_INC / "mpi.h").exists(), reason="MPICH headers are not installed")
def test_mpi_only_form_compiles(tmp_path):
    """AMAT_AMD_WITH_MPI without PETSc: Maps(MPI_Comm) and the MPI halo transport."""
    src = tmp_path / "mpi_only.cpp"
    src.write_text(
        '#include "amat_amd/aMatFree.hpp"\n'
        "int main(int argc, char** argv) {\n"
        "  MPI_Init(&argc, &argv);\n"
        "  par::Maps<double, unsigned long, unsigned int> maps(MPI_COMM_WORLD);\n"
        "  (void)maps.get_comm();\n"
        "  MPI_Finalize();\n"
        "  return 0;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-DAMAT_AMD_WITH_MPI",
                        f"-I{ROOT / 'include'}", f"-I{MPI_INC}", str(src)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
