"""Builds the native parts of amat_amd in-tree.

    python -m amat_amd.build            # library + oracle (+ reference objects when present)

* amat_amd/libamat_amd.so -- the C ABI of include/amat_amd.h: host C++ and the
  gfx950 HIP kernels, compiled with hipcc for gfx950 only.
* tests/cpp/dropin_example -- a C++ program that uses the header-only
  par::aMatFree drop-in (include/amat_amd/aMatFree.hpp) over the C ABI.
* oracle/liboracle.so (and oracle/_ref/ when the reference tree is mounted) via
  oracle/Makefile -- test infrastructure, never linked into the library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent
CSRC = ROOT / "amat_amd" / "csrc"
BUILD = ROOT / "_build"
LIB = ROOT / "amat_amd" / "libamat_amd.so"
ROCM = pathlib.Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"
REFERENCE = pathlib.Path("/root/reference")

SOURCES = [
    "amx_common.cpp",
    "amx_maps.cpp",
    "amx_fem.cpp",
    "amx_op.cpp",
    "amx_solver.cpp",
    "amx_serve.cpp",
    "amx_gmsh.cpp",
    "amx_kernels.hip",
    "amx_solver_kernels.hip",
    "amx_mrhs.hip",
    "amx_free.hip",
    "amx_noalloc.hip",
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fopenmp", "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    noisy = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def _compile(src: str) -> pathlib.Path:
    obj = BUILD / (src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *COMMON, *lang, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_library() -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fopenmp", *map(str, objs), "-o", str(tmp),
          f"-Wl,-rpath,{ROCM / 'lib'}", f"-Wl,-rpath,{ROCM / 'lib' / 'llvm' / 'lib'}"])
    os.replace(tmp, LIB)
    return LIB


def build_dropin_example() -> pathlib.Path | None:
    src = ROOT / "tests" / "cpp" / "dropin_example.cpp"
    if not src.exists():
        return None
    out = src.with_suffix("")
    _run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{ROOT / 'oracle'}", str(src),
          f"-L{LIB.parent}", "-lamat_amd", f"-Wl,-rpath,{LIB.parent}", f"-L{ROOT / 'oracle'}", "-loracle",
          f"-Wl,-rpath,{ROOT / 'oracle'}", "-o", str(out)])
    return out


MPI_HOME = pathlib.Path(os.environ.get("MPI_HOME", "/opt/conda"))


def build_petsc_shell_example() -> pathlib.Path | None:
    """tests/cpp/petsc_shell_example: the BUILD_WITH_PETSC form of the C++
    drop-in against the test double tests/cpp/petsc_stub/petsc.h and MPICH
    (skipped when MPICH is absent).  libmpi is linked by path and the system
    library directory leads the run path, so that the HIP runtime finds the
    system libstdc++ rather than the older one beside MPICH."""
    src = ROOT / "tests" / "cpp" / "petsc_shell_example.cpp"
    libmpi = MPI_HOME / "lib" / "libmpi.so.12"
    if not src.exists() or not (MPI_HOME / "include" / "mpi.h").exists() or not libmpi.exists():
        return None
    out = src.with_suffix("")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-DBUILD_WITH_PETSC", f"-I{ROOT / 'include'}",
          f"-I{ROOT / 'tests' / 'cpp' / 'petsc_stub'}", f"-I{MPI_HOME / 'include'}", str(src), f"-L{LIB.parent}",
          "-lamat_amd", str(libmpi), f"-Wl,-rpath,{LIB.parent}:/usr/lib/x86_64-linux-gnu:{MPI_HOME / 'lib'}",
          "-o", str(out)])
    return out


def build_oracle() -> None:
    if shutil.which("make") is None:
        raise RuntimeError("make is required to build the oracle")
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    if REFERENCE.is_dir():
        _run(["make", "-s", "-C", str(ROOT / "oracle"), "ref", f"REF={REFERENCE}"])


def main() -> None:
    build_library()
    build_oracle()
    build_dropin_example()
    build_petsc_shell_example()
    print(f"built {LIB}")


if __name__ == "__main__":
    main()
