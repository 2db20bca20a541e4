"""The header-only C++ drop-in (include/amat_amd/aMatFree.hpp) used the way
examples/src/example03a.cpp uses the reference (and example07a.cpp sets up
method 2): a compiled C++ program runs the GPU operator and checks it against
the oracle."""
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_dropin_example():
    exe = ROOT / "tests" / "cpp" / "dropin_example"
    assert exe.exists(), "build with python -m amat_amd.build"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("maxdiff_matvec=0.000e+00 maxdiff_mf=0.000e+00") == 2
    assert "error-path code=11" in r.stdout
    # method 2: the callback form is stored (bit-exact), the in-kernel form within 1e-13 of max |v|
    assert "dropin free_callback_7x6x5 maxdiff_matvec=0.000e+00" in r.stdout
    rel = float(r.stdout.split("free_inkernel_7x6x5 relmax=")[1].split()[0])
    assert rel <= 1e-13
