"""The C ABI library loads and exports every entry point of include/amat_amd.h
(no compute calls: these run without a GPU)."""
import ctypes
import re

from conftest import ROOT


def header_functions():
    text = (ROOT / "include" / "amat_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(amx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    fns = header_functions()
    assert len(fns) >= 50
    assert "amx_op_matvec_dev" in fns and "amx_maps_set_map" in fns


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "amat_amd" / "libamat_amd.so"))
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from amat_amd import _capi
    assert set(header_functions()) == set(_capi.SIGNATURES), set(header_functions()) ^ set(_capi.SIGNATURES)


def test_version_and_status_without_device():
    import amat_amd
    assert amat_amd.version().startswith("amat_amd")
    assert amat_amd.device_count() >= 0


def test_op_create_reports_errors():
    import numpy as np
    import pytest
    import amat_amd
    from amat_amd import Maps, aMatFree, AmatError
    m = Maps()
    with pytest.raises(AmatError) as ei:
        aMatFree(m)  # maps not set
    assert ei.value.name == "INVALID_STATE"
    m.set_map(1, np.arange(8, dtype=np.uint32)[None], None, 8, np.arange(8), 0, 7, 8)
    with pytest.raises(AmatError) as ei:
        aMatFree(m, 7)
    assert ei.value.name == "UNKNOWN_BC_METH"
    with pytest.raises(AmatError):
        aMatFree(m, 0, device=10 ** 6)
