"""Element matrices of the examples through the C ABI (amx_fem_*).

Row-major numpy outputs; element node coordinates xe are (nodes, 3) arrays in
the node order of examples/include/shapeFunc.hpp.
"""
from __future__ import annotations

import numpy as np

from ._capi import call


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def gauss(n: int) -> np.ndarray:
    """integration<double>(n).Pts_n_Wts: [x0, w0, x1, w1, ...]."""
    out = np.zeros(2 * n)
    call("amx_fem_gauss", n, out.ctypes.data)
    return out


def ke_hex8(xe, ngt: int = 2) -> np.ndarray:
    x = _c(xe).reshape(8, 3)
    ke = np.zeros((8, 8))
    call("amx_fem_ke_hex8", x.ctypes.data, ngt, ke.ctypes.data)
    return ke


def ke_hex20_iso(xe, E: float, nu: float, ngt: int = 3) -> np.ndarray:
    x = _c(xe).reshape(20, 3)
    ke = np.zeros((60, 60))
    call("amx_fem_ke_hex20_iso", x.ctypes.data, float(E), float(nu), ngt, ke.ctypes.data)
    return ke


def ke_tet10(xe, ngt: int = 2) -> np.ndarray:
    x = _c(xe).reshape(10, 3)
    ke = np.zeros((10, 10))
    call("amx_fem_ke_tet10", x.ctypes.data, ngt, ke.ctypes.data)
    return ke


def fe_tet10(xe, be, ngt: int = 2) -> np.ndarray:
    x = _c(xe).reshape(10, 3)
    b = _c(be).reshape(10)
    fe = np.zeros(10)
    call("amx_fem_fe_tet10", x.ctypes.data, b.ctypes.data, ngt, fe.ctypes.data)
    return fe


def ke_tet10_batch(coords, ngt: int = 2) -> np.ndarray:
    x = _c(coords).reshape(-1, 10, 3)
    ke = np.zeros((x.shape[0], 10, 10))
    call("amx_fem_ke_tet10_batch", x.shape[0], x.ctypes.data, ngt, ke.ctypes.data)
    return ke


def fe_tet10_batch(coords, be, ngt: int = 2) -> np.ndarray:
    x = _c(coords).reshape(-1, 10, 3)
    b = _c(be).reshape(-1, 10)
    fe = np.zeros((x.shape[0], 10))
    call("amx_fem_fe_tet10_batch", x.shape[0], x.ctypes.data, b.ctypes.data, ngt, fe.ctypes.data)
    return fe
