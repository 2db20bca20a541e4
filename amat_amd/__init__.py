"""amat_amd: MI355X-native assembly-free element-by-element operator.

The product is the C ABI in include/amat_amd.h (amat_amd/libamat_amd.so:
host C++ and gfx950 HIP kernels).  This package holds its Python binding
(Maps, aMatFree mirroring paralab/aMat's classes), mesh and workload builders,
and the build script.
"""
from .amat import (BC_METH, BCOP, ELEM, GHOST_TO_OWNER, HALO, MATFREE_TYPE, MATRIX_TYPE, OWNER_TO_GHOST, PREC,
                   SCATTER, AmatError, Error, Maps, MultiRHS, aMatFree, device_count, version)

__all__ = ["AmatError", "BC_METH", "BCOP", "ELEM", "Error", "GHOST_TO_OWNER", "HALO", "MATFREE_TYPE", "MATRIX_TYPE",
           "Maps", "MultiRHS", "OWNER_TO_GHOST", "PREC", "SCATTER", "aMatFree", "device_count", "version"]
