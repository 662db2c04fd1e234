"""elemental_amd -- MI355X-native dense factor-and-solve behind Elemental's API.

The HIP library (libelemental_amd.so, built by `python -m elemental_amd.build`) exports
  * the device ABI of include/elemental_amd.h (elx_dgetrf, elx_dgetrs, elx_dgesv,
    elx_dpotrf, elx_dpotrs, elx_dposv, elx_dgemm, elx_dtrsm, elx_dsyrk, elx_dpermute_rows,
    elx_dgesv_mixed, elx_dposv_mixed), and
  * Elemental's own C++ (include/El.hpp) and C (include/El.h) interfaces for the path.
`elemental_amd.device` wraps the device ABI for column-major torch tensors.
"""
from ._lib import LIB_PATH, ElxError, IrStats, load

__all__ = ["LIB_PATH", "ElxError", "IrStats", "load"]
