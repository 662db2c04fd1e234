"""ctypes binding of libelemental_amd.so (the C ABI in include/elemental_amd.h and
include/El.h). Loading fails loudly when the library has not been built: there is no
fallback path."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libelemental_amd.so")

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_flt = ctypes.c_float
c_vp = ctypes.c_void_p

ELX_STATUS = {
    0: "ELX_SUCCESS",
    1: "ELX_ERROR_INVALID_ARG",
    2: "ELX_ERROR_HIP",
    3: "ELX_ERROR_ALLOC",
    4: "ELX_ERROR_NOT_SUPPORTED",
    5: "ELX_ERROR_NUMERIC",
}


class ElxError(RuntimeError):
    def __init__(self, fn, status):
        super().__init__(f"{fn} returned {ELX_STATUS.get(status, status)}")
        self.status = status


class IrStats(ctypes.Structure):
    _fields_ = [
        ("converged", c_i32),
        ("outer_iters", c_i32),
        ("inner_iters", c_i32),
        ("fell_back", c_i32),
        ("factor_info", c_i32),
        ("near_tie_pivots", c_i32),
        ("backward_error", c_dbl),
        ("factor_seconds", c_dbl),
        ("refine_seconds", c_dbl),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> argtypes (restype c_int unless noted)
_SIGS = {
    "elx_device_count": [],
    "elx_finalize": [],
    "elx_dgetrf": [c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp],
    "elx_dgetrs": [c_vp, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64],
    "elx_dgetrs_noalloc": [c_vp, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64],
    "elx_dgesv": [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp],
    "elx_dpotrf": [c_vp, c_int, c_i64, c_vp, c_i64, c_i64, c_vp],
    "elx_dpotrs": [c_vp, c_int, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64],
    "elx_dposv": [c_vp, c_int, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp],
    "elx_dgemm": [c_vp, c_int, c_int, c_i64, c_i64, c_i64, c_dbl, c_vp, c_i64, c_vp, c_i64,
                  c_dbl, c_vp, c_i64],
    "elx_dtrsm": [c_vp, c_int, c_int, c_int, c_int, c_i64, c_i64, c_dbl, c_vp, c_i64, c_vp,
                  c_i64],
    "elx_dsyrk": [c_vp, c_int, c_int, c_i64, c_i64, c_dbl, c_vp, c_i64, c_dbl, c_vp, c_i64],
    "elx_dpermute_rows": [c_vp, c_int, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64],
    "elx_dgetrf_nopiv": [c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp],
    "elx_sgetrf": [c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp],
    "elx_sgetrf_nopiv": [c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp],
    "elx_sgetrs": [c_vp, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64],
    "elx_sgesv": [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp],
    "elx_spotrf": [c_vp, c_int, c_i64, c_vp, c_i64, c_i64, c_vp],
    "elx_spotrs": [c_vp, c_int, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64],
    "elx_sposv": [c_vp, c_int, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp],
    "elx_sgemm": [c_vp, c_int, c_int, c_i64, c_i64, c_i64, c_flt, c_vp, c_i64, c_vp, c_i64,
                  c_flt, c_vp, c_i64],
    "elx_strsm": [c_vp, c_int, c_int, c_int, c_int, c_i64, c_i64, c_flt, c_vp, c_i64, c_vp,
                  c_i64],
    "elx_ssyrk": [c_vp, c_int, c_int, c_i64, c_i64, c_flt, c_vp, c_i64, c_flt, c_vp, c_i64],
    "elx_spermute_rows": [c_vp, c_int, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64],
    "elx_dgetrf_full": [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp],
    "elx_sgetrf_full": [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp],
    "elx_dgetrs_full": [c_vp, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64],
    "elx_sgetrs_full": [c_vp, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64],
    "elx_druiz_equil": [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp],
    "elx_dsymmetric_ruiz_equil": [c_vp, c_i64, c_vp, c_i64, c_vp, c_i32],
    "elx_dgeom_equil": [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp],
    "elx_dgesv_mixed": [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_int, c_int,
                        c_dbl, c_i32, ctypes.POINTER(IrStats)],
    "elx_dposv_mixed": [c_vp, c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_int,
                        c_int, c_dbl, c_i32, ctypes.POINTER(IrStats)],
    "elx_dgesv_mixed_batched": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                c_int, c_int, c_dbl, c_i32, c_vp],
    "elx_dpotrf_vbatched": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp],
    "elx_spotrf_vbatched": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp],
    "elx_dpotrs_vbatched": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "elx_spotrs_vbatched": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "elx_dposv_vbatched_graph": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp],
    "elx_sposv_vbatched_graph": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp],
    "elx_graph_launch": [c_vp, c_vp],
    "elx_graph_destroy": [c_vp],
    "elx_batched_set_variant": [c_int, c_int, c_int, c_int, c_int],
    "elx_batched_get_variant": [c_int, c_int, c_vp, c_vp, c_vp],
}

_lib = None


def load():
    """Load the built library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m elemental_amd.build` "
            "(there is no CPU fallback for the HIP path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, argtypes in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = c_int
    lib.elx_version.restype = ctypes.c_char_p
    lib.elx_version.argtypes = []
    _lib = lib
    return lib


def call(name, *args):
    """Call an elx_* entry point and raise ElxError on a non-zero status."""
    st = getattr(load(), name)(*args)
    if st != 0:
        raise ElxError(name, st)
    return st
