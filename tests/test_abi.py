"""The C-ABI library loads on the CPU and exports every entry point the headers declare.
No compute is called here (there is no GPU in the build container): only argument
validation, which returns before any device work."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["include/elemental_amd.h", "include/El.h"]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(os.path.join(ROOT, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s+(\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_the_path():
    names = declared_functions()
    for required in ["elx_dgetrf", "elx_dgetrs", "elx_dgesv", "elx_dpotrf", "elx_dpotrs",
                     "elx_dposv", "elx_dgemm", "elx_dtrsm", "elx_dsyrk", "elx_dgesv_mixed",
                     "ElLUPartialPiv_d", "ElSolveAfterLUPartialPiv_d", "ElCholesky_d",
                     "ElLinearSolve_d", "ElHPDSolve_d",
                     # the float instantiations and the unpivoted LU
                     "elx_sgetrf", "elx_sgetrs", "elx_sgesv", "elx_spotrf", "elx_spotrs",
                     "elx_sposv", "elx_sgemm", "elx_strsm", "elx_ssyrk", "elx_spermute_rows",
                     "elx_dgetrf_nopiv", "elx_sgetrf_nopiv",
                     "ElLU_d", "ElLU_s", "ElSolveAfterLU_d", "ElSolveAfterLU_s",
                     "ElLUPartialPiv_s", "ElSolveAfterLUPartialPiv_s", "ElCholesky_s",
                     "ElSolveAfterCholesky_s", "ElLinearSolve_s", "ElHPDSolve_s", "ElGemm_s",
                     "ElTrsm_s", "ElHerk_s", "ElSyrk_s", "ElMatrixCreate_s",
                     "elx_dgetrf_full", "elx_sgetrf_full", "elx_dgetrs_full", "elx_sgetrs_full",
                     "ElLUFullPiv_d", "ElLUFullPiv_s", "ElSolveAfterLUFullPiv_d",
                     "ElSolveAfterLUFullPiv_s"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from elemental_amd import LIB_PATH, load
    lib = load()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert getattr(lib, n) is not None


def test_cpp_api_is_exported():
    from elemental_amd import LIB_PATH
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in ["El::LU<double>(El::Matrix<double>&, El::Permutation&)",
                "El::Cholesky<double>(El::UpperOrLower, El::Matrix<double>&)",
                "El::LinearSolve<double>(El::Matrix<double> const&, El::Matrix<double>&)",
                "El::lu::SolveAfter<double>(", "El::cholesky::SolveAfter<double>(",
                "El::HPDSolve<double>(",
                "El::LU<float>(El::Matrix<float>&, El::Permutation&)",
                "El::LU<float>(El::Matrix<float>&)", "El::LU<double>(El::Matrix<double>&)",
                "El::lu::SolveAfter<float>(El::Orientation, El::Matrix<float> const&, El::Matrix<float>&)",
                "El::Cholesky<float>(", "El::LinearSolve<float>(", "El::HPDSolve<float>(",
                "El::Gemm<float>(", "El::Trsm<float>(", "El::Herk<float>(", "El::Syrk<float>(",
                "El::LU<double>(El::Matrix<double>&, El::Permutation&, El::Permutation&)"]:
        assert sym in out, sym


def test_argument_errors_without_device_work():
    from elemental_amd import load
    lib = load()
    info = ctypes.c_int32(0)
    # lda < m
    st = lib.elx_dgetrf(None, 10, 10, None, 5, 128, None, ctypes.byref(info))
    assert st == 1
    # block size out of range (any positive block size is accepted)
    st = lib.elx_dgetrf(None, 10, 10, None, 10, 0, None, ctypes.byref(info))
    assert st == 1
    st = lib.elx_sgetrf(None, 10, 10, None, 5, 128, None, ctypes.byref(info))
    assert st == 1
    st = lib.elx_dpotrf(None, 7, 10, None, 10, 128, ctypes.byref(info))
    assert st == 1
    assert lib.elx_version().startswith(b"elemental_amd")


def test_c_api_errors_map_like_el_try():
    """Non-square A is a logic_error -> EL_LOGIC_ERROR, as EL_TRY maps it
    (include/El/core/CReflect.hpp:12-26)."""
    from elemental_amd import load
    lib = load()
    A = ctypes.c_void_p()
    assert lib.ElMatrixCreate_d(ctypes.byref(A)) == 0
    assert lib.ElMatrixResize_d(A, 3, 4) == 0
    assert lib.ElCholesky_d(0, A) == 4  # EL_LOGIC_ERROR
    assert lib.ElMatrixResize_d(A, -1, 4) == 4
    assert lib.ElMatrixDestroy_d(A) == 0
    bs = ctypes.c_int(0)
    assert lib.ElBlocksize(ctypes.byref(bs)) == 0 and bs.value == 128
