"""Elemental's own interfaces on the GPU path: the C API through ctypes, exactly as the
reference's Python binding calls it (python/lapack_like/factor.py:525-548,
python/lapack_like/solve.py:49-59), and the C++ API through a driver mirroring
tests/lapack_like/LU.cpp and tests/lapack_like/Cholesky.cpp. Results are compared with the
oracle on the same inputs."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS = 2.0 ** -53
vp = ctypes.c_void_p


@pytest.fixture(scope="module")
def lib():
    from elemental_amd import load
    L = load()
    L.ElMatrixCreate_d.argtypes = [ctypes.POINTER(vp)]
    L.ElMatrixDestroy_d.argtypes = [vp]
    L.ElMatrixAttach_d.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.ElPermutationCreate.argtypes = [ctypes.POINTER(vp)]
    L.ElPermutationDestroy.argtypes = [vp]
    L.ElPermutationSwapDestinations.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_int)]
    for name in ("ElLUPartialPiv_d", "ElLinearSolve_d"):
        getattr(L, name).argtypes = [vp, vp]
    L.ElSolveAfterLUPartialPiv_d.argtypes = [ctypes.c_int, vp, vp, vp]
    L.ElCholesky_d.argtypes = [ctypes.c_int, vp]
    L.ElSolveAfterCholesky_d.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
    L.ElHPDSolve_d.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
    return L


class Mat:
    """An ElMatrix_d attached to a Fortran-ordered numpy array."""

    def __init__(self, lib, a):
        self.lib = lib
        self.a = np.array(a, dtype=np.float64, order="F", copy=True)
        self.h = vp()
        assert lib.ElMatrixCreate_d(ctypes.byref(self.h)) == 0
        m, n = self.a.shape
        assert lib.ElMatrixAttach_d(self.h, m, n,
                                    self.a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    max(m, 1)) == 0

    def __del__(self):
        self.lib.ElMatrixDestroy_d(self.h)


def perm_dests(lib, P):
    cnt = ctypes.c_int()
    lib.ElPermutationSwapDestinations(P, None, ctypes.byref(cnt))
    buf = (ctypes.c_int * max(cnt.value, 1))()
    lib.ElPermutationSwapDestinations(P, buf, ctypes.byref(cnt))
    return np.array(buf[: cnt.value], dtype=np.int32)


def test_c_api_lu_and_solve(lib, orc):
    n = 500
    orc.rng_seed()
    A = orc.gaussian(n, n)
    B = orc.gaussian(n, 4)
    F_ref, d_ref, _ = orc.lu(A)
    M = Mat(lib, A)
    P = vp()
    assert lib.ElPermutationCreate(ctypes.byref(P)) == 0
    assert lib.ElLUPartialPiv_d(M.h, P) == 0
    d = perm_dests(lib, P)
    agree = orc.pivot_agreement(A, d_ref, d)
    assert agree["mismatches"] == 0 or agree["near_ties"] == 1
    assert orc.lu_factor_error(A, M.a, d) <= 8 * n * EPS
    X = Mat(lib, B)
    assert lib.ElSolveAfterLUPartialPiv_d(0, M.h, P, X.h) == 0
    assert orc.backward_error(A, X.a, B) <= n * EPS
    lib.ElPermutationDestroy(P)


def test_c_api_linear_and_hpd_solve(lib, orc):
    n = 400
    orc.rng_seed(3)
    A = orc.gaussian(n, n)
    B = orc.gaussian(n, 2)
    X = Mat(lib, B)
    Am = Mat(lib, A)
    assert lib.ElLinearSolve_d(Am.h, X.h) == 0
    assert np.array_equal(Am.a, A)  # LinearSolve works on a copy of A
    assert orc.backward_error(A, X.a, B) <= n * EPS
    G = orc.gaussian(n, n)
    H = np.asfortranarray(G @ G.T / n + np.eye(n))
    for uplo in (0, 1):
        Y = Mat(lib, B)
        Hm = Mat(lib, H)
        assert lib.ElHPDSolve_d(uplo, 0, Hm.h, Y.h) == 0
        assert orc.backward_error(H, Y.a, B) <= n * EPS
        C = Mat(lib, H)
        assert lib.ElCholesky_d(uplo, C.h) == 0
        assert orc.chol_factor_error(H, C.a, uplo) <= 8 * n * EPS
        Z = Mat(lib, B)
        assert lib.ElSolveAfterCholesky_d(uplo, 0, C.h, Z.h) == 0
        assert np.abs(Z.a - Y.a).max() <= 1e-10 * np.abs(Y.a).max()


def test_c_api_errors_like_el_try(lib):
    Z = np.random.default_rng(0).standard_normal((50, 50))
    Z[:, 9] = 0.0
    M = Mat(lib, Z)
    P = vp()
    lib.ElPermutationCreate(ctypes.byref(P))
    assert lib.ElLUPartialPiv_d(M.h, P) == 5  # SingularMatrixException -> EL_RUNTIME_ERROR
    N = Mat(lib, np.diag([1.0, 2.0, -1.0, 4.0]))
    assert lib.ElCholesky_d(0, N.h) == 5      # NonHPDMatrixException -> EL_RUNTIME_ERROR
    lib.ElPermutationDestroy(P)


def test_cpp_api_driver():
    exe = os.path.join(ROOT, "tests", "cpp", "el_api_test")
    assert os.path.exists(exe), "build() compiles tests/cpp/el_api_test"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
