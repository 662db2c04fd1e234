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


class MatS:
    """An ElMatrix_s attached to a Fortran-ordered float32 numpy array."""

    def __init__(self, lib, a):
        self.lib = lib
        self.a = np.array(a, dtype=np.float32, order="F", copy=True)
        self.h = vp()
        assert lib.ElMatrixCreate_s(ctypes.byref(self.h)) == 0
        m, n = self.a.shape
        assert lib.ElMatrixAttach_s(self.h, m, n,
                                    self.a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    max(m, 1)) == 0

    def __del__(self):
        self.lib.ElMatrixDestroy_s(self.h)


def test_c_api_float_entry_points(lib, orc):
    """The _s entry points (C_PROTO_FIELD(s, float) in factor-C.cpp / solve-C.cpp)."""
    lib.ElMatrixCreate_s.argtypes = [ctypes.POINTER(vp)]
    lib.ElMatrixDestroy_s.argtypes = [vp]
    lib.ElMatrixAttach_s.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    for name in ("ElLUPartialPiv_s", "ElLinearSolve_s"):
        getattr(lib, name).argtypes = [vp, vp]
    lib.ElSolveAfterLUPartialPiv_s.argtypes = [ctypes.c_int, vp, vp, vp]
    lib.ElCholesky_s.argtypes = [ctypes.c_int, vp]
    lib.ElHPDSolve_s.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
    n = 300
    orc.rng_seed(50)
    A = orc.gaussian(n, n).astype(np.float32)
    B = orc.gaussian(n, 2).astype(np.float32)
    F_ref, d_ref, _ = orc.lu(A)
    M = MatS(lib, A)
    P = vp()
    assert lib.ElPermutationCreate(ctypes.byref(P)) == 0
    assert lib.ElLUPartialPiv_s(M.h, P) == 0
    d = perm_dests(lib, P)
    agree = orc.pivot_agreement(A, d_ref, d)
    assert agree["mismatches"] == 0 or agree["near_ties"] == 1, agree
    assert np.abs(M.a - F_ref).max() <= 1e3 * n * 2.0 ** -24 * np.abs(F_ref).max()
    X = MatS(lib, B)
    assert lib.ElSolveAfterLUPartialPiv_s(0, M.h, P, X.h) == 0
    X_ref = orc.lu_solve_after(F_ref, d_ref, B)
    assert np.abs(X.a - X_ref).max() <= 1e-2 * np.abs(X_ref).max()
    Y = MatS(lib, B)
    assert lib.ElLinearSolve_s(MatS(lib, A).h, Y.h) == 0
    assert np.abs(Y.a - X_ref).max() <= 1e-2 * np.abs(X_ref).max()
    G = orc.gaussian(n, n)
    H = (G @ G.T / n + np.eye(n)).astype(np.float32)
    C = MatS(lib, H)
    assert lib.ElCholesky_s(0, C.h) == 0
    Z = MatS(lib, B)
    assert lib.ElHPDSolve_s(0, 0, MatS(lib, H).h, Z.h) == 0
    Z_ref, _ = orc.hpd_solve(H, B)
    assert np.abs(Z.a - Z_ref).max() <= 1e-3 * np.abs(Z_ref).max()
    lib.ElPermutationDestroy(P)


def test_c_api_lu_without_pivoting(lib, orc):
    """ElLU_d / ElSolveAfterLU_d (factor-C.cpp:271,289): LU.cpp:21-44 and
    LU/SolveAfter.hpp:15-38."""
    lib.ElLU_d.argtypes = [vp]
    lib.ElSolveAfterLU_d.argtypes = [ctypes.c_int, vp, vp]
    rng = np.random.default_rng(51)
    n = 260
    A = rng.standard_normal((n, n))
    A += np.diag(np.abs(A).sum(0) + 1.0)
    A = np.asfortranarray(A)
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    F_ref, info = orc.lu_nopiv(A)
    M = Mat(lib, A)
    assert lib.ElLU_d(M.h) == 0
    assert np.abs(M.a - F_ref).max() <= 1e3 * n * EPS * np.abs(F_ref).max()
    X = Mat(lib, B)
    assert lib.ElSolveAfterLU_d(0, M.h, X.h) == 0
    assert orc.backward_error(A, X.a, B) <= n * EPS
    Z = np.zeros((4, 4))
    Zm = Mat(lib, Z)
    assert lib.ElLU_d(Zm.h) == 5  # SingularMatrixException -> EL_RUNTIME_ERROR


def test_c_api_lu_full_pivoting(lib, orc):
    """ElLUFullPiv_d / ElSolveAfterLUFullPiv_d (factor-C.cpp:279-283, 305-312)."""
    lib.ElLUFullPiv_d.argtypes = [vp, vp, vp]
    lib.ElSolveAfterLUFullPiv_d.argtypes = [ctypes.c_int, vp, vp, vp, vp]
    rng = np.random.default_rng(73)
    n = 200
    A = np.asfortranarray(rng.standard_normal((n, n)))
    B = np.asfortranarray(rng.standard_normal((n, 2)))
    F_ref, p_ref, q_ref, _ = orc.lu_full(A)
    M = Mat(lib, A)
    P, Q = vp(), vp()
    lib.ElPermutationCreate(ctypes.byref(P))
    lib.ElPermutationCreate(ctypes.byref(Q))
    assert lib.ElLUFullPiv_d(M.h, P, Q) == 0
    assert np.array_equal(perm_dests(lib, P), p_ref) and np.array_equal(perm_dests(lib, Q), q_ref)
    assert np.array_equal(M.a, F_ref)
    X = Mat(lib, B)
    assert lib.ElSolveAfterLUFullPiv_d(0, M.h, P, Q, X.h) == 0
    assert orc.backward_error(A, X.a, B) <= n * EPS
    lib.ElPermutationDestroy(P)
    lib.ElPermutationDestroy(Q)
