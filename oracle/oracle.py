"""Python binding of the CPU oracle (liboracle.so, built from el_oracle.cpp by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, as the checker or as the timed CPU baseline. Every function
restates the reference (file:line citations are in el_oracle.cpp). Arrays are numpy,
column-major (Fortran order), float64 or float32 (the factorizations, solves and level-3
routines run in the dtype of their input, through the orc_* or orc_s* entry points); pivots
are int32 swap destinations.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)

_lib = None

ELEMENTAL_SEED = (21 << 16) | 0  # src/core/random.cpp:24-32, rank 0, deterministic default
EPS = 2.0 ** -53                  # El::limits::Epsilon<double>() (include/El/core/limits.hpp:129)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_set_backend.argtypes = [ctypes.c_char_p]
        L.orc_backend.restype = ctypes.c_char_p
        L.orc_backward_error.restype = _dbl
        L.orc_lu_factor_error.restype = _dbl
        L.orc_chol_factor_error.restype = _dbl
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need Fortran float64"
    return a.ctypes.data_as(_dp)


def _pt(a):
    """Pointer to a Fortran-ordered float64 or float32 array."""
    assert a.flags.f_contiguous or a.ndim == 1, "need Fortran order"
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    assert a.dtype == np.float32, a.dtype
    return a.ctypes.data_as(_fp)


def _dt(A):
    """float32 for float32 input, float64 otherwise."""
    return np.float32 if np.asarray(A).dtype == np.float32 else np.float64


def _fn(name, dt):
    """orc_<name> for float64, orc_s<name> for float32."""
    return getattr(lib(), "orc_" + ("s" if dt == np.float32 else "") + name)


def _sc(x, dt):
    return ctypes.c_float(x) if dt == np.float32 else _dbl(x)


def _ip_(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _ld(a):
    return max(a.shape[0], 1)


def set_backend(name="naive"):
    st = lib().orc_set_backend(name.encode())
    if st != 0:
        raise RuntimeError(f"BLAS backend {name!r} not available")
    return lib().orc_backend().decode()


def set_threads(n):
    lib().orc_set_threads(int(n))


def rng_seed(seed=ELEMENTAL_SEED):
    lib().orc_rng_seed(ctypes.c_uint64(seed))


def gaussian(m, n, mean=0.0, stddev=1.0):
    """El::Gaussian(A, m, n) from the oracle's mt19937 stream (continues the stream)."""
    A = np.zeros((m, n), order="F")
    lib().orc_gaussian(_p(A), _i64(m), _i64(n), _i64(_ld(A)), _dbl(mean), _dbl(stddev))
    return A


def gepp_growth(n):
    A = np.zeros((n, n), order="F")
    lib().orc_geppgrowth(_p(A), _i64(n), _i64(_ld(A)))
    return A


def lu(A, bsize=128):
    """El::LU(A, P) on a copy, in A's dtype; returns (factors, dests, info)."""
    dt = _dt(A)
    F = np.array(A, dtype=dt, order="F", copy=True)
    m, n = F.shape
    d = np.zeros(max(min(m, n), 1), np.int32)
    info = _fn("lu", dt)(_pt(F), _i64(m), _i64(n), _i64(_ld(F)), _i64(bsize), _ip_(d))
    return F, d[: min(m, n)], info


def lu_nopiv(A, bsize=128):
    """El::LU(A) without pivoting on a copy, in A's dtype; returns (factors, info)."""
    dt = _dt(A)
    F = np.array(A, dtype=dt, order="F", copy=True)
    m, n = F.shape
    info = _fn("lu_nopiv", dt)(_pt(F), _i64(m), _i64(n), _i64(_ld(F)), _i64(bsize))
    return F, info


def lu_full(A):
    """El::LU(A, P, Q) with full pivoting on a copy, in A's dtype; returns (F, p, q, info)."""
    dt = _dt(A)
    F = np.array(A, dtype=dt, order="F", copy=True)
    m, n = F.shape
    k = max(min(m, n), 1)
    p = np.zeros(k, np.int32)
    q = np.zeros(k, np.int32)
    info = _fn("lu_full", dt)(_pt(F), _i64(m), _i64(n), _i64(_ld(F)), _ip_(p), _ip_(q))
    return F, p[: min(m, n)], q[: min(m, n)], info


def lu_full_solve_after(F, p, q, B, orient=0):
    dt = _dt(F)
    F = np.asfortranarray(F, dtype=dt)
    X = np.array(B, dtype=dt, order="F", copy=True)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    pp = np.ascontiguousarray(p, np.int32)
    qq = np.ascontiguousarray(q, np.int32)
    _fn("lu_full_solve_after", dt)(orient, _pt(F), _i64(F.shape[0]), _i64(_ld(F)), _ip_(pp),
                                   _ip_(qq), _i64(pp.size), _pt(X), _i64(X.shape[1]),
                                   _i64(_ld(X)))
    return X


def lu_trace_column(A, col, bsize=128):
    dt = _dt(A)
    F = np.array(A, dtype=dt, order="F", copy=True)
    m, n = F.shape
    out = np.zeros(m - col, dtype=dt)
    _fn("lu_trace_column", dt)(_pt(F), _i64(m), _i64(n), _i64(_ld(F)), _i64(bsize), _i64(col),
                               _pt(out))
    return out


def lu_solve_after(F, dests, B, orient=0):
    """lu::SolveAfter(orient, F, P, B); dests None gives the unpivoted form."""
    dt = _dt(F)
    F = np.asfortranarray(F, dtype=dt)
    X = np.array(B, dtype=dt, order="F", copy=True)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    n = F.shape[0]
    if dests is None:
        dp, nd = ctypes.cast(None, _ip), 0
    else:
        d = np.ascontiguousarray(dests, np.int32)
        dp, nd = _ip_(d), d.size
    _fn("lu_solve_after", dt)(orient, _pt(F), _i64(n), _i64(_ld(F)), dp, _i64(nd), _pt(X),
                              _i64(X.shape[1]), _i64(_ld(X)))
    return X


def linear_solve(A, B, bsize=128):
    """El::LinearSolve(A, B) in A's dtype; returns (X, dests, info)."""
    dt = _dt(A)
    A = np.asfortranarray(A, dtype=dt)
    X = np.array(B, dtype=dt, order="F", copy=True)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    n = A.shape[0]
    d = np.zeros(max(n, 1), np.int32)
    info = _fn("linear_solve", dt)(_pt(A), _i64(n), _i64(_ld(A)), _pt(X), _i64(X.shape[1]),
                                   _i64(_ld(X)), _i64(bsize), _ip_(d))
    return X, d[:n], info


def cholesky(A, uplo=0, bsize=128):
    dt = _dt(A)
    F = np.array(A, dtype=dt, order="F", copy=True)
    n = F.shape[0]
    info = _fn("cholesky", dt)(uplo, _pt(F), _i64(n), _i64(_ld(F)), _i64(bsize))
    return F, info


def cholesky_solve_after(F, B, uplo=0, orient=0):
    dt = _dt(F)
    F = np.asfortranarray(F, dtype=dt)
    X = np.array(B, dtype=dt, order="F", copy=True)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    n = F.shape[0]
    _fn("cholesky_solve_after", dt)(uplo, orient, _pt(F), _i64(n), _i64(_ld(F)), _pt(X),
                                    _i64(X.shape[1]), _i64(_ld(X)))
    return X


def hpd_solve(A, B, uplo=0, orient=0, bsize=128):
    dt = _dt(A)
    A = np.asfortranarray(A, dtype=dt)
    X = np.array(B, dtype=dt, order="F", copy=True)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    n = A.shape[0]
    info = _fn("hpd_solve", dt)(uplo, orient, _pt(A), _i64(n), _i64(_ld(A)), _pt(X),
                                _i64(X.shape[1]), _i64(_ld(X)), _i64(bsize))
    return X, info


def perm_explicit(dests, size):
    d = np.ascontiguousarray(dests, np.int32)
    p = np.zeros(size, np.int32)
    lib().orc_perm_explicit(_ip_(d), _i64(d.size), _i64(size), _ip_(p))
    return p


def perm_parity(dests):
    d = np.ascontiguousarray(dests, np.int32)
    return lib().orc_perm_parity(_ip_(d), _i64(d.size))


def dd_residual(A, X, B):
    """R = B - A X in double-double; returns (Rhi, Rlo)."""
    A = np.asfortranarray(A, dtype=np.float64)
    X = np.asfortranarray(X.reshape(X.shape[0], -1), dtype=np.float64)
    B = np.asfortranarray(B.reshape(B.shape[0], -1), dtype=np.float64)
    m, k = A.shape
    nrhs = X.shape[1]
    Rh = np.zeros((m, nrhs), order="F")
    Rl = np.zeros((m, nrhs), order="F")
    lib().orc_dd_residual(_i64(m), _i64(k), _p(A), _i64(_ld(A)), _p(X), _i64(_ld(X)), _i64(nrhs),
                          _p(B), _i64(_ld(B)), _p(Rh), _p(Rl), _i64(max(m, 1)))
    return Rh, Rl


def backward_error(A, X, B):
    """max_c ||b_c - A x_c||inf / (||A||inf ||x_c||inf + ||b_c||inf), residual in double-double."""
    A = np.asfortranarray(A, dtype=np.float64)
    X = np.asfortranarray(X.reshape(X.shape[0], -1), dtype=np.float64)
    B = np.asfortranarray(B.reshape(B.shape[0], -1), dtype=np.float64)
    n = A.shape[0]
    eta = np.zeros(X.shape[1])
    return lib().orc_backward_error(_i64(n), _p(A), _i64(_ld(A)), _p(X), _i64(_ld(X)),
                                    _i64(X.shape[1]), _p(B), _i64(_ld(B)), _p(eta))


def lu_factor_error(A0, F, dests):
    A0 = np.asfortranarray(A0, dtype=np.float64)
    F = np.asfortranarray(F, dtype=np.float64)
    d = np.ascontiguousarray(dests, np.int32)
    return lib().orc_lu_factor_error(_i64(A0.shape[0]), _p(A0), _i64(_ld(A0)), _p(F), _i64(_ld(F)),
                                     _ip_(d), _i64(d.size))


def chol_factor_error(A0, F, uplo=0):
    A0 = np.asfortranarray(A0, dtype=np.float64)
    F = np.asfortranarray(F, dtype=np.float64)
    return lib().orc_chol_factor_error(uplo, _i64(A0.shape[0]), _p(A0), _i64(_ld(A0)), _p(F),
                                       _i64(_ld(F)))


def dd_two_prod(a, b):
    hi, lo = _dbl(), _dbl()
    lib().orc_dd_two_prod(_dbl(a), _dbl(b), ctypes.byref(hi), ctypes.byref(lo))
    return hi.value, lo.value


def dd_add(a, b):
    hi, lo = _dbl(), _dbl()
    lib().orc_dd_add(_dbl(a[0]), _dbl(a[1]), _dbl(b[0]), _dbl(b[1]), ctypes.byref(hi), ctypes.byref(lo))
    return hi.value, lo.value


def blas(name, *args):
    return getattr(lib(), name)(*args)


def gemm(ta, tb, alpha, A, B, beta, C):
    dt = _dt(C)
    C = np.array(C, dtype=dt, order="F", copy=True)
    A = np.asfortranarray(A, dtype=dt)
    B = np.asfortranarray(B, dtype=dt)
    m, n = C.shape
    k = A.shape[1] if ta == "N" else A.shape[0]
    f = lib().orc_sgemm if dt == np.float32 else lib().orc_dgemm
    f(ctypes.c_char(ta.encode()), ctypes.c_char(tb.encode()), _i64(m), _i64(n), _i64(k),
      _sc(alpha, dt), _pt(A), _i64(_ld(A)), _pt(B), _i64(_ld(B)), _sc(beta, dt), _pt(C),
      _i64(_ld(C)))
    return C


def trsm(side, uplo, trans, diag, alpha, A, B):
    dt = _dt(B)
    X = np.array(B, dtype=dt, order="F", copy=True)
    A = np.asfortranarray(A, dtype=dt)
    m, n = X.shape
    f = lib().orc_strsm if dt == np.float32 else lib().orc_dtrsm
    f(ctypes.c_char(side.encode()), ctypes.c_char(uplo.encode()), ctypes.c_char(trans.encode()),
      ctypes.c_char(diag.encode()), _i64(m), _i64(n), _sc(alpha, dt), _pt(A), _i64(_ld(A)),
      _pt(X), _i64(_ld(X)))
    return X


def syrk(uplo, trans, alpha, A, beta, C):
    dt = _dt(C)
    C = np.array(C, dtype=dt, order="F", copy=True)
    A = np.asfortranarray(A, dtype=dt)
    n = C.shape[0]
    k = A.shape[1] if trans == "N" else A.shape[0]
    f = lib().orc_ssyrk if dt == np.float32 else lib().orc_dsyrk
    f(ctypes.c_char(uplo.encode()), ctypes.c_char(trans.encode()), _i64(n), _i64(k),
      _sc(alpha, dt), _pt(A), _i64(_ld(A)), _sc(beta, dt), _pt(C), _i64(_ld(C)))
    return C


def host_dgetrf(A):
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    n = F.shape[0]
    ip = np.zeros(n, np.int32)
    info = lib().orc_host_dgetrf(_i64(n), _p(F), _i64(_ld(F)), _ip_(ip))
    return F, ip, info


# ---------------------------------------------------------------------------------------
# Near-tie classification of pivot disagreements
# ---------------------------------------------------------------------------------------
def pivot_agreement(A, dests_ref, dests_test, bsize=128, rel_tie=None):
    """Compare two pivot sequences of LU(A). Returns dict(mismatches, near_ties, first,
    gap): at the first disagreement the oracle's candidate column is recomputed and the
    relative magnitude gap between the two chosen rows reported. A disagreement is a
    near-tie when that gap is below rel_tie (1e-8 for float64, 1e-4 for float32 input). Later
    steps run on different Schur complements, so only the first disagreement is classified;
    later ones are counted."""
    if rel_tie is None:
        rel_tie = 1e-4 if _dt(A) == np.float32 else 1e-8
    dr = np.asarray(dests_ref)
    dt = np.asarray(dests_test)
    mism = np.nonzero(dr != dt)[0]
    out = {"mismatches": int(mism.size), "near_ties": 0, "first": None, "gap": None}
    if mism.size == 0:
        return out
    c = int(mism[0])
    col = lu_trace_column(A, c, bsize)
    # positions are in the swapped order at step c: candidate index = dest - c
    a = abs(col[dr[c] - c])
    b = abs(col[dt[c] - c]) if 0 <= dt[c] - c < col.size else 0.0
    gap = (a - b) / a if a > 0 else 0.0
    out.update(first=c, gap=float(gap), near_ties=int(gap <= rel_tie))
    return out


def ruiz_damp(s):
    """DampScaling (src/lapack_like/equilibrate/RuizEquil.cpp:13-21): 1 for a zero norm,
    otherwise max(norm, eps^0.33) with eps = El::limits::Epsilon<double>() = 2^-53."""
    tol = EPS ** 0.33
    return np.where(s == 0.0, 1.0, np.maximum(s, tol))


def ruiz_equil(A):
    """El::RuizEquil (src/lapack_like/equilibrate/RuizEquil.cpp:23-60) on a copy of A: four
    sweeps; each takes the column max norms (ColumnMaxNorms, src/blas_like/level1/
    ColumnNorms.cpp:89-104), damps them, multiplies them into dCol and every column by the
    reciprocal of its scale (DiagonalSolve RIGHT: A(i,j) *= 1/d(j), include/El/blas_like/
    level1/DiagonalSolve.hpp:42-56), then the same for the rows (RowMaxNorms, RowNorms.cpp:
    58-72; DiagonalSolve LEFT). Returns (A_eq, dRow, dCol) with A_eq = diag(dRow)^-1 A
    diag(dCol)^-1."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    dRow = np.ones(m)
    dCol = np.ones(n)
    for _ in range(4):
        if m > 0 and n > 0:
            cs = ruiz_damp(np.abs(A).max(axis=0))
        else:
            cs = np.ones(n)
        dCol *= cs
        A *= (1.0 / cs)[None, :]
        if m > 0 and n > 0:
            rs = ruiz_damp(np.abs(A).max(axis=1))
        else:
            rs = np.ones(m)
        dRow *= rs
        A *= (1.0 / rs)[:, None]
    return A, dRow, dCol


def symmetric_ruiz_equil(A, maxiter=3):
    """El::SymmetricRuizEquil (src/lapack_like/equilibrate/SymmetricRuizEquil.cpp:29-55) on a
    copy of the square A: maxiter sweeps (default 3, include/El/lapack_like/equilibrate.hpp:
    80-84); each takes the column max norms, damps them (DampScaling, :13-21), takes their
    square roots (SquareRootScaling, :23-27), multiplies them into d, then every entry by the
    reciprocal of its column's scale (DiagonalSolve RIGHT) and after that of its row's
    (DiagonalSolve LEFT). Returns (A_eq, d) with A_eq = diag(d)^-1 A diag(d)^-1."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    n = A.shape[0]
    d = np.ones(n)
    for _ in range(maxiter):
        s = np.sqrt(ruiz_damp(np.abs(A).max(axis=0))) if n > 0 else np.ones(0)
        d *= s
        inv = 1.0 / s
        A *= inv[None, :]
        A *= inv[:, None]
    return A, d


def geom_equil(A):
    """El::GeomEquil (src/lapack_like/equilibrate/GeomEquil.cpp:26-116) on a copy of A: ratio
    max|a| / min nonzero |a| (MinAbsNonzero, src/lapack_like/equilibrate/Util.hpp:15-31);
    up to six sweeps, each scaling every nonzero column and then every nonzero row by
    max(sqrt(mn*mx), sqrt(1/1000)*mx) (entries multiplied by 1/scale, the scale multiplied into
    dCol / dRow), stopping after sweep iter >= 3 once the new ratio is >= 9/10 of the previous;
    then every nonzero column multiplied by 1/max|a| (its max into dCol). Returns
    (A_eq, dRow, dCol) with A_eq = diag(dRow)^-1 A diag(dCol)^-1."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    dRow = np.ones(m)
    dCol = np.ones(n)
    if m == 0 or n == 0:
        return A, dRow, dCol

    def maxmin(X, axis=None):
        ab = np.abs(X)
        mx = ab.max(axis=axis)
        mn = np.where(ab > 0, ab, np.inf).min(axis=axis)
        return mx, np.minimum(mn, mx)

    mx, mn = maxmin(A)
    if mx == 0.0:
        return A, dRow, dCol
    ratio = mx / mn
    sqrt_damp = np.sqrt(1.0 / 1000.0)
    for it in range(6):
        cmx, cmn = maxmin(A, axis=0)
        nz = cmx > 0
        sc = np.maximum(np.sqrt(cmn * cmx), sqrt_damp * cmx)
        inv = np.where(nz, 1.0 / np.where(nz, sc, 1.0), 1.0)
        A *= inv[None, :]
        dCol = np.where(nz, dCol * sc, dCol)
        rmx, rmn = maxmin(A, axis=1)
        nz = rmx > 0
        sc = np.maximum(np.sqrt(rmn * rmx), sqrt_damp * rmx)
        inv = np.where(nz, 1.0 / np.where(nz, sc, 1.0), 1.0)
        A *= inv[:, None]
        dRow = np.where(nz, dRow * sc, dRow)
        mx, mn = maxmin(A)
        new_ratio = mx / mn
        if it >= 3 and new_ratio >= ratio * (9.0 / 10.0):
            break
        ratio = new_ratio
    cmx, _ = maxmin(A, axis=0)
    nz = cmx > 0
    A *= np.where(nz, 1.0 / np.where(nz, cmx, 1.0), 1.0)[None, :]
    dCol = np.where(nz, dCol * cmx, dCol)
    return A, dRow, dCol
