"""Pins the CPU oracle (the restatement of Elemental's algorithms) against known answers
that do not come from the oracle: exact rational LU, 60-digit solutions, exact Cholesky
factors, the closed form of GEPPGrowth, libstdc++'s normal_distribution restated in Python,
exact rational double-double checks, and host LAPACK where it is a valid cross-check."""
from fractions import Fraction

import numpy as np
import pytest

EPS = 2.0 ** -53


def test_gaussian_stream_matches_independent_restatement(orc, golden):
    # Elemental: mt19937 seeded (21<<16)|rank, one fresh normal_distribution per entry,
    # column-major fill.
    orc.rng_seed()
    A = orc.gaussian(8, 4)
    ref = golden["gaussian_first32"]
    assert np.array_equal(A.ravel(order="F"), ref)
    # the stream continues: the next draws differ from a reseeded stream
    B = orc.gaussian(2, 1)
    orc.rng_seed()
    C = orc.gaussian(2, 1)
    assert np.array_equal(C.ravel(), ref[:2]) and not np.array_equal(B, C)


def test_lu_pivots_and_factors_match_exact_rational(orc, golden):
    for idx in range(int(golden["lu_count"])):
        A = golden[f"lu{idx}_A"]
        F, d, info = orc.lu(A, bsize=3)  # blocked with uneven blocks
        assert info == 0
        assert np.array_equal(d, golden[f"lu{idx}_dests"]), idx
        Fe = golden[f"lu{idx}_F"]
        n = A.shape[0]
        assert np.allclose(F, Fe, rtol=0, atol=64 * n * EPS * max(1.0, np.abs(Fe).max())), idx


def test_lu_solve_matches_high_precision(orc, golden):
    for idx in range(int(golden["lu_count"])):
        A = golden[f"lu{idx}_A"]
        b = golden[f"lu{idx}_b"]
        x = golden[f"lu{idx}_x"]
        F, d, info = orc.lu(A)
        X = orc.lu_solve_after(F, d, b)
        cond = np.linalg.cond(A, np.inf)
        n = A.shape[0]
        assert np.max(np.abs(X[:, 0] - x)) <= 16 * n * cond * EPS * max(1, np.abs(x).max()), idx
        X2, d2, info2 = orc.linear_solve(A, b)
        assert info2 == 0 and np.array_equal(d2, d)
        assert np.max(np.abs(X2[:, 0] - x)) <= 16 * n * cond * EPS * max(1, np.abs(x).max()), idx


def test_lu_transpose_solve(orc):
    orc.rng_seed(11)
    A = orc.gaussian(70, 70)
    B = orc.gaussian(70, 3)
    F, d, _ = orc.lu(A, bsize=16)
    X = orc.lu_solve_after(F, d, B, orient=1)
    assert np.allclose(A.T @ X, B, atol=1e-10)


def test_gepp_growth_closed_form(orc):
    # Every candidate has |.| = 1, so ties go to the first row: identity pivots,
    # L = the -1 strictly lower part, U(i, n-1) = 2^i (src/matrices/deterministic/misc/
    # GEPPGrowth.cpp:15-30).
    for n in (1, 2, 7, 40):
        A = orc.gepp_growth(n)
        F, d, info = orc.lu(A, bsize=8)
        assert info == 0
        assert np.array_equal(d, np.arange(n, dtype=np.int32))
        L = np.tril(F, -1)
        assert np.array_equal(L, np.tril(-np.ones((n, n)), -1))
        assert np.array_equal(F[:, n - 1], 2.0 ** np.arange(n))
        U = np.triu(F)
        assert np.array_equal(U[:, : n - 1], np.eye(n)[:, : n - 1])


def test_lu_known_pivots_permutation_and_triangular(orc):
    n = 9
    p = np.array([3, 7, 0, 8, 1, 5, 2, 6, 4])
    P = np.eye(n)[p]  # row i of P A = row p[i]
    F, d, info = orc.lu(P * (2.0 ** np.arange(n))[None, :], bsize=4)
    assert info == 0
    # the row of A holding column j's nonzero is argsort(p)[j]: the explicit vector
    # (A(pe,:) = P A) lists exactly those rows, and P A = L U with L = I
    pe = orc.perm_explicit(d, n)
    assert np.array_equal(pe, np.argsort(p).astype(np.int32))
    A = P * (2.0 ** np.arange(n))[None, :]
    assert np.array_equal(A[pe], np.triu(F))
    # upper triangular with dominant diagonal: identity pivots, factors = A
    T = np.triu(np.arange(1.0, n * n + 1).reshape(n, n)) + 100 * np.eye(n)
    F, d, _ = orc.lu(np.asfortranarray(T), bsize=4)
    assert np.array_equal(d, np.arange(n, dtype=np.int32)) and np.array_equal(F, T)


def test_lu_singular_reports_column(orc):
    A = np.asfortranarray(np.arange(1.0, 26).reshape(5, 5))
    A[:, 2] = 0.0
    F, d, info = orc.lu(A, bsize=2)
    assert info == 3  # throws at global column 2
    B = np.asfortranarray(np.ones((4, 4)))
    assert orc.lu(B)[2] == 2


def test_cholesky_exact(orc, golden):
    for idx in range(int(golden["chol_count"])):
        G = golden[f"chol{idx}_G"]
        A = golden[f"chol{idx}_A"]
        for bs in (1, 3, 128):
            F, info = orc.cholesky(A, uplo=0, bsize=bs)
            assert info == 0 and np.array_equal(np.tril(F), G), (idx, bs)
            F, info = orc.cholesky(A, uplo=1, bsize=bs)
            assert info == 0 and np.array_equal(np.triu(F), G.T), (idx, bs)


def test_cholesky_non_hpd(orc):
    A = -np.eye(4, order="F")
    assert orc.cholesky(A)[1] == 1
    A = np.asfortranarray(np.diag([4.0, 1.0, -2.0, 3.0]))
    assert orc.cholesky(A, uplo=0)[1] == 3
    assert orc.cholesky(A, uplo=1)[1] == 3


def test_hpd_solve(orc):
    orc.rng_seed(3)
    G = orc.gaussian(60, 60)
    A = np.asfortranarray(G @ G.T / 60 + np.eye(60))
    B = orc.gaussian(60, 4)
    for uplo in (0, 1):
        X, info = orc.hpd_solve(A, B, uplo=uplo, bsize=16)
        assert info == 0
        assert orc.backward_error(A, X, B) <= 60 * EPS


def test_dd_primitives_exact(orc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = rng.standard_normal(2) * 10.0 ** rng.integers(-20, 20, 2)
        hi, lo = orc.dd_two_prod(a, b)
        assert Fraction(hi) + Fraction(lo) == Fraction(a) * Fraction(b)
        x = (a, 0.0)
        y = (b, b * 1e-17)
        s_hi, s_lo = orc.dd_add(x, y)
        exact = Fraction(a) + Fraction(b) + Fraction(b * 1e-17)
        got = Fraction(s_hi) + Fraction(s_lo)
        assert abs(got - exact) <= abs(exact) * Fraction(2.0 ** -100) + Fraction(2.0 ** -1070)


def test_dd_residual_against_rationals(orc):
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.standard_normal((7, 7)))
    X = np.asfortranarray(rng.standard_normal((7, 2)))
    B = np.asfortranarray(A @ X)  # residual is pure rounding error of the product
    Rh, Rl = orc.dd_residual(A, X, B)
    for c in range(2):
        for i in range(7):
            exact = Fraction(B[i, c]) - sum(Fraction(A[i, j]) * Fraction(X[j, c]) for j in range(7))
            got = Fraction(Rh[i, c]) + Fraction(Rl[i, c])
            assert abs(got - exact) <= Fraction(2.0 ** -95) * (1 + abs(exact))


def test_backward_error_definition(orc):
    A = np.asfortranarray(np.diag([2.0, 4.0]))
    X = np.asfortranarray([[1.0], [1.0]])
    B = np.asfortranarray([[2.0], [4.0 + 2.0 ** -40]])
    eta = orc.backward_error(A, X, B)
    assert eta == pytest.approx(2.0 ** -40 / (4.0 * 1.0 + (4.0 + 2.0 ** -40)), rel=1e-12)


def test_trsm_variants_against_numpy(orc):
    rng = np.random.default_rng(2)
    n, k = 13, 5
    T = np.asfortranarray(rng.standard_normal((n, n)) + 5 * np.eye(n))
    for side in "LR":
        for uplo in "LU":
            for trans in "NT":
                for diag in "NU":
                    B = np.asfortranarray(rng.standard_normal((n, k) if side == "L" else (k, n)))
                    X = orc.trsm(side, uplo, trans, diag, 1.0, T, B)
                    M = np.tril(T) if uplo == "L" else np.triu(T)
                    if diag == "U":
                        M = M - np.diag(np.diag(M)) + np.eye(n)
                    if trans == "T":
                        M = M.T
                    R = (M @ X) if side == "L" else (X @ M)
                    assert np.allclose(R, B, atol=1e-11), (side, uplo, trans, diag)


def test_gemm_syrk_against_numpy(orc):
    rng = np.random.default_rng(4)
    A = np.asfortranarray(rng.standard_normal((9, 6)))
    B = np.asfortranarray(rng.standard_normal((6, 7)))
    C = np.asfortranarray(rng.standard_normal((9, 7)))
    assert np.allclose(orc.gemm("N", "N", -1.0, A, B, 1.0, C), C - A @ B)
    assert np.allclose(orc.gemm("T", "N", 2.0, np.asfortranarray(A.T), B, 0.0, C), 2 * A @ B)
    assert np.allclose(orc.gemm("N", "T", 1.0, A, np.asfortranarray(B.T), 1.0, C), C + A @ B)
    S = np.asfortranarray(rng.standard_normal((9, 9)))
    L = orc.syrk("L", "N", -1.0, A, 1.0, S)
    assert np.allclose(np.tril(L), np.tril(S - A @ A.T)) and np.array_equal(np.triu(L, 1), np.triu(S, 1))
    U = orc.syrk("U", "T", -1.0, np.asfortranarray(A.T), 1.0, S)
    assert np.allclose(np.triu(U), np.triu(S - A @ A.T)) and np.array_equal(np.tril(U, -1), np.tril(S, -1))


def test_lu_config1_against_host_lapack(orc):
    """Config 1 (n=1024 Gaussian from Elemental's stream): oracle pivots vs host dgetrf;
    the blocking differs, so only near-ties may disagree. Backward error <= n u."""
    n = 1024
    orc.rng_seed()
    A = orc.gaussian(n, n)
    b = orc.gaussian(n, 1)
    F, d, info = orc.lu(A)
    assert info == 0
    import scipy.linalg as sl
    _, piv = sl.lu_factor(A)
    agree = orc.pivot_agreement(A, d, piv.astype(np.int32))
    assert agree["mismatches"] == 0 or agree["near_ties"] == 1
    x = orc.lu_solve_after(F, d, b)
    assert orc.backward_error(A, x, b) <= n * EPS
    assert orc.lu_factor_error(A, F, d) <= n * EPS * 8
    # the reference's own LU test criterion (tests/lapack_like/LU.cpp:43-53)
    Y = orc.lu_solve_after(F, d, b)
    rel = np.abs(A @ Y - b).max() / (EPS * n * max(np.abs(A).sum(0).max(), np.abs(Y).sum(0).max()))
    assert rel <= 100


def test_perm_parity_and_explicit_closed_forms(orc):
    # Permutation::Parity flips once per non-trivial swap (Permutation.cpp:213-214)
    assert orc.perm_parity(np.arange(5, dtype=np.int32)) == 0
    assert orc.perm_parity(np.array([1, 1, 2], np.int32)) == 1
    assert orc.perm_parity(np.array([2, 2, 2], np.int32)) == 0  # (0 2)(1 2) -> 3-cycle, even
    d = np.array([2, 2, 2], np.int32)
    # swap 0<->2 then 1<->2 on rows [0,1,2]: [2,1,0] -> [2,0,1]
    assert np.array_equal(orc.perm_explicit(d, 3), np.array([2, 0, 1], np.int32))


def test_factor_error_measures_planted_defects(orc):
    """The factor-error checks used by the GPU tests must see a planted error of known size."""
    rng = np.random.default_rng(9)
    n = 12
    A = np.asfortranarray(rng.standard_normal((n, n)))
    F, d, _ = orc.lu(A, bsize=4)
    e0 = orc.lu_factor_error(A, F, d)
    assert e0 <= 8 * n * EPS
    G = F.copy(order="F")
    G[5, 7] += 1e-6  # one U entry off by 1e-6: ||PA - LU|| grows by ~1e-6 relative
    e1 = orc.lu_factor_error(A, G, d)
    assert 1e-8 < e1 < 1e-5
    d2 = d.copy()
    d2[0] = (d2[0] + 1) % n  # wrong first pivot: a gross error
    assert orc.lu_factor_error(A, F, d2) > 1e-3
    H = np.asfortranarray(A @ A.T / n + np.eye(n))
    C, info = orc.cholesky(H, uplo=0, bsize=4)
    assert info == 0 and orc.chol_factor_error(H, C, 0) <= 8 * n * EPS
    C[6, 2] *= 1 + 1e-6
    assert orc.chol_factor_error(H, C, 0) > 1e-9


def test_trace_column_is_the_pivot_candidate_column(orc):
    # lu_trace_column(A, c) returns the candidates of step c (rows c.., after the first c
    # steps): column 0 is A(:,0); every column's first largest entry sits at dests[c] - c and
    # equals U(c,c) in magnitude.
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.standard_normal((40, 40)))
    F, d, _ = orc.lu(A, bsize=8)
    assert np.array_equal(orc.lu_trace_column(A, 0, 8), A[:, 0])
    for c in (1, 7, 8, 9, 23, 39):
        col = orc.lu_trace_column(A, c, 8)
        k = int(np.argmax(np.abs(col)))
        assert k == d[c] - c and abs(col[k]) == abs(F[c, c]), c


def test_host_dgetrf_matches_scipy(orc):
    import scipy.linalg as sl
    rng = np.random.default_rng(8)
    A = np.asfortranarray(rng.standard_normal((64, 64)))
    for name in ("openblas", "mkl"):
        try:
            orc.set_backend(name)
            break
        except RuntimeError:
            continue
    else:
        pytest.skip("no host LAPACK")
    try:
        F, ip, info = orc.host_dgetrf(A)
    finally:
        orc.set_backend("naive")
    lu_, piv = sl.lu_factor(A)
    assert info == 0 and np.array_equal(ip, piv) and np.allclose(F, lu_, atol=1e-12)


# ------------------------------------------------------------------------------------------
# float instantiations (the orc_s* entry points) and the unpivoted LU
# ------------------------------------------------------------------------------------------
EPS32 = 2.0 ** -24


def test_float_gepp_growth_closed_form(orc):
    # the closed form holds in float too (all values are small integers and powers of two)
    for n in (1, 2, 7, 40):
        A = orc.gepp_growth(n).astype(np.float32)
        F, d, info = orc.lu(A, bsize=8)
        assert F.dtype == np.float32 and info == 0
        assert np.array_equal(d, np.arange(n, dtype=np.int32))
        assert np.array_equal(np.tril(F, -1), np.tril(-np.ones((n, n), np.float32), -1))
        assert np.array_equal(F[:, n - 1], (2.0 ** np.arange(n)).astype(np.float32))


def test_float_cholesky_exact(orc, golden):
    # integer G with small entries: G G^T and every step of the factorization are exact in float
    for idx in range(int(golden["chol_count"])):
        G = golden[f"chol{idx}_G"]
        A = golden[f"chol{idx}_A"]
        if np.abs(A).max() >= 2 ** 24:
            continue
        for bs in (1, 3, 128):
            F, info = orc.cholesky(A.astype(np.float32), uplo=0, bsize=bs)
            assert info == 0 and np.array_equal(np.tril(F), G.astype(np.float32)), (idx, bs)
            F, info = orc.cholesky(A.astype(np.float32), uplo=1, bsize=bs)
            assert info == 0 and np.array_equal(np.triu(F), G.T.astype(np.float32)), (idx, bs)


def test_float_lu_matches_exact_rational(orc, golden):
    """The golden matrices are small integers (exact in float): the float oracle must find the
    exact pivots and factors within a float rounding budget."""
    for idx in range(int(golden["lu_count"])):
        A = golden[f"lu{idx}_A"]
        assert np.array_equal(A.astype(np.float32).astype(np.float64), A)
        F, d, info = orc.lu(A.astype(np.float32), bsize=3)
        assert info == 0
        assert np.array_equal(d, golden[f"lu{idx}_dests"]), idx
        Fe = golden[f"lu{idx}_F"]
        n = A.shape[0]
        assert np.allclose(F, Fe, rtol=0, atol=64 * n * EPS32 * max(1.0, np.abs(Fe).max())), idx


def test_float_solves_against_float64(orc):
    orc.rng_seed(21)
    A = orc.gaussian(90, 90)
    B = orc.gaussian(90, 3)
    H = np.asfortranarray(A @ A.T / 90 + np.eye(90))
    X64, _, _ = orc.linear_solve(A, B, bsize=16)
    X32, d32, info = orc.linear_solve(A.astype(np.float32), B.astype(np.float32), bsize=16)
    assert info == 0 and X32.dtype == np.float32
    cond = np.linalg.cond(A)
    assert np.abs(X32 - X64).max() <= 90 * cond * EPS32 * np.abs(X64).max()
    Y64, _ = orc.hpd_solve(H, B, bsize=16)
    Y32, info = orc.hpd_solve(H.astype(np.float32), B.astype(np.float32), bsize=16)
    assert info == 0 and np.abs(Y32 - Y64).max() <= 90 * np.linalg.cond(H) * EPS32 * np.abs(Y64).max()
    for side in "LR":
        T = np.asfortranarray(np.triu(A) + 10 * np.eye(90)).astype(np.float32)
        C = np.asfortranarray(B if side == "L" else B.T).astype(np.float32)
        Xs = orc.trsm(side, "U", "N", "N", 1.0, T, C)
        R = (np.triu(T).astype(np.float64) @ Xs) if side == "L" else (Xs @ np.triu(T).astype(np.float64))
        assert np.allclose(R, C, atol=1e-4)


def _exact_lu_pair(n, seed):
    """L unit lower and U upper with small integers and power-of-two diagonal: A = L U is exact,
    and so is every multiplier and update of the unpivoted elimination."""
    rng = np.random.default_rng(seed)
    L = np.tril(rng.integers(-3, 4, (n, n)).astype(np.float64), -1) + np.eye(n)
    U = np.triu(rng.integers(-3, 4, (n, n)).astype(np.float64), 1) + np.diag(2.0 ** rng.integers(0, 3, n))
    return L, U, np.asfortranarray(L @ U)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lu_nopiv_exact_known_factors(orc, dtype):
    for n in (1, 5, 12):
        L, U, A = _exact_lu_pair(n, n)
        if np.abs(A).max() >= 2 ** 20:
            continue
        for bs in (1, 3, 128):
            F, info = orc.lu_nopiv(A.astype(dtype), bsize=bs)
            assert info == 0 and F.dtype == dtype
            assert np.array_equal(np.tril(F, -1), np.tril(L, -1).astype(dtype)), (n, bs)
            assert np.array_equal(np.triu(F), U.astype(dtype)), (n, bs)


def test_lu_nopiv_equals_pivoted_on_column_dominant(orc):
    """A strictly column diagonally dominant matrix stays so under elimination, so partial
    pivoting picks the diagonal every step; with the same restated BLAS the unpivoted form
    (LU.cpp:21-44) then gives the pivoted one's factors bit for bit."""
    rng = np.random.default_rng(13)
    for n, bs in [(40, 8), (130, 128), (75, 16)]:
        A = rng.standard_normal((n, n))
        A += np.diag(np.abs(A).sum(0) + 1.0)
        A = np.asfortranarray(A)
        F, d, info = orc.lu(A, bsize=bs)
        assert info == 0 and np.array_equal(d, np.arange(n, dtype=np.int32))
        G, info2 = orc.lu_nopiv(A, bsize=bs)
        assert info2 == 0 and np.array_equal(F, G)
        B = np.asfortranarray(rng.standard_normal((n, 2)))
        assert np.array_equal(orc.lu_solve_after(G, None, B), orc.lu_solve_after(F, d, B))


def test_lu_nopiv_singular_reports_column(orc):
    L, U, A = _exact_lu_pair(9, 3)
    U[4, 4] = 0.0  # exact zero pivot at column 4 once columns 0..3 are eliminated
    A = np.asfortranarray(L @ U)
    for bs in (1, 3, 128):
        F, info = orc.lu_nopiv(A, bsize=bs)
        assert info == 5, bs
        # lu::Unb throws inside the diagonal block, before the block's Trsm of A21: the
        # columns of earlier blocks are finished (exact multipliers), and inside the failing
        # block only its own rows carry the multipliers of columns before the zero pivot
        k0 = (4 // bs) * bs
        assert np.array_equal(np.tril(F, -1)[:, :k0], np.tril(L, -1)[:, :k0]), bs
        r1 = min(k0 + bs, 9)
        assert np.array_equal(np.tril(F, -1)[k0:r1, k0:4], np.tril(L, -1)[k0:r1, k0:4]), bs
        if r1 < 9 and k0 < 4:
            assert not np.array_equal(F[r1:, k0:4], L[r1:, k0:4]), bs  # A21 not yet solved


# ------------------------------------------------------------------------------------------
# LU with full pivoting (lu::Full)
# ------------------------------------------------------------------------------------------
def _exact_full_pivot_lu(A):
    """Complete pivoting in exact rational arithmetic with the reference's choice rule (the
    first entry of largest magnitude in column-major order)."""
    m, n = A.shape
    M = [[Fraction(A[i, j]) for j in range(n)] for i in range(m)]
    p, q = [], []
    for k in range(min(m, n)):
        best, bi, bj = Fraction(0), 0, 0
        for j in range(k, n):
            for i in range(k, m):
                if abs(M[i][j]) > best:
                    best, bi, bj = abs(M[i][j]), i - k, j - k
        ip, jp = k + bi, k + bj
        p.append(ip)
        q.append(jp)
        M[k], M[ip] = M[ip], M[k]
        for r in range(m):
            M[r][k], M[r][jp] = M[r][jp], M[r][k]
        if M[k][k] == 0:
            return M, p, q, k + 1
        for i in range(k + 1, m):
            M[i][k] = M[i][k] / M[k][k]
            for j in range(k + 1, n):
                M[i][j] -= M[i][k] * M[k][j]
    return M, p, q, 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lu_full_matches_exact_rational(orc, dtype):
    rng = np.random.default_rng(70)
    for (m, n) in [(6, 6), (7, 5), (5, 8)]:
        A = rng.integers(-9, 10, (m, n)).astype(np.float64)
        Me, pe, qe, infoe = _exact_full_pivot_lu(A)
        F, p, q, info = orc.lu_full(A.astype(dtype))
        assert info == infoe
        assert list(p) == pe[: len(p)] and list(q) == qe[: len(q)], (m, n)
        Fe = np.array([[float(x) for x in row] for row in Me])
        eps = 2.0 ** -53 if dtype == np.float64 else 2.0 ** -24
        assert np.allclose(F, Fe, rtol=0, atol=64 * max(m, n) * eps * np.abs(Fe).max())


def test_lu_full_closed_forms(orc):
    # diagonal: the pivots take the diagonal in decreasing magnitude, L = I, U = sorted diagonal
    d = np.array([3.0, -7.0, 0.5, 9.0, -2.0])
    F, p, q, info = orc.lu_full(np.asfortranarray(np.diag(d)))
    assert info == 0 and np.array_equal(np.abs(np.diag(F)), np.sort(np.abs(d))[::-1])
    assert np.array_equal(np.tril(F, -1), np.zeros((5, 5))) and np.array_equal(p, q)
    # rank one: the Schur complement after one step is exactly zero -> singular at step 2
    u = np.array([1.0, -2.0, 4.0])
    v = np.array([2.0, 1.0, -8.0, 0.5])
    F, p, q, info = orc.lu_full(np.asfortranarray(np.outer(u, v)))
    assert info == 2 and p[0] == 2 and q[0] == 2  # |4 * -8| = 32 is the largest entry
    # ties: all ones -> the first entry in column-major order, then singular
    F, p, q, info = orc.lu_full(np.ones((4, 4), order="F"))
    assert info == 2 and p[0] == 0 and q[0] == 0
    # the solve: P A Q = L U
    rng = np.random.default_rng(71)
    A = np.asfortranarray(rng.standard_normal((30, 30)))
    B = np.asfortranarray(rng.standard_normal((30, 3)))
    F, p, q, info = orc.lu_full(A)
    for orient in (0, 1):
        X = orc.lu_full_solve_after(F, p, q, B, orient=orient)
        assert np.allclose((A if orient == 0 else A.T) @ X, B, atol=1e-11)
