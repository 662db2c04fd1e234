"""The rest of the drop-in boundary on the GPU, against the oracle on the same inputs:

* the float instantiations (elx_s*: LU, SolveAfter, LinearSolve, Cholesky, HPDSolve, Gemm,
  Trsm, Syrk, PermuteRows), which the reference provides through the same PROTO macros as
  the double ones (LU.cpp:233-313, Cholesky.cpp:206-266, Linear.cpp:320-343, HPD.cpp:109-133);
* LU without pivoting and its SolveAfter (LU.cpp:21-44, LU/SolveAfter.hpp:15-38);
* the state a failed factorization leaves behind: the reference throws at the first exactly
  zero pivot (Panel.hpp:52-53, Local.hpp:52-53) or non-positive Cholesky pivot
  (LowerVariant3.hpp:26-29) with A partially overwritten, and the device sequence stops at the
  same point.

Bars: pivots identical to the oracle (or one counted near-tie), factor error <= 8 n u of the
field, single kernels whose FMA order is the restated BLAS's bit for bit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53
EPS32 = 2.0 ** -24


def say(*a):
    print("[case]", *a, flush=True)


def lu_factor_error32(A, F, d):
    """||P A - L U||_inf / ||A||_inf of float factors, in float64 (exact products of floats)."""
    n = A.shape[0]
    p = np.arange(n)
    for j, t in enumerate(d):
        p[[j, t]] = p[[t, j]]
    L = np.tril(F.astype(np.float64), -1) + np.eye(n)
    U = np.triu(F.astype(np.float64))
    R = A.astype(np.float64)[p] - L @ U
    return np.abs(R).sum(1).max() / np.abs(A.astype(np.float64)).sum(1).max()


# ---------------------------------------------------------------------------------- float
def test_float_lu_vs_oracle(dev, orc):
    orc.rng_seed(31)
    for (m, n, nb) in [(1, 1, 128), (64, 64, 16), (128, 128, 128), (300, 300, 96),
                       (700, 700, 128), (400, 250, 128)]:
        say("float lu", m, n, nb)
        A = orc.gaussian(m, n).astype(np.float32)
        F_ref, d_ref, info_ref = orc.lu(A, bsize=nb)
        dA = dev.to_device(A)
        assert dA.dtype == torch.float32
        d, info = dev.getrf(dA, nb=nb)
        F = dev.to_host(dA)
        d = d.cpu().numpy()
        assert int(info.item()) == info_ref == 0
        agree = orc.pivot_agreement(A, d_ref, d, bsize=nb)
        assert agree["mismatches"] == 0 or agree["near_ties"] == 1, agree
        if m == n:
            assert lu_factor_error32(A, F, d) <= 8 * n * EPS32
        if n <= nb and agree["mismatches"] == 0:
            assert np.array_equal(F, F_ref), "single float panel must match lu::Panel bit for bit"


def test_float_solves_vs_oracle(dev, orc):
    orc.rng_seed(32)
    for n in (50, 300):
        A = orc.gaussian(n, n).astype(np.float32)
        B = orc.gaussian(n, 3).astype(np.float32)
        say("float linear solve", n)
        X_ref, d_ref, _ = orc.linear_solve(A, B)
        dA, dB = dev.to_device(A), dev.to_device(B)
        d, info = dev.gesv(dA, dB, want_dests=True)
        X = dev.to_host(dB)
        assert int(info.item()) == 0 and X.dtype == np.float32
        cond = np.linalg.cond(A.astype(np.float64), np.inf)
        assert np.abs(X - X_ref).max() <= 16 * n * cond * EPS32 * np.abs(X_ref).max()
        say("float lu solve after", n)
        F, d2, _ = orc.lu(A)
        for orient in (0, 1):
            dF = dev.to_device(F)
            dd = torch.from_numpy(d2).cuda()
            dB = dev.to_device(B)
            dev.getrs(dF, dd, dB, orient=orient)
            ref = orc.lu_solve_after(F, d2, B, orient=orient)
            assert np.abs(dev.to_host(dB) - ref).max() <= 16 * n * cond * EPS32 * np.abs(ref).max()


def test_float_cholesky_and_hpd_solve(dev, orc, golden):
    for idx in range(int(golden["chol_count"])):
        G = golden[f"chol{idx}_G"]
        A = golden[f"chol{idx}_A"]
        if np.abs(A).max() >= 2 ** 24:
            continue
        for uplo in (0, 1):
            dA = dev.to_device(A.astype(np.float32))
            info = dev.potrf(dA, uplo=uplo, nb=4)
            F = dev.to_host(dA)
            assert int(info.item()) == 0
            tri = np.tril(F) if uplo == 0 else np.triu(F).T
            assert np.array_equal(tri, G.astype(np.float32)), "exact integer factor in float"
    orc.rng_seed(33)
    for n in (100, 400):
        Gm = orc.gaussian(n, n)
        H = np.asfortranarray(Gm @ Gm.T / n + np.eye(n)).astype(np.float32)
        B = orc.gaussian(n, 5).astype(np.float32)
        for uplo in (0, 1):
            say("float hpd", n, uplo)
            F_ref, _ = orc.cholesky(H, uplo=uplo)
            dH = dev.to_device(H)
            info = dev.potrf(dH, uplo=uplo)
            F = dev.to_host(dH)
            tri = np.tril if uplo == 0 else np.triu
            assert int(info.item()) == 0
            assert np.abs(tri(F) - tri(F_ref)).max() <= 64 * n * EPS32 * np.abs(F_ref).max()
            if n <= 128:
                assert np.array_equal(tri(F), tri(F_ref))
            X_ref, _ = orc.hpd_solve(H, B, uplo=uplo)
            dH, dB = dev.to_device(H), dev.to_device(B)
            assert int(dev.posv(dH, dB, uplo=uplo).item()) == 0
            X = dev.to_host(dB)
            assert np.abs(X - X_ref).max() <= 1e-3 * np.abs(X_ref).max()


def test_float_level3_vs_oracle(dev, orc):
    rng = np.random.default_rng(34)
    for (m, n, k) in [(37, 53, 19), (300, 257, 130)]:
        for ta in (0, 1):
            for tb in (0, 1):
                say("sgemm", m, n, k, ta, tb)
                A = rng.standard_normal((m, k) if ta == 0 else (k, m)).astype(np.float32)
                B = rng.standard_normal((k, n) if tb == 0 else (n, k)).astype(np.float32)
                C = rng.standard_normal((m, n)).astype(np.float32)
                dC = dev.to_device(C)
                dev.gemm(ta, tb, -1.0, dev.to_device(A), dev.to_device(B), 0.5, dC)
                ref = orc.gemm("NT"[ta], "NT"[tb], -1.0, A, B, 0.5, C)
                assert np.abs(dev.to_host(dC) - ref).max() <= 8 * k * EPS32 * (np.abs(ref).max() + k)
    for order in (17, 128, 300):
        for side in (0, 1):
            for uplo in (0, 1):
                T = (rng.standard_normal((order, order)) + 4 * np.eye(order)).astype(np.float32)
                B = rng.standard_normal((order, 9) if side == 0 else (9, order)).astype(np.float32)
                dB = dev.to_device(B)
                dev.trsm(side, uplo, 0, 0, 1.0, dev.to_device(T), dB)
                ref = orc.trsm("LR"[side], "LU"[uplo], "N", "N", 1.0, T, B)
                got = dev.to_host(dB)
                if order <= 128:
                    assert np.array_equal(got, ref), "small float triangular solve bit for bit"
                else:
                    assert np.abs(got - ref).max() <= 1e-3 * np.abs(ref).max()
    A = rng.standard_normal((200, 64)).astype(np.float32)
    S = rng.standard_normal((200, 200)).astype(np.float32)
    for uplo in (0, 1):
        dS = dev.to_device(S)
        dev.syrk(uplo, 0, -1.0, dev.to_device(A), 1.0, dS)
        ref = orc.syrk("LU"[uplo], "N", -1.0, A, 1.0, S)
        tri = np.tril if uplo == 0 else np.triu
        got = dev.to_host(dS)
        assert np.abs(tri(got) - tri(ref)).max() <= 1e-4 * np.abs(ref).max()
        other = np.triu if uplo == 0 else np.tril
        assert np.array_equal(other(got, 1 if uplo == 0 else -1), other(S, 1 if uplo == 0 else -1))
    n = 333
    dests = np.array([rng.integers(j, n) for j in range(n)], np.int32)
    B = rng.standard_normal((n, 3)).astype(np.float32)
    p = orc.perm_explicit(dests, n)
    dB = dev.to_device(B)
    dev.permute_rows(torch.from_numpy(dests).cuda(), dB)
    assert np.array_equal(dev.to_host(dB), B[p])


# ------------------------------------------------------------------------- no pivoting
def dominant(rng, n, dtype=np.float64):
    A = rng.standard_normal((n, n))
    A += np.diag(np.abs(A).sum(0) + 1.0)
    return np.asfortranarray(A).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lu_nopiv_vs_oracle(dev, orc, dtype):
    rng = np.random.default_rng(35)
    eps = EPS if dtype == np.float64 else EPS32
    for (n, nb) in [(1, 128), (9, 4), (100, 128), (128, 128), (300, 96), (700, 128)]:
        say("lu nopiv", n, nb, dtype.__name__)
        A = dominant(rng, n, dtype)
        F_ref, info_ref = orc.lu_nopiv(A, bsize=nb)
        dA = dev.to_device(A)
        info = dev.getrf_nopiv(dA, nb=nb)
        F = dev.to_host(dA)
        assert int(info.item()) == info_ref == 0
        if n <= nb:
            assert np.array_equal(F, F_ref), "one lu::Unb block matches the oracle bit for bit"
        L = np.tril(F.astype(np.float64), -1) + np.eye(n)
        U = np.triu(F.astype(np.float64))
        R = A.astype(np.float64) - L @ U
        assert np.abs(R).sum(1).max() <= 8 * n * eps * np.abs(A.astype(np.float64)).sum(1).max()
        assert np.abs(F - F_ref).max() <= 1e3 * n * eps * np.abs(F_ref).max()
        B = rng.standard_normal((n, 4)).astype(dtype)
        for orient in (0, 1):
            dB = dev.to_device(B)
            dev.getrs(dev.to_device(F_ref), None, dB, orient=orient)
            ref = orc.lu_solve_after(F_ref, None, B, orient=orient)
            assert np.abs(dev.to_host(dB) - ref).max() <= 1e3 * n * eps * np.abs(ref).max()


def test_lu_nopiv_exact_factors(dev):
    rng = np.random.default_rng(36)
    n = 200
    L = np.tril(rng.integers(-2, 3, (n, n)).astype(np.float64), -1) / 64 + np.eye(n)
    U = np.triu(rng.integers(-2, 3, (n, n)).astype(np.float64), 1) / 64 + 4 * np.eye(n)
    A = np.asfortranarray(L @ U)
    for nb in (16, 128):
        dA = dev.to_device(A)
        info = dev.getrf_nopiv(dA, nb=nb)
        F = dev.to_host(dA)
        assert int(info.item()) == 0
        assert np.abs(np.tril(F, -1) - np.tril(L, -1)).max() <= 1e-12
        assert np.abs(np.triu(F) - U).max() <= 1e-12


# ------------------------------------------------- the state a failed factorization leaves
def test_lu_singular_state_matches_reference(dev, orc):
    """A zero column at 200 of 300 (nb 128): an exactly zero pivot at global column 200. The
    reference throws inside the second panel: block 0 finished, the second panel's rows
    swapped and updated through column 199, the columns outside that panel not permuted for
    it, the trailing block not updated."""
    orc.rng_seed(37)
    n = 300
    A = orc.gaussian(n, n)
    A[:, 200] = 0.0
    F_ref, d_ref, info_ref = orc.lu(A)
    assert info_ref == 201
    dA = dev.to_device(A)
    d, info = dev.getrf(dA)
    F = dev.to_host(dA)
    assert int(info.item()) == 201
    d = d.cpu().numpy()
    agree = orc.pivot_agreement(A, d_ref[:200], d[:200])
    assert agree["mismatches"] == 0, agree
    assert d[200] == 200
    scale = np.abs(F_ref).max()
    assert np.abs(F - F_ref).max() <= 1e-11 * scale
    # the trailing block (columns 256..) holds A after block 0 only: far from a finished LU
    assert np.abs(F[:, 256:] - F_ref[:, 256:]).max() <= 1e-11 * scale


def test_linear_solve_singular_state(dev, orc):
    orc.rng_seed(38)
    n = 260
    A = orc.gaussian(n, n)
    A[:, 140] = 0.0
    B = orc.gaussian(n, 2)
    X_ref, d_ref, info_ref = orc.linear_solve(A, B)
    assert info_ref == 141
    dA, dB = dev.to_device(A), dev.to_device(B)
    info = dev.gesv(dA, dB)
    assert int(info.item()) == 141
    # B as RowEchelon left it when it threw: permuted and updated by block 0 only
    assert np.abs(dev.to_host(dB) - X_ref).max() <= 1e-11 * np.abs(X_ref).max()


def test_cholesky_non_hpd_state(dev, orc):
    orc.rng_seed(39)
    n = 300
    G = orc.gaussian(n, n)
    A = np.asfortranarray(G @ G.T / n + np.eye(n))
    A[170, 170] = -1.0  # fails at (or before) column 170, inside the second block
    for uplo in (0, 1):
        F_ref, info_ref = orc.cholesky(A, uplo=uplo)
        assert 128 < info_ref <= 171
        dA = dev.to_device(A)
        info = dev.potrf(dA, uplo=uplo)
        assert int(info.item()) == info_ref
        F = dev.to_host(dA)
        assert np.abs(F - F_ref).max() <= 1e-11 * np.abs(F_ref).max()


def test_lu_nopiv_singular_state(dev, orc):
    rng = np.random.default_rng(40)
    n = 260
    A = dominant(rng, n)
    A[:, 150] = 0.0
    A[150, 150] = 0.0
    F_ref, info_ref = orc.lu_nopiv(A)
    assert info_ref > 0
    dA = dev.to_device(A)
    info = dev.getrf_nopiv(dA)
    assert int(info.item()) == info_ref
    assert np.abs(dev.to_host(dA) - F_ref).max() <= 1e-11 * np.abs(F_ref).max()


def test_blocksize_above_128(dev, orc):
    """Blocksize() is not limited to 128 in the reference: a wider block gives the same
    pivots (exact arithmetic); the factors agree to rounding."""
    orc.rng_seed(41)
    A = orc.gaussian(600, 600)
    F_ref, d_ref, _ = orc.lu(A, bsize=256)
    dA = dev.to_device(A)
    d, info = dev.getrf(dA, nb=256)
    assert int(info.item()) == 0
    agree = orc.pivot_agreement(A, d_ref, d.cpu().numpy(), bsize=256)
    assert agree["mismatches"] == 0 or agree["near_ties"] == 1, agree
    assert orc.lu_factor_error(A, dev.to_host(dA), d.cpu().numpy()) <= 8 * 600 * EPS


# -------------------------------------------------------------------------- full pivoting
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lu_full_pivoting_vs_oracle(dev, orc, dtype):
    """lu::Full (LU/Full.hpp:15-58): every step's pivot is the first largest entry of the
    trailing block in column-major order; pivots identical to the oracle's, and the factors
    bit for bit (each entry receives the same FMA sequence as the restated dger)."""
    rng = np.random.default_rng(72)
    for (m, n) in [(1, 1), (7, 7), (60, 45), (45, 60), (300, 300), (700, 700)]:
        say("lu full", m, n, dtype.__name__)
        A = np.asfortranarray(rng.standard_normal((m, n))).astype(dtype)
        F_ref, p_ref, q_ref, info_ref = orc.lu_full(A)
        dA = dev.to_device(A)
        p, q, info = dev.getrf_full(dA)
        F = dev.to_host(dA)
        assert int(info.item()) == info_ref == 0
        assert np.array_equal(p.cpu().numpy(), p_ref) and np.array_equal(q.cpu().numpy(), q_ref)
        assert np.array_equal(F, F_ref)
        if m == n:
            B = rng.standard_normal((n, 3)).astype(dtype)
            for orient in (0, 1):
                dB = dev.to_device(B)
                dev.getrs_full(dev.to_device(F_ref), torch.from_numpy(p_ref).cuda(),
                               torch.from_numpy(q_ref).cuda(), dB, orient=orient)
                ref = orc.lu_full_solve_after(F_ref, p_ref, q_ref, B, orient=orient)
                tol = (1e-11 if dtype == np.float64 else 1e-3) * np.abs(ref).max()
                assert np.abs(dev.to_host(dB) - ref).max() <= tol


def test_lu_full_pivoting_singular_and_ties(dev, orc):
    u = np.array([1.0, -2.0, 4.0, 0.5])
    v = np.array([2.0, 1.0, -8.0, 0.5, 3.0])
    for A in (np.asfortranarray(np.outer(u, v)), np.ones((6, 6), order="F"),
              np.zeros((5, 5), order="F")):
        F_ref, p_ref, q_ref, info_ref = orc.lu_full(A)
        dA = dev.to_device(A)
        p, q, info = dev.getrf_full(dA)
        assert int(info.item()) == info_ref > 0
        k = info_ref
        assert np.array_equal(p.cpu().numpy()[:k], p_ref[:k])
        assert np.array_equal(q.cpu().numpy()[:k], q_ref[:k])
        assert np.array_equal(dev.to_host(dA), F_ref)  # A as the reference leaves it
