"""Parity of the plain FP64 GPU path with the oracle (the CPU restatement of Elemental).

Bars:
* pivots: identical swap destinations; a disagreement is allowed only if it is a counted
  near-tie (the oracle's candidates at the first differing step are within 1e-8 relative);
* factors: ||PA - LU|| / ||A|| <= c n u (c = 8) and elementwise close to the oracle;
* solutions: DD-residual backward error <= n u, and close to the oracle's;
* kernels whose operation order equals the restated BLAS (panel, small triangular solves,
  unblocked Cholesky) are compared bit for bit.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def say(*a):
    print("[case]", *a, flush=True)


def test_dgemm_all_orientations(dev, orc):
    rng = np.random.default_rng(0)
    for (m, n, k) in [(1, 1, 1), (37, 53, 19), (128, 128, 128), (300, 257, 130), (513, 65, 7)]:
        for ta in (0, 1):
            for tb in (0, 1):
                say("dgemm", m, n, k, ta, tb)
                A = rng.standard_normal((m, k) if ta == 0 else (k, m))
                B = rng.standard_normal((k, n) if tb == 0 else (n, k))
                C = rng.standard_normal((m, n))
                dA, dB, dC = dev.to_device(A), dev.to_device(B), dev.to_device(C)
                dev.gemm(ta, tb, -1.0, dA, dB, 0.5, dC)
                ref = 0.5 * C - (A if ta == 0 else A.T) @ (B if tb == 0 else B.T)
                got = dev.to_host(dC)
                assert np.abs(got - ref).max() <= 4 * k * EPS * (np.abs(A).max() * np.abs(B).max() * k + 1)


def test_dgemm_beta_zero_ignores_nan(dev):
    A = np.ones((40, 3))
    B = np.ones((3, 50))
    dC = dev.to_device(np.full((40, 50), np.nan))
    dev.gemm(0, 0, 1.0, dev.to_device(A), dev.to_device(B), 0.0, dC)
    assert np.array_equal(dev.to_host(dC), np.full((40, 50), 3.0))


def test_dsyrk_triangles(dev, orc):
    rng = np.random.default_rng(1)
    for (n, k) in [(5, 3), (130, 64), (257, 128)]:
        for uplo in (0, 1):
            for orient in (0, 1):
                say("dsyrk", n, k, uplo, orient)
                A = rng.standard_normal((n, k) if orient == 0 else (k, n))
                C = rng.standard_normal((n, n))
                dC = dev.to_device(C)
                dev.syrk(uplo, orient, -1.0, dev.to_device(A), 1.0, dC)
                got = dev.to_host(dC)
                full = C - (A @ A.T if orient == 0 else A.T @ A)
                tri = np.tril if uplo == 0 else np.triu
                other = np.triu if uplo == 0 else np.tril
                assert np.allclose(tri(got), tri(full), atol=1e-12 * k)
                assert np.array_equal(other(got, 1 if uplo == 0 else -1), other(C, 1 if uplo == 0 else -1))


@pytest.mark.parametrize("side", [0, 1])
def test_trsm_small_bitwise_and_blocked(dev, orc, side):
    rng = np.random.default_rng(2)
    sides = "LR"
    for order in (1, 17, 128, 300):
        for uplo in (0, 1):
            for orient in (0, 1):
                for diag in (0, 1):
                    T = rng.standard_normal((order, order)) + 4 * np.eye(order)
                    # 45 right-hand sides, and one (the solve's one-column kernel on the left)
                    for other in (45, 1):
                        say("trsm", sides[side], order, uplo, orient, diag, other)
                        B = rng.standard_normal((order, other) if side == 0 else (other, order))
                        dB = dev.to_device(B)
                        dev.trsm(side, uplo, orient, diag, 1.0, dev.to_device(T), dB)
                        got = dev.to_host(dB)
                        ref = orc.trsm("LR"[side], "LU"[uplo], "NT"[orient], "NU"[diag], 1.0,
                                       np.asfortranarray(T), np.asfortranarray(B))
                        if order <= 128:
                            assert np.array_equal(got, ref), "small triangular solve must match the restated BLAS bit for bit"
                        else:
                            scale = np.abs(ref).max()
                            assert np.abs(got - ref).max() <= 1e-11 * scale


def check_lu(dev, orc, A, nb=128, expect_exact=False):
    m, n = A.shape
    F_ref, d_ref, info_ref = orc.lu(A, bsize=nb)
    dA = dev.to_device(A)
    d, info = dev.getrf(dA, nb=nb)
    torch.cuda.synchronize()
    F = dev.to_host(dA)
    d = d.cpu().numpy()
    assert int(info.item()) == info_ref
    agree = orc.pivot_agreement(A, d_ref, d, bsize=nb)
    if agree["mismatches"]:
        say("pivot disagreement", agree)
        assert agree["near_ties"] == 1, f"pivot disagreement that is not a near-tie: {agree}"
        # the GPU's own factors must still be a backward-stable LU with its own pivots
        if m == n and n > 0:
            err = orc.lu_factor_error(A, F, d)
            assert err <= 8 * n * EPS, err
        return agree
    if expect_exact:
        assert np.array_equal(F, F_ref), "a single panel must match lu::Panel bit for bit"
    if m == n and n > 0:
        err = orc.lu_factor_error(A, F, d)
        assert err <= 8 * n * EPS, err
    scale = np.abs(F_ref).max() if F_ref.size else 1.0
    assert np.abs(F - F_ref).max() <= 1e3 * max(m, n) * EPS * scale
    return agree


def test_lu_shapes(dev, orc):
    orc.rng_seed()
    for (m, n, nb) in [(1, 1, 128), (5, 5, 2), (64, 64, 16), (127, 127, 128), (128, 128, 128),
                       (129, 129, 128), (300, 300, 96), (400, 250, 128), (250, 400, 128)]:
        say("lu", m, n, nb)
        A = orc.gaussian(m, n)
        check_lu(dev, orc, A, nb)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_trsm_many_rhs_blocked_bitwise(dev, orc, dtype):
    """Left lower unit solves with 256+ right-hand sides (the LU's U12 = L11^-1 A12) run in
    32-row blocks, each entry updated one FMA at a time in increasing k: bit for bit the
    restated BLAS, at every order up to 128, ragged column counts and padded leading
    dimensions; and the U12 block of a two-panel LU equals the oracle's bit for bit."""
    rng = np.random.default_rng(5)
    for order, ncols in [(1, 256), (31, 300), (32, 257), (33, 1000), (100, 513), (128, 4000)]:
        say("trsm llnu", order, ncols, dtype.__name__)
        T = (np.tril(rng.standard_normal((order, order)), -1) + np.eye(order)).astype(dtype)
        B = rng.standard_normal((order, ncols)).astype(dtype)
        dB = dev.to_device(B)
        dev.trsm(0, 0, 0, 1, 1.0, dev.to_device(T), dB)
        ref = orc.trsm("L", "L", "N", "U", 1.0, np.asfortranarray(T.astype(np.float64)),
                       np.asfortranarray(B.astype(np.float64))) if dtype == np.float64 else None
        got = dev.to_host(dB)
        if dtype == np.float64:
            assert np.array_equal(got, ref)
        else:
            # float: against the small kernel (under 256 columns) column by column
            for c0 in range(0, ncols, 200):
                part = dev.to_device(B[:, c0:c0 + 200])
                dev.trsm(0, 0, 0, 1, 1.0, dev.to_device(T), part)
                assert np.array_equal(dev.to_host(part), got[:, c0:c0 + 200])
    orc.rng_seed(23)
    A = orc.gaussian(256, 256)
    F_ref, d_ref, _ = orc.lu(A, bsize=128)
    dA = dev.to_device(A)
    d, info = dev.getrf(dA, nb=128)
    F = dev.to_host(dA)
    assert int(info.item()) == 0 and np.array_equal(d.cpu().numpy(), d_ref)
    assert np.array_equal(F[:128, :], F_ref[:128, :]), "first panel and U12 bit for bit"


def test_lu_single_panel_bitwise(dev, orc):
    """n <= nb (and tall single panels: one cooperative panel over many workgroups) runs only
    the panel kernel, whose per-entry FMA order is lu::Panel's: the factors are asserted equal
    to the oracle's bit for bit."""
    orc.rng_seed(17)
    for (m, n, nb) in [(1, 1, 128), (7, 7, 128), (100, 100, 128), (128, 128, 128),
                       (300, 64, 64), (4000, 128, 128), (20000, 96, 96), (32768, 128, 128)]:
        say("single panel bitwise", m, n, nb)
        A = orc.gaussian(m, n)
        check_lu(dev, orc, A, nb, expect_exact=True)


@pytest.mark.parametrize("exchange", ["memory", "xcd"])
def test_lu_one_xcd_panels(dev, orc, exchange, monkeypatch):
    """Blocksize() <= 32, with the panel exchange through uncached memory (the default) and
    with ELX_PANEL_EXCHANGE=xcd, where the panels run on the CUs of one XCD and exchange their
    candidates through its L2 (up to 1024 rows per CU). Same bars: single panels bit for bit,
    blocked factorizations with identical pivots."""
    monkeypatch.setenv("ELX_PANEL_EXCHANGE", exchange)
    orc.rng_seed(18)
    for (m, n, nb) in [(40, 40, 32), (1000, 32, 32), (17000, 16, 16), (32768, 32, 32)]:
        say("one-XCD panel bitwise", m, n, nb)
        A = orc.gaussian(m, n)
        check_lu(dev, orc, A, nb, expect_exact=True)
        F_ref, d_ref, _ = orc.lu(A.astype(np.float32), bsize=nb)
        dA = dev.to_device(A.astype(np.float32))
        d, info = dev.getrf(dA, nb=nb)
        assert int(info.item()) == 0
        assert np.array_equal(d.cpu().numpy(), d_ref)
        assert np.array_equal(dev.to_host(dA), F_ref), "float one-XCD panel bit for bit"
    for (n, nb) in [(700, 32), (1500, 16)]:
        say("one-XCD panels, blocked LU", n, nb)
        check_lu(dev, orc, orc.gaussian(n, n), nb)


def test_lu_config1_n1024(dev, orc):
    """Config 1: n=1024, A and b from Elemental's Gaussian stream; pivots, factor error,
    DD backward error of the solve."""
    n = 1024
    orc.rng_seed()
    A = orc.gaussian(n, n)
    b = orc.gaussian(n, 1)
    say("lu config1", n)
    agree = check_lu(dev, orc, A)
    dA = dev.to_device(A)
    d, info = dev.getrf(dA)
    dB = dev.to_device(b)
    dev.getrs(dA, d, dB)
    x = dev.to_host(dB)
    eta = orc.backward_error(A, x, b)
    say("config1 backward error / (n u) =", eta / (n * EPS), "pivot agreement", agree)
    assert eta <= n * EPS
    F_ref, d_ref, _ = orc.lu(A)
    x_ref = orc.lu_solve_after(F_ref, d_ref, b)
    assert np.abs(x - x_ref).max() <= 1e-9 * np.abs(x_ref).max()


def test_lu_gepp_growth_exact(dev, orc):
    for n in (2, 33, 129, 600):
        say("gepp growth", n)
        A = orc.gepp_growth(n)
        dA = dev.to_device(A)
        d, info = dev.getrf(dA, nb=128)
        F = dev.to_host(dA)
        assert np.array_equal(d.cpu().numpy(), np.arange(n, dtype=np.int32))
        assert np.array_equal(np.tril(F, -1), np.tril(-np.ones((n, n)), -1))
        if n <= 1000:
            assert np.array_equal(F[:, n - 1], 2.0 ** np.arange(n))


def test_lu_ties_go_to_first_row(dev, orc):
    # many exact ties in every column: integer matrix with entries in {-1, 0, 1}
    rng = np.random.default_rng(9)
    for n in (8, 50, 200):
        say("ties", n)
        A = rng.integers(-1, 2, size=(n, n)).astype(np.float64) + np.eye(n) * 0
        F_ref, d_ref, info_ref = orc.lu(A, bsize=16)
        dA = dev.to_device(A)
        d, info = dev.getrf(dA, nb=16)
        if info_ref == 0:
            assert np.array_equal(d.cpu().numpy()[: n // 2], d_ref[: n // 2])


def test_lu_singular(dev, orc):
    A = np.arange(1.0, 26).reshape(5, 5)
    A[:, 2] = 0.0
    dA = dev.to_device(A)
    d, info = dev.getrf(dA, nb=2)
    assert int(info.item()) == 3
    Z = np.zeros((300, 300))
    Z[:, :130] = np.random.default_rng(0).standard_normal((300, 130))
    dZ = dev.to_device(Z)
    d, info = dev.getrf(dZ)
    assert int(info.item()) == 131


def test_lu_solve_after_orientations(dev, orc):
    orc.rng_seed(5)
    for n in (7, 130, 520):
        A = orc.gaussian(n, n)
        B = orc.gaussian(n, 5)
        F, d, _ = orc.lu(A)
        for orient in (0, 1, 2):
            say("solve after", n, orient)
            dF = dev.to_device(F)
            dd = torch.from_numpy(d).cuda()
            dB = dev.to_device(B)
            dev.getrs(dF, dd, dB, orient=orient)
            got = dev.to_host(dB)
            ref = orc.lu_solve_after(F, d, B, orient=min(orient, 1))
            assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def test_linear_solve_row_echelon(dev, orc):
    orc.rng_seed(6)
    for n in (1, 100, 129, 700):
        say("linear solve", n)
        A = orc.gaussian(n, n)
        B = orc.gaussian(n, 3)
        X_ref, d_ref, info_ref = orc.linear_solve(A, B)
        dA, dB = dev.to_device(A), dev.to_device(B)
        d, info = dev.gesv(dA, dB, want_dests=True)
        X = dev.to_host(dB)
        assert int(info.item()) == 0
        agree = orc.pivot_agreement(A, d_ref, d.cpu().numpy())
        assert agree["mismatches"] == 0 or agree["near_ties"] == 1
        assert orc.backward_error(A, X, B) <= n * EPS
        assert np.abs(X - X_ref).max() <= 1e-9 * max(1.0, np.abs(X_ref).max())


def test_permute_rows(dev, orc):
    rng = np.random.default_rng(3)
    n = 777
    dests = np.array([rng.integers(j, n) for j in range(n)], np.int32)
    B = rng.standard_normal((n, 4))
    p = orc.perm_explicit(dests, n)
    for inverse in (False, True):
        dB = dev.to_device(B)
        dev.permute_rows(torch.from_numpy(dests).cuda(), dB, inverse=inverse)
        got = dev.to_host(dB)
        ref = B[p] if not inverse else B[np.argsort(p)]
        assert np.array_equal(got, ref)


def hpd(orc, n, seed):
    orc.rng_seed(seed)
    G = orc.gaussian(n, n)
    return np.asfortranarray(G @ G.T / n + np.eye(n))


def test_cholesky_exact_known_answers(dev, golden):
    for idx in range(int(golden["chol_count"])):
        G = golden[f"chol{idx}_G"]
        A = golden[f"chol{idx}_A"]
        for uplo in (0, 1):
            dA = dev.to_device(A)
            info = dev.potrf(dA, uplo=uplo, nb=4)
            F = dev.to_host(dA)
            assert int(info.item()) == 0
            if uplo == 0:
                assert np.array_equal(np.tril(F), G)
            else:
                assert np.array_equal(np.triu(F), G.T)


def test_cholesky_vs_oracle(dev, orc):
    for n, nb in [(1, 128), (50, 16), (128, 128), (300, 128), (1000, 128)]:
        A = hpd(orc, n, n)
        for uplo in (0, 1):
            say("cholesky", n, nb, uplo)
            F_ref, info_ref = orc.cholesky(A, uplo=uplo, bsize=nb)
            dA = dev.to_device(A)
            info = dev.potrf(dA, uplo=uplo, nb=nb)
            F = dev.to_host(dA)
            assert int(info.item()) == info_ref == 0
            tri = np.tril if uplo == 0 else np.triu
            assert orc.chol_factor_error(A, F, uplo) <= 8 * n * EPS
            assert np.abs(tri(F) - tri(F_ref)).max() <= 64 * n * EPS * np.abs(F_ref).max()
            other = np.triu if uplo == 0 else np.tril
            k = 1 if uplo == 0 else -1
            assert np.array_equal(other(F, k), other(A, k))  # untouched triangle
            if n <= nb:  # one unblocked step: same operation order as the restated BLAS
                assert np.array_equal(tri(F), tri(F_ref))


def test_cholesky_non_hpd(dev):
    A = np.diag([4.0, 1.0, -2.0, 3.0])
    for uplo in (0, 1):
        info = dev.potrf(dev.to_device(A), uplo=uplo)
        assert int(info.item()) == 3
    A = -np.eye(300)
    A[:200, :200] = np.eye(200)
    info = dev.potrf(dev.to_device(A))
    assert int(info.item()) == 201


def test_hpd_solve(dev, orc):
    for n in (10, 257, 1000):
        A = hpd(orc, n, 100 + n)
        orc.rng_seed(n)
        B = orc.gaussian(n, 16)
        for uplo in (0, 1):
            say("hpd solve", n, uplo)
            X_ref, _ = orc.hpd_solve(A, B, uplo=uplo)
            dA, dB = dev.to_device(A), dev.to_device(B)
            info = dev.posv(dA, dB, uplo=uplo)
            X = dev.to_host(dB)
            assert int(info.item()) == 0
            assert orc.backward_error(A, X, B) <= n * EPS
            assert np.abs(X - X_ref).max() <= 1e-10 * np.abs(X_ref).max()
