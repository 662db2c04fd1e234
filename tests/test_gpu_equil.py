"""Ruiz equilibration on the GPU (elx_druiz_equil, El::RuizEquil): bitwise equal to the
oracle, and the mixed solver's equilibrated factorization on badly scaled systems."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def say(*a):
    print("[case]", *a, flush=True)


def scaled(rng, m, n, spread):
    return np.asfortranarray((2.0 ** rng.uniform(-spread, spread, m))[:, None] *
                             rng.standard_normal((m, n)) *
                             (2.0 ** rng.uniform(-spread, spread, n))[None, :])


@pytest.mark.parametrize("m,n,ld_extra", [(300, 200, 0), (1000, 1000, 24), (3, 70000, 5),
                                          (1, 1, 0), (257, 64, 3)])
def test_ruiz_bitwise(dev, orc, m, n, ld_extra):
    say("ruiz", m, n, ld_extra)
    rng = np.random.default_rng(m * 7 + n)
    A = scaled(rng, m, n, 40)
    if n > 3:
        A[:, 2] = 0.0
    if m > 3:
        A[1, :] = 0.0
        A[2, 0] = 1e-200
    Ae, dR, dC = orc.ruiz_equil(A)
    T = torch.zeros((n, m + ld_extra), dtype=torch.float64, device="cuda").t()[:m, :]
    T.copy_(torch.from_numpy(A))
    gR, gC = dev.ruiz_equil(T)
    torch.cuda.synchronize()
    assert np.array_equal(T.cpu().numpy(), Ae)
    assert np.array_equal(gR.cpu().numpy(), dR)
    assert np.array_equal(gC.cpu().numpy(), dC)


@pytest.mark.parametrize("n", [700, 3000])
def test_mixed_equilibrated_badly_scaled(dev, orc, n):
    """Rows and columns scaled by 2^-24 .. 2^24: the fp16 LU of the equilibrated matrix is a
    good preconditioner again; every system converges to n u without the FP64 fallback."""
    rng = np.random.default_rng(n)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    core = (Q1 * 10.0 ** (-2 * np.arange(n) / (n - 1))[None, :]) @ Q2.T
    A = np.asfortranarray((2.0 ** rng.uniform(-24, 24, n))[:, None] * core *
                          (2.0 ** rng.uniform(-24, 24, n))[None, :])
    B = rng.standard_normal((n, 3))
    dA, dB = dev.to_device(A), dev.to_device(B)
    dX = dev.colmajor_empty(n, 3)
    plain = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="auto")
    say(n, "without equilibration", {k: plain[k] for k in ("converged", "outer_iters",
                                                           "inner_iters", "fell_back")})
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="auto", equilibrate=True)
    X = dev.to_host(dX)
    eta = orc.backward_error(A, X, B)
    say(n, "equilibrated", st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and st["fell_back"] == 0
    assert eta <= n * EPS
    assert np.array_equal(dev.to_host(dA), A)  # A is not modified


@pytest.mark.parametrize("n,ld_extra,maxiter", [(300, 0, 3), (1000, 17, 3), (1, 0, 3),
                                                (257, 3, 1), (64, 0, 0)])
def test_symmetric_ruiz_bitwise(dev, orc, n, ld_extra, maxiter):
    say("symmetric ruiz", n, ld_extra, maxiter)
    rng = np.random.default_rng(n + maxiter)
    G = rng.standard_normal((n, n))
    D = 2.0 ** rng.uniform(-40, 40, n)
    A = np.asfortranarray(D[:, None] * (G @ G.T / n + np.eye(n)) * D[None, :])
    if n > 4:
        A[3, :] = 0.0
        A[:, 3] = 0.0
    Ae, d = orc.symmetric_ruiz_equil(A, maxiter=maxiter)
    T = torch.zeros((n, n + ld_extra), dtype=torch.float64, device="cuda").t()[:n, :]
    T.copy_(torch.from_numpy(A))
    gd = dev.symmetric_ruiz_equil(T, maxiter=maxiter)
    torch.cuda.synchronize()
    assert np.array_equal(T.cpu().numpy(), Ae)
    assert np.array_equal(gd.cpu().numpy(), d)


@pytest.mark.parametrize("uplo", [0, 1])
def test_mixed_hpd_equilibrated_badly_scaled(dev, orc, uplo):
    """An HPD matrix with its rows and columns scaled by 2^-20 .. 2^20: the fp16 Cholesky of the
    symmetrically equilibrated matrix preconditions the refinement; every column reaches n u
    without the FP64 fallback, and only the uplo triangle is read."""
    n = 1500
    rng = np.random.default_rng(11 + uplo)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    core = (Q * 10.0 ** (-2 * np.arange(n) / (n - 1))[None, :]) @ Q.T
    D = 2.0 ** rng.uniform(-20, 20, n)
    A = D[:, None] * ((core + core.T) / 2) * D[None, :]
    Ast = A.copy(order="F")
    if uplo == 0:
        Ast[np.triu_indices(n, 1)] = np.nan
    else:
        Ast[np.tril_indices(n, -1)] = np.nan
    B = rng.standard_normal((n, 2))
    dA, dB = dev.to_device(Ast), dev.to_device(B)
    dX = dev.colmajor_empty(n, 2)
    st = dev.posv_mixed(dA, dB, dX, uplo=uplo, factor="fp16", refine="auto", equilibrate=True)
    X = dev.to_host(dX)
    eta = orc.backward_error(A, X, B)
    say("hpd equilibrated", uplo, st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and st["fell_back"] == 0 and st["factor_info"] == 0
    assert eta <= n * EPS
    assert np.array_equal(dev.to_host(dA), Ast, equal_nan=True)


@pytest.mark.parametrize("m,n,ld_extra", [(300, 200, 0), (1000, 1000, 24), (3, 70000, 5),
                                          (1, 1, 0), (257, 64, 3), (64, 300, 0)])
def test_geom_bitwise(dev, orc, m, n, ld_extra):
    say("geom", m, n, ld_extra)
    rng = np.random.default_rng(m * 3 + n)
    A = scaled(rng, m, n, 40)
    if n > 3:
        A[:, 2] = 0.0
    if m > 3:
        A[1, :] = 0.0
        A[2, 0] = 1e-200
    Ae, dR, dC = orc.geom_equil(A)
    T = torch.zeros((n, m + ld_extra), dtype=torch.float64, device="cuda").t()[:m, :]
    T.copy_(torch.from_numpy(A))
    gR, gC = dev.geom_equil(T)
    torch.cuda.synchronize()
    assert np.array_equal(T.cpu().numpy(), Ae)
    assert np.array_equal(gR.cpu().numpy(), dR)
    assert np.array_equal(gC.cpu().numpy(), dC)


def test_geom_zero_matrix(dev):
    T = torch.zeros((5, 4), dtype=torch.float64, device="cuda").t().contiguous().t()
    gR, gC = dev.geom_equil(T)
    torch.cuda.synchronize()
    assert torch.all(T == 0) and torch.all(gR == 1) and torch.all(gC == 1)


@pytest.mark.parametrize("kind", ["gesv", "posv"])
def test_mixed_equilibrated_escalation(dev, orc, kind):
    """Scaled (2^-3 .. 2^3) and ill-conditioned (kappa 1e6) systems: the FP16 factors of the
    equilibrated matrix do not precondition well enough, so AUTO refactors FP32-accurately
    (still equilibrated) or falls back to FP64, and the result meets n u. (Wider scalings make
    the normwise backward error too weak a test: with rows spanning 2^32, even a one-shot FP16
    solve scores about n u.)"""
    n = 800
    rng = np.random.default_rng(31 if kind == "gesv" else 32)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sig = 10.0 ** (-6 * np.arange(n) / (n - 1))
    if kind == "gesv":
        Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        core = (Q1 * sig[None, :]) @ Q2.T
        Dr, Dc = 2.0 ** rng.uniform(-3, 3, n), 2.0 ** rng.uniform(-3, 3, n)
    else:
        core = (Q1 * sig[None, :]) @ Q1.T
        core = (core + core.T) / 2
        Dr = Dc = 2.0 ** rng.uniform(-3, 3, n)
    A = np.asfortranarray(Dr[:, None] * core * Dc[None, :])
    B = rng.standard_normal((n, 2))
    dA, dB = dev.to_device(A), dev.to_device(B)
    dX = dev.colmajor_empty(n, 2)
    if kind == "gesv":
        st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="auto", equilibrate=True)
    else:
        st = dev.posv_mixed(dA, dB, dX, factor="fp16", refine="auto", equilibrate=True)
    X = dev.to_host(dX)
    eta = orc.backward_error(A, X, B)
    say("equilibrated escalation", kind, st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and eta <= n * EPS
    assert st["fell_back"] in (1, 2)
