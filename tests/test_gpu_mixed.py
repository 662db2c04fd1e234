"""Mixed-precision LU + FP64 iterative refinement on the GPU.

Bar: every solve reports FP64 normwise backward error <= n * 2^-53, checked
here independently with the oracle's double-double residual; the factorization alone
(no refinement) must be a useful preconditioner (backward error well below 1)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def say(*a):
    print("[case]", *a, flush=True)


def randsvd(n, kappa, seed):
    """A = U diag(sigma) V^T, U and V Haar (QR of Gaussians with sign-fixed R), sigma
    geometric from 1 to 1/kappa."""
    rng = np.random.default_rng(seed)
    def haar():
        Q, R = np.linalg.qr(rng.standard_normal((n, n)))
        return Q * np.sign(np.diag(R))[None, :]
    U, V = haar(), haar()
    sigma = kappa ** (-np.arange(n) / max(n - 1, 1))
    return np.asfortranarray((U * sigma[None, :]) @ V.T)


@pytest.mark.parametrize("factor", ["fp16", "bf16", "fp32"])
@pytest.mark.parametrize("n", [200, 1000, 2500])
def test_ir_converges_well_conditioned(dev, orc, factor, n):
    say("mixed", factor, n)
    A = randsvd(n, 1e2, n)
    b = np.random.default_rng(n + 1).standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor=factor, refine="auto")
    x = dev.to_host(dX)
    eta = orc.backward_error(A, x, b)
    say(factor, n, st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and st["factor_info"] == 0
    assert eta <= n * EPS
    # the device's own backward error agrees with the double-double one
    assert st["backward_error"] <= 2 * n * EPS
    assert abs(st["backward_error"] - eta) <= 0.5 * max(eta, st["backward_error"]) + 1e-20


@pytest.mark.parametrize("factor", ["fp16", "bf16", "fp32"])
def test_factor_alone_is_a_preconditioner(dev, orc, factor):
    n = 1024
    A = randsvd(n, 1e2, 7)
    b = np.random.default_rng(8).standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor=factor, refine="none")
    assert st["status"] in (0, 5)
    x = dev.to_host(dX)
    eta = orc.backward_error(A, x, b)
    u_low = {"fp16": 2.0 ** -11, "bf16": 2.0 ** -8, "fp32": 2.0 ** -22}[factor]
    say(factor, "one-shot backward error", eta, "u_low", u_low)
    assert eta < 50 * u_low


@pytest.mark.parametrize("factor", ["fp16", "bf16", "fp32"])
def test_rank_nb_update_kernel_matches_library(dev, orc, factor, monkeypatch):
    """The rank-1024 trailing update of the mixed LU (n = 3000: two of them) on the
    hand-written 64-deep-slab MFMA kernel and on rocBLAS gemm_ex (ELX_LOWP_GEMM=library): the
    same pivots and one-shot solves of the same quality. Both accumulate in FP32 from the same
    fp16/bf16 packs, so only the order of the FP32 additions differs."""
    n = 3000
    A = randsvd(n, 1e2, 27)
    b = np.random.default_rng(28).standard_normal((n, 1))
    etas = {}
    for gemm in ("own", "library"):
        if gemm == "library":
            monkeypatch.setenv("ELX_LOWP_GEMM", "library")
        else:
            monkeypatch.delenv("ELX_LOWP_GEMM", raising=False)
        dA, dB = dev.to_device(A), dev.to_device(b)
        dX = dev.colmajor_empty(n, 1)
        st = dev.gesv_mixed(dA, dB, dX, factor=factor, refine="none")
        assert st["status"] in (0, 5) and st["factor_info"] == 0
        etas[gemm] = orc.backward_error(A, dev.to_host(dX), b)
    say(factor, "one-shot backward error own / rocBLAS", etas)
    u_low = {"fp16": 2.0 ** -11, "bf16": 2.0 ** -8, "fp32": 2.0 ** -22}[factor]
    assert etas["own"] < 50 * u_low
    assert etas["own"] <= 2.0 * etas["library"] and etas["library"] <= 2.0 * etas["own"]


def test_mixed_multiple_rhs(dev, orc):
    n = 700
    A = randsvd(n, 1e2, 3)
    B = np.random.default_rng(4).standard_normal((n, 3))
    dA, dB = dev.to_device(A), dev.to_device(B)
    dX = dev.colmajor_empty(n, 3)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="ir")
    X = dev.to_host(dX)
    assert st["converged"] == 1
    assert orc.backward_error(A, X, B) <= n * EPS


def test_mixed_pivoting_needed(dev, orc):
    """A matrix that needs row pivoting (tiny leading entries) still converges."""
    n = 600
    rng = np.random.default_rng(12)
    A = rng.standard_normal((n, n))
    A[np.arange(n), np.arange(n)] *= 1e-8
    A = np.asfortranarray(A)
    b = rng.standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="ir")
    x = dev.to_host(dX)
    assert st["converged"] == 1
    assert orc.backward_error(A, x, b) <= n * EPS


def test_gmres_ir_ill_conditioned(dev, orc):
    """kappa * u16 = 5: plain refinement converges slowly or not at all; GMRES-IR
    preconditioned by the fp16 LU reaches FP64 backward error without the fallback."""
    n = 1500
    say("gmres-ir fp16 1e4")
    A = randsvd(n, 1e4, 21)
    b = np.random.default_rng(22).standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="gmres")
    x = dev.to_host(dX)
    eta = orc.backward_error(A, x, b)
    say(st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and st["fell_back"] == 0, st
    assert eta <= n * EPS
    assert st["inner_iters"] > 0


@pytest.mark.parametrize("kappa", [1e4, 1e6])
@pytest.mark.parametrize("factor", ["fp16", "bf16"])
def test_auto_always_reaches_fp64_accuracy(dev, orc, kappa, factor):
    """AUTO goes IR -> GMRES-IR -> FP32-accurate refactorization -> FP64 LU: whatever
    kappa * u_low is, the answer meets the FP64 backward-error bound, and up to kappa = 1e6
    the FP32-accurate factors suffice (no FP64 LU)."""
    n = 1500
    say("auto", factor, kappa)
    A = randsvd(n, kappa, 21)
    b = np.random.default_rng(22).standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor=factor, refine="auto")
    x = dev.to_host(dX)
    eta = orc.backward_error(A, x, b)
    say(factor, kappa, st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1, st
    assert eta <= n * EPS
    assert st["fell_back"] in (0, 2), st


@pytest.mark.parametrize("kappa", [1e4, 1e6])
def test_fp32_accurate_factor_ill_conditioned(dev, orc, kappa):
    """FP32-accurate factors (fp16 high/low parts, three MFMA products per update) precondition
    kappa = 1e6 well enough for plain IR to reach FP64 accuracy in a few steps."""
    n = 1500
    say("fp32-accurate", kappa)
    A = randsvd(n, kappa, 41)
    b = np.random.default_rng(42).standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp32", refine="ir")
    x = dev.to_host(dX)
    eta = orc.backward_error(A, x, b)
    say(kappa, st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and st["fell_back"] == 0, st
    assert st["outer_iters"] <= 12, st
    assert eta <= n * EPS


def test_fallback_when_low_precision_factor_breaks_down(dev, orc):
    """A column of magnitude 1e-50 is exactly zero in fp32: the low-precision LU reports a
    zero pivot, AUTO solves with the FP64 LU instead; IR/GMRES modes report the failure."""
    n = 400
    rng = np.random.default_rng(31)
    A = rng.standard_normal((n, n))
    A[:, 5] *= 1e-50
    A = np.asfortranarray(A)
    b = rng.standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="auto")
    say("breakdown", st)
    assert st["factor_info"] != 0 and st["fell_back"] == 1 and st["converged"] == 1
    x = dev.to_host(dX)
    assert orc.backward_error(A, x, b) <= n * EPS
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="ir")
    assert st["factor_info"] != 0 and st["converged"] == 0 and st["status"] == 5


def spd_kappa(n, kappa, seed):
    """A = Q diag(sigma) Q^T, Q Haar, sigma geometric from 1 to 1/kappa (HPD, condition kappa)."""
    rng = np.random.default_rng(seed)
    Q, R = np.linalg.qr(rng.standard_normal((n, n)))
    Q = Q * np.sign(np.diag(R))[None, :]
    sigma = kappa ** (-np.arange(n) / max(n - 1, 1))
    A = (Q * sigma[None, :]) @ Q.T
    return np.asfortranarray((A + A.T) / 2)


@pytest.mark.parametrize("uplo", [0, 1])
@pytest.mark.parametrize("factor", ["fp16", "bf16", "fp32"])
@pytest.mark.parametrize("n", [200, 1000, 2300])
def test_hpd_mixed_solve(dev, orc, uplo, factor, n):
    """El::HPDSolve semantics (only the uplo triangle is read: the other one holds NaNs here)
    to FP64 backward error <= n u, from a low-precision Cholesky + FP64 refinement."""
    say("posv mixed", factor, n, "uplo", uplo)
    A = spd_kappa(n, 1e2, n + 7)
    b = np.random.default_rng(n + 8).standard_normal((n, 2))
    Ast = A.copy(order="F")
    if uplo == 0:
        Ast[np.triu_indices(n, 1)] = np.nan
    else:
        Ast[np.tril_indices(n, -1)] = np.nan
    dA, dB = dev.to_device(Ast), dev.to_device(b)
    dX = dev.colmajor_empty(n, 2)
    st = dev.posv_mixed(dA, dB, dX, uplo=uplo, factor=factor, refine="auto")
    X = dev.to_host(dX)
    eta = orc.backward_error(A, X, b)
    say(st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and st["factor_info"] == 0
    assert eta <= n * EPS
    if factor == "fp16":
        assert st["fell_back"] == 0
    assert np.array_equal(dev.to_host(dA), Ast, equal_nan=True)  # A is not modified


@pytest.mark.parametrize("n", [300, 2500, 9000])
@pytest.mark.parametrize("kind", ["permutation", "lower_half_zero"])
def test_mixed_structured_pivoting(dev, orc, n, kind):
    """Structured matrices where most of the panel's workgroups hold only zeros in a panel's
    columns: a permutation matrix (exact factors, exact solution) and a matrix whose leading
    panels have their nonzeros only in the lower half of the rows. Every system must converge
    without the FP64 fallback."""
    say("structured", kind, n)
    rng = np.random.default_rng(n)
    if kind == "permutation":
        A = np.zeros((n, n))
        A[rng.permutation(n), np.arange(n)] = rng.choice([-3.0, -1.0, 2.0, 5.0], n)
    else:
        A = randsvd(n, 1e2, n)
        A[: n // 2, : n // 2] = 0.0
        A[: n // 2, n // 2:] = np.asfortranarray(randsvd(n - n // 2, 1e1, n + 1)[: n // 2])
        A[n // 2:, : n // 2] += np.eye(n - n // 2, n // 2)
        A = np.asfortranarray(A)
    b = rng.standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="auto")
    x = dev.to_host(dX)
    eta = orc.backward_error(A, x, b)
    say(st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["factor_info"] == 0 and st["converged"] == 1 and st["fell_back"] == 0
    assert eta <= n * EPS
    if kind == "permutation":
        assert st["outer_iters"] <= 1


@pytest.mark.parametrize("kind", ["gesv", "posv"])
@pytest.mark.parametrize("factor", ["fp16", "fp32"])
def test_block_refinement_many_rhs(dev, orc, kind, factor):
    """Many right-hand sides go through the block refinement (groups of 64 columns, rocBLAS
    strsm on the low-precision factors, one FP64 residual product per step): 70 columns, one
    of them zero, one scaled by 2^100 and one by 2^-100; every column meets the FP64
    backward-error bound on its own and the zero column gives x = 0."""
    n = 900
    say("block refinement", kind, factor, n)
    A = randsvd(n, 1e3, 21) if kind == "gesv" else spd_kappa(n, 1e3, 21)
    B = np.random.default_rng(22).standard_normal((n, 70))
    B[:, 5] = 0.0
    B[:, 6] *= 2.0 ** 100
    B[:, 66] *= 2.0 ** -100
    dA, dB = dev.to_device(A), dev.to_device(B)
    dX = dev.colmajor_empty(n, 70)
    if kind == "gesv":
        st = dev.gesv_mixed(dA, dB, dX, factor=factor, refine="ir")
    else:
        st = dev.posv_mixed(dA, dB, dX, factor=factor, refine="ir")
    X = dev.to_host(dX)
    say(st)
    assert st["converged"] == 1 and st["fell_back"] == 0
    assert np.all(X[:, 5] == 0.0)
    worst = 0.0
    for k in range(70):
        if k == 5:
            continue
        eta = orc.backward_error(A, X[:, k:k + 1], B[:, k:k + 1])
        worst = max(worst, eta)
        assert eta <= n * EPS, (k, eta / (n * EPS))
    assert st["backward_error"] <= 2 * n * EPS
    say("worst oracle eta/(n u) =", worst / (n * EPS))


@pytest.mark.parametrize("factor", ["fp16", "bf16"])
def test_block_one_shot_solve(dev, orc, factor):
    """refine = none with several right-hand sides: one block solve with the low-precision
    factors, each column a useful approximation, reported without refinement."""
    n = 1024
    A = randsvd(n, 1e2, 7)
    B = np.random.default_rng(9).standard_normal((n, 5))
    dA, dB = dev.to_device(A), dev.to_device(B)
    dX = dev.colmajor_empty(n, 5)
    st = dev.gesv_mixed(dA, dB, dX, factor=factor, refine="none")
    assert st["status"] in (0, 5) and st["outer_iters"] == 0
    X = dev.to_host(dX)
    u_low = {"fp16": 2.0 ** -11, "bf16": 2.0 ** -8}[factor]
    for k in range(5):
        eta = orc.backward_error(A, X[:, k:k + 1], B[:, k:k + 1])
        assert eta < 50 * u_low, (k, eta)


def test_hpd_mixed_not_positive_definite(dev):
    """A symmetric indefinite matrix: the low-precision Cholesky reports the column, the FP64
    fallback fails the same way, and the call says so instead of returning an answer."""
    n = 300
    rng = np.random.default_rng(3)
    G = rng.standard_normal((n, n))
    A = np.asfortranarray(G @ G.T / n + np.eye(n))
    A[150, 150] = -50.0
    b = rng.standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(n, 1)
    st = dev.posv_mixed(dA, dB, dX, factor="fp16", refine="auto")
    assert st["factor_info"] > 0 and st["converged"] == 0 and st["status"] == 5


def test_batched_mixed_lu_ir(dev, orc):
    """The batched LU-IR configuration in small: nonsymmetric randsvd systems of log-uniform
    order and condition (1e1 .. 1e8), 4 right-hand sides each, fp16 factors with AUTO
    refinement: every system meets the FP64 backward-error bound, the ill-conditioned ones
    through the FP32-accurate refactorization or the FP64 fallback."""
    rng = np.random.default_rng(77)
    cnt = 14
    ns = np.exp(rng.uniform(np.log(64), np.log(700), cnt)).astype(int)
    kappas = 10.0 ** rng.uniform(1, 8, cnt)
    kappas[0], kappas[1] = 1e8, 10.0
    As = [randsvd(int(n), float(k), 100 + i) for i, (n, k) in enumerate(zip(ns, kappas))]
    Bs = [rng.standard_normal((int(n), 4)) for n in ns]
    dA = [dev.to_device(A) for A in As]
    dB = [dev.to_device(B) for B in Bs]
    dX = [dev.colmajor_empty(int(n), 4) for n in ns]
    say("batched lu-ir", list(ns), [f"{k:.0e}" for k in kappas])
    status, stats = dev.gesv_mixed_batched(dA, dB, dX, factor="fp16", refine="auto")
    assert status == 0
    for i in range(cnt):
        X = dev.to_host(dX[i])
        eta = orc.backward_error(As[i], X, Bs[i])
        say(i, ns[i], f"{kappas[i]:.1e}", stats[i]["fell_back"], eta / (ns[i] * EPS))
        assert stats[i]["converged"] == 1 and eta <= ns[i] * EPS, (i, ns[i], kappas[i], stats[i])
    assert stats[0]["fell_back"] in (1, 2) and stats[1]["fell_back"] == 0


def padded(a, extra):
    """A column-major device copy of a with leading dimension rows + extra."""
    a = np.asarray(a)
    m, n = a.shape
    t = torch.zeros((n, m + extra), dtype=torch.float64, device="cuda").t()[:m, :]
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
    return t


@pytest.mark.parametrize("n", [1, 2, 127, 129, 300])
@pytest.mark.parametrize("kind", ["gesv", "posv"])
def test_mixed_edge_shapes(dev, orc, n, kind):
    """Orders around the 128-row block (and 1), leading dimensions above the order for A, B
    and X, several right-hand sides: the solve meets n u and leaves the padding untouched."""
    say("edge", kind, n)
    rng = np.random.default_rng(1000 + n)
    if kind == "gesv":
        A = rng.standard_normal((n, n)) + n * np.eye(n)
    else:
        G = rng.standard_normal((n, n))
        A = G @ G.T / n + np.eye(n)
    B = rng.standard_normal((n, 3))
    dA, dB = padded(A, 3), padded(B, 5)
    dX = padded(np.zeros((n, 3)), 2)
    if kind == "gesv":
        st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="auto")
    else:
        st = dev.posv_mixed(dA, dB, dX, factor="fp16", refine="auto")
    X = dX.cpu().numpy()
    eta = orc.backward_error(A, X, B)
    assert st["converged"] == 1 and eta <= max(n, 1) * EPS, (st, eta)
    # the padding rows of X's storage are untouched
    base = dX.as_strided((n + 2, 3), (1, n + 2))
    assert torch.all(base[n:, :] == 0)


def test_mixed_no_right_hand_sides(dev):
    n = 50
    A = np.eye(n) * 3.0
    dA = dev.to_device(A)
    dB = dev.colmajor_empty(n, 0)
    dX = dev.colmajor_empty(n, 0)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="auto")
    assert st["converged"] == 1 and st["status"] == 0


def test_ir_keeps_only_decreasing_candidates(dev, orc):
    """Refined.hpp:96-105: a refinement candidate x + dx replaces x only when its residual is
    smaller. With fp16 factors of a kappa = 1e10 system the preconditioned operator is far
    from the identity: the first IR correction dx = M^-1 (b - A x0) (computed here by the same
    factorization applied to the residual) makes the residual larger, so plain IR must return
    x0 = M^-1 b itself, bit for bit the unrefined solve."""
    n = 400
    A = randsvd(n, 1e10, 21)
    b = np.random.default_rng(22).standard_normal((n, 1))
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX0 = dev.colmajor_empty(n, 1)
    dev.gesv_mixed(dA, dB, dX0, factor="fp16", refine="none")
    x0 = dev.to_host(dX0)
    r0 = b - A @ x0
    dD = dev.colmajor_empty(n, 1)
    dev.gesv_mixed(dA, dev.to_device(r0), dD, factor="fp16", refine="none")
    x1 = x0 + dev.to_host(dD)
    grow = np.abs(b - A @ x1).max() / np.abs(r0).max()
    say("keep-if-decreasing: residual ratio of the first candidate", grow)
    assert grow > 1.05, "premise: the first correction must increase the residual"
    dX = dev.colmajor_empty(n, 1)
    st = dev.gesv_mixed(dA, dB, dX, factor="fp16", refine="ir")
    say(st)
    assert np.array_equal(dev.to_host(dX), x0)
    assert st["converged"] == 0 and st["status"] == 5
