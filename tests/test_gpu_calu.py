"""Tournament-pivoting panel (communication-avoiding LU) of the mixed-precision solve.

The panel is chosen by ELX_MIXED_PANEL (calu16 / calu32: block widths 16 and 32; partial:
lu::Panel's partial pivoting), read at every solve. Its pivot order differs from partial
pivoting by design, so these tests hold it to the mixed solve's own bar: FP64 normwise
backward error <= n * 2^-53 (double-double residual from the oracle), the low-precision
factorization alone a useful preconditioner, refinement iteration counts close to those of
partial pivoting, and the FP64 fallback on an exactly zero pivot. Sizes cover a single
workgroup (n <= 128), one merge level, two merge levels (more than 512/WS workgroups, n > 2048
for WS = 32 and n > 4096 for WS = 16), and a last panel and block narrower than their widths.
"""
import numpy as np
import pytest

from test_gpu_mixed import EPS, randsvd, say

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["calu16", "calu32"])
def calu(request, monkeypatch):
    monkeypatch.setenv("ELX_MIXED_PANEL", request.param)
    return request.param


def solve(dev, A, b, **kw):
    dA, dB = dev.to_device(A), dev.to_device(b)
    dX = dev.colmajor_empty(A.shape[0], b.shape[1])
    st = dev.gesv_mixed(dA, dB, dX, **kw)
    return st, dev.to_host(dX)


@pytest.mark.parametrize("n", [1, 2, 50, 129, 300, 1000, 2500, 5000])
def test_calu_converges(dev, orc, calu, n):
    say(calu, n)
    A = randsvd(n, 1e2, n + 11)
    b = np.random.default_rng(n + 12).standard_normal((n, 1))
    st, x = solve(dev, A, b, factor="fp16", refine="auto")
    eta = orc.backward_error(A, x, b)
    say(calu, n, st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["factor_info"] == 0 and st["converged"] == 1 and st["fell_back"] == 0
    assert eta <= n * EPS


@pytest.mark.parametrize("factor", ["fp16", "bf16", "fp32"])
def test_calu_factor_alone_is_a_preconditioner(dev, orc, calu, factor):
    n = 3000
    A = randsvd(n, 1e2, 17)
    b = np.random.default_rng(18).standard_normal((n, 1))
    st, x = solve(dev, A, b, factor=factor, refine="none")
    assert st["status"] in (0, 5) and st["factor_info"] == 0
    eta = orc.backward_error(A, x, b)
    u_low = {"fp16": 2.0 ** -11, "bf16": 2.0 ** -8, "fp32": 2.0 ** -22}[factor]
    say(calu, factor, "one-shot backward error", eta, "u_low", u_low)
    assert eta < 50 * u_low


def test_calu_iterations_close_to_partial_pivoting(dev, orc, monkeypatch):
    n = 4500
    A = randsvd(n, 1e2, 21)
    b = np.random.default_rng(22).standard_normal((n, 1))
    iters = {}
    for mode in ("partial", "calu16", "calu32"):
        monkeypatch.setenv("ELX_MIXED_PANEL", mode)
        st, x = solve(dev, A, b, factor="fp16", refine="ir")
        say(mode, st)
        assert st["converged"] == 1 and st["fell_back"] == 0
        assert orc.backward_error(A, x, b) <= n * EPS
        iters[mode] = st["outer_iters"]
    assert iters["calu16"] <= iters["partial"] + 3
    assert iters["calu32"] <= iters["partial"] + 3


@pytest.mark.parametrize("n", [300, 2500, 9000])
@pytest.mark.parametrize("kind", ["permutation", "lower_half_zero"])
def test_calu_structured(dev, orc, calu, n, kind):
    """Most leaves hold only zeros in a block's columns (a permutation matrix, and a matrix
    whose leading columns are nonzero only in the lower half of the rows): the tournament
    must still bring every nonzero row forward."""
    say(calu, kind, n)
    rng = np.random.default_rng(n + 5)
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
    st, x = solve(dev, A, b, factor="fp16", refine="auto")
    eta = orc.backward_error(A, x, b)
    say(st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["factor_info"] == 0 and st["converged"] == 1 and st["fell_back"] == 0
    assert eta <= n * EPS
    if kind == "permutation":
        assert st["outer_iters"] <= 1


def test_calu_zero_pivot_falls_back(dev, orc, calu):
    """A column that is exactly zero in fp32 stops the panel with info at that column; AUTO
    then solves with the FP64 LU."""
    n = 2600
    rng = np.random.default_rng(41)
    A = rng.standard_normal((n, n))
    A[:, 37] *= 1e-50
    A = np.asfortranarray(A)
    b = rng.standard_normal((n, 1))
    st, x = solve(dev, A, b, factor="fp16", refine="auto")
    say(calu, "breakdown", st)
    assert st["factor_info"] == 38 and st["fell_back"] == 1 and st["converged"] == 1
    assert orc.backward_error(A, x, b) <= n * EPS


@pytest.mark.parametrize("kappa", [1e4, 1e6])
def test_calu_ill_conditioned_auto(dev, orc, calu, kappa):
    n = 1800
    A = randsvd(n, kappa, 51)
    b = np.random.default_rng(52).standard_normal((n, 2))
    st, X = solve(dev, A, b, factor="fp16", refine="auto")
    eta = orc.backward_error(A, X, b)
    say(calu, kappa, st, "oracle eta/(n u) =", eta / (n * EPS))
    assert st["converged"] == 1 and eta <= n * EPS
