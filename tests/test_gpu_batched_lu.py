"""The device-side batched mixed LU-IR (elx_dgesv_mixed_batched with FP16 factors and AUTO
refinement): every system factored and refined by its own workgroup, the FP64 LU fallback per
system, no host round trip between the launches.

Each system must reach what El::LinearSolve gives in FP64 (src/lapack_like/solve/Linear.cpp:
212-218), measured as the normwise backward error ||B - A X|| / (||A|| ||X|| + ||B||) <= n u
per column, computed in double-double by the oracle. The batch spans both size bins (n <= 512
and n <= 2048, with the edges 1, 31, 32, 33, 512, 513, 2048), right-hand-side counts 0 .. 9
(chunks of 4), condition numbers 1e1 .. 1e8 (randsvd, geometric singular values), padded
leading dimensions, and an exactly singular matrix, which no tier can solve."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def say(*a):
    print("[case]", *a, flush=True)


def randsvd(n, kappa, rng):
    def haar():
        Q, R = np.linalg.qr(rng.standard_normal((n, n)))
        return Q * np.sign(np.diag(R))[None, :]
    U, V = haar(), haar()
    sigma = kappa ** (-np.arange(n) / max(n - 1, 1))
    return np.asfortranarray((U * sigma[None, :]) @ V.T)


def padded(a, extra):
    a = np.asarray(a)
    m, n = a.shape
    t = torch.zeros((n, m + extra), dtype=torch.float64, device="cuda").t()[:m, :]
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
    return t


def test_batched_lu_ir_device(dev, orc):
    rng = np.random.default_rng(2024)
    ns = [1, 31, 32, 33, 64, 100, 511, 512, 513, 700, 1024, 2048, 300, 90]
    kap = [1e1, 1e2, 1e3, 1e5, 1e8, 1e1, 1e4, 1e2, 1e6, 1e1, 1e3, 1e2, 1e7, 1e1]
    nrhs = [1, 4, 5, 0, 9, 2, 4, 1, 3, 4, 4, 2, 1, 4]
    As, Bs = [], []
    for n, k in zip(ns, kap):
        As.append(randsvd(n, k, rng) if n > 1 else np.array([[0.75]]))
    for n, r in zip(ns, nrhs):
        Bs.append(rng.standard_normal((n, r)))
    # an exactly singular system (a zero column): both tiers meet a zero pivot
    sing = len(ns)
    Z = rng.standard_normal((80, 80))
    Z[:, 17] = 0.0
    As.append(np.asfortranarray(Z))
    Bs.append(rng.standard_normal((80, 2)))
    ns.append(80)
    dA = [padded(A, 3) for A in As]
    dB = [padded(B, 2) for B in Bs]
    dX = [padded(np.zeros_like(B), 1) for B in Bs]
    say("batched lu-ir device", ns)
    status, stats = dev.gesv_mixed_batched(dA, dB, dX, factor="fp16", refine="auto")
    assert status == 5  # the singular system misses the tolerance
    for i, n in enumerate(ns):
        st = stats[i]
        if i == sing:
            say("singular", st)
            assert st["converged"] == 0 and st["fell_back"] == 1 and st["factor_info"] == 18
            continue
        if Bs[i].shape[1] == 0:
            assert st["converged"] == 1
            continue
        X = dev.to_host(dX[i])
        eta = orc.backward_error(As[i], np.asfortranarray(X), np.asfortranarray(Bs[i]))
        say(i, n, f"{kap[i]:.0e}", Bs[i].shape[1], st["fell_back"], st["outer_iters"],
            st["inner_iters"], eta / (max(n, 1) * EPS))
        assert st["converged"] == 1 and eta <= max(n, 1) * EPS, (i, n, kap[i], st, eta)
        # the device's own backward error is the DD one up to its rounding
        assert st["backward_error"] <= max(n, 1) * EPS
    # well-conditioned systems stay with the fp16 factors
    for i in (0, 5, 9, 13):
        assert stats[i]["fell_back"] == 0, (i, stats[i])
    # the padding of X past each order is untouched
    for i, t in enumerate(dX):
        full = t.as_strided((ns[i] + 1, t.shape[1]), (1, ns[i] + 1)).cpu().numpy()
        assert np.all(full[ns[i]:, :] == 0.0)


def test_batched_lu_ir_matches_one_at_a_time(dev, orc):
    """The same systems through the batched path and one by one through elx_dgesv_mixed:
    both meet n u; the batched path reports per-system statistics for every system."""
    rng = np.random.default_rng(99)
    ns = [int(v) for v in np.exp(rng.uniform(np.log(64), np.log(2048), 24)).astype(int)]
    kap = 10.0 ** rng.uniform(1, 8, len(ns))
    As = [randsvd(n, k, rng) for n, k in zip(ns, kap)]
    Bs = [rng.standard_normal((n, 4)) for n in ns]
    dA = [dev.to_device(A) for A in As]
    dB = [dev.to_device(B) for B in Bs]
    dX = [dev.colmajor_empty(n, 4) for n in ns]
    say("batched vs single", ns)
    status, stats = dev.gesv_mixed_batched(dA, dB, dX, factor="fp16", refine="auto")
    assert status == 0
    for i, n in enumerate(ns):
        X = dev.to_host(dX[i])
        eta = orc.backward_error(As[i], X, Bs[i])
        assert stats[i]["converged"] == 1 and eta <= n * EPS, (i, n, kap[i], stats[i], eta)
        if i < 4:
            Xs = dev.colmajor_empty(n, 4)
            st = dev.gesv_mixed(dA[i], dB[i], Xs, factor="fp16", refine="auto")
            eta1 = orc.backward_error(As[i], dev.to_host(Xs), Bs[i])
            assert st["converged"] == 1 and eta1 <= n * EPS
