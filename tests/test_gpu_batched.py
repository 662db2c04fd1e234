"""Batched variable-size Cholesky (lower) on the GPU against the oracle.

Every matrix of the batch must get what El::Cholesky(LOWER, A_i) and
cholesky::SolveAfter(LOWER, NORMAL, L_i, B_i) give one matrix at a time
(src/lapack_like/factor/Cholesky/LowerVariant3.hpp:15-68, Cholesky/SolveAfter.hpp:15-43):
factor within 8 n u of A in the normwise sense (u of the arithmetic type), the same
non-HPD column (info = j+1) for a matrix that is not positive definite, and FP64 solves to
backward error <= n u. The batch mixes every block-boundary size around 16 and 32 with
random sizes, in an order unrelated to size."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS64 = 2.0 ** -53
EPS32 = 2.0 ** -24


def say(*a):
    print("[case]", *a, flush=True)


def spd(n, rng):
    G = rng.standard_normal((n, n))
    return np.asfortranarray(G @ G.T / max(n, 1) + np.eye(n))


def sizes(rng):
    edge = [1, 2, 15, 16, 17, 31, 32, 33, 47, 63, 64, 65, 96, 127, 128, 129, 255, 256, 257,
            383, 511, 512]
    rand = list(rng.integers(16, 513, size=26))
    s = edge + rand
    rng.shuffle(s)
    return [int(v) for v in s]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_cholesky_factor_and_solve(dev, orc, dtype):
    rng = np.random.default_rng(41)
    ns = sizes(rng)
    say("batched cholesky", dtype, len(ns), "matrices, n in", min(ns), max(ns))
    As = [spd(n, rng) for n in ns]
    # one matrix that is not positive definite at global column 40 (0-based): its diagonal
    # there is made very negative
    bad = ns.index(max(n for n in ns if n > 64))
    As[bad][40, 40] = -1e3
    mats = [dev.to_device(A, dtype=dtype) for A in As]
    batch = dev.Batch(mats, dtype)
    info = dev.potrf_vbatched(batch).cpu().numpy()
    torch.cuda.synchronize()
    u = EPS64 if dtype == torch.float64 else EPS32
    for i, (A, n) in enumerate(zip(As, ns)):
        F = dev.to_host(mats[i]).astype(np.float64)
        Fo, info_o = orc.cholesky(A, uplo=0)
        assert info[i] == info_o, (i, n, info[i], info_o)
        if info_o:
            assert info_o == 41
            continue
        err = orc.chol_factor_error(A, np.asfortranarray(F), 0)
        assert err <= 8 * n * u, (i, n, err)
        if dtype == torch.float64:
            # same algorithm up to the blocking: entrywise close to the oracle's factor
            L, Lo = np.tril(F), np.tril(Fo)
            assert np.abs(L - Lo).max() <= 64 * n * u * np.abs(Lo).max(), (i, n)
    # solves, for the HPD matrices: 1..8 right-hand sides, one with 11 (two column chunks)
    keep = [i for i in range(len(ns)) if info[i] == 0]
    nr = [int(rng.integers(1, 9)) for _ in keep]
    nr[0] = 11
    Bs = [rng.standard_normal((ns[i], r)) for i, r in zip(keep, nr)]
    fac = dev.Batch([mats[i] for i in keep], dtype)
    rhs = dev.Batch([dev.to_device(B, dtype=dtype) for B in Bs], dtype)
    dev.potrs_vbatched(fac, rhs)
    torch.cuda.synchronize()
    for j, i in enumerate(keep):
        X = dev.to_host(rhs.mats[j]).astype(np.float64)
        eta = orc.backward_error(As[i], np.asfortranarray(X), np.asfortranarray(Bs[j]))
        n = ns[i]
        # n u, with room for the two roundings (two divisions) of the smallest orders
        bound = max(n, 4) * u if dtype == torch.float64 else 4 * max(n, 4) * u
        assert eta <= bound, (i, n, nr[j], eta / (n * u))


def test_batched_cholesky_known_answers(dev, golden):
    """Exact integer Cholesky factors (every floating-point step exact) come back exactly,
    in a batch with other sizes around them."""
    rng = np.random.default_rng(5)
    mats, expect = [], []
    for idx in range(int(golden["chol_count"])):
        mats.append(dev.to_device(golden[f"chol{idx}_A"]))
        expect.append(golden[f"chol{idx}_G"])
        A = spd(int(rng.integers(20, 90)), rng)
        mats.append(dev.to_device(A))
        expect.append(None)
    batch = dev.Batch(mats, torch.float64)
    info = dev.potrf_vbatched(batch).cpu().numpy()
    assert (info == 0).all()
    for t, G in zip(mats, expect):
        if G is not None:
            assert np.array_equal(np.tril(dev.to_host(t)), G)


def test_batched_empty_and_zero_sizes(dev):
    batch = dev.Batch([], torch.float64)
    dev.potrf_vbatched(batch)
    mats = [dev.colmajor_empty(0, 0), dev.to_device(np.eye(3) * 4.0)]
    info = dev.potrf_vbatched(dev.Batch(mats, torch.float64)).cpu().numpy()
    assert (info == 0).all() and np.array_equal(dev.to_host(mats[1]), np.eye(3) * 2.0)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("nb", [16, 32])
def test_batched_cholesky_every_variant(dev, orc, nb, kind):
    """Each size bin's kernel with both block widths (the autotuner's candidates), one and
    two or three workgroups per CU, both kernel kinds (the register-direct one in bins 1 and 2):
    same bars as above, and the non-HPD column exact."""
    rng = np.random.default_rng(43 + nb + kind)
    ns = [1, 15, 16, 17, 33, 100, 128, 129, 200, 256, 257, 300, 511, 512]
    As = [spd(n, rng) for n in ns]
    As[9][70, 70] = -1e3  # n = 256: not positive definite at column 70
    try:
        for occ in (1, 2):
            for b in range(3):
                dev.batched_variant(torch.float64, b, nb, min(occ, 3 - b) if not kind else occ,
                                    kind if b >= 1 else 0)
            say("batched variant", nb, occ, kind)
            mats = [dev.to_device(A) for A in As]
            info = dev.potrf_vbatched(dev.Batch(mats, torch.float64)).cpu().numpy()
            for i, (A, n) in enumerate(zip(As, ns)):
                Fo, info_o = orc.cholesky(A, uplo=0)
                assert info[i] == info_o, (i, n, info[i], info_o)
                if info_o:
                    assert info_o == 71
                    continue
                F = dev.to_host(mats[i])
                assert orc.chol_factor_error(A, np.asfortranarray(F), 0) <= 8 * n * EPS64
                L, Lo = np.tril(F), np.tril(Fo)
                assert np.abs(L - Lo).max() <= 64 * n * EPS64 * np.abs(Lo).max(), (i, n)
    finally:
        for b in range(3):
            dev.batched_variant(torch.float64, b, 0)
    assert dev.batched_variant(torch.float64, 0)[0] in (16, 32)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_wide_panels(dev, orc, dtype):
    """The wide-panel kind (64-column panels, bins 1 and 2) with one and two workgroups per CU:
    sizes around the panel width and the 256-row pass, one matrix not positive definite in its
    first panel and one in its fourth; same bars as above."""
    rng = np.random.default_rng(47)
    ns = [129, 150, 191, 192, 193, 255, 256, 257, 300, 319, 320, 321, 383, 384, 385, 448, 500,
          511, 512]
    As = [spd(n, rng) for n in ns]
    As[2][40, 40] = -1e3    # n = 191: column 40, first panel
    As[17][200, 200] = -1e3  # n = 511: column 200, fourth panel
    u = EPS64 if dtype == torch.float64 else EPS32
    try:
        for occ in (1, 2):
            for b in (1, 2):
                dev.batched_variant(dtype, b, 64, occ, 2)
            say("wide panels", dtype, occ)
            # every third matrix a view with 9 padding rows below it (leading dimension n + 9),
            # the padding holding 7.0: it must come back untouched
            mats, pads = [], []
            for i, A in enumerate(As):
                if i % 3 == 0:
                    n = A.shape[0]
                    big = dev.colmajor_empty(n + 9, n, dtype=dtype)
                    big.fill_(7.0)
                    big[:n].copy_(dev.to_device(A, dtype=dtype))
                    mats.append(big[:n])
                    pads.append(big)
                else:
                    mats.append(dev.to_device(A, dtype=dtype))
                    pads.append(None)
            info = dev.potrf_vbatched(dev.Batch(mats, dtype)).cpu().numpy()
            for big, A in zip(pads, As):
                if big is not None:
                    assert (dev.to_host(big)[A.shape[0]:] == 7.0).all()
            for i, (A, n) in enumerate(zip(As, ns)):
                Fo, info_o = orc.cholesky(A, uplo=0)
                assert info[i] == info_o, (i, n, info[i], info_o)
                if info_o:
                    assert info_o in (41, 201)
                    continue
                F = dev.to_host(mats[i]).astype(np.float64)
                assert orc.chol_factor_error(A, np.asfortranarray(F), 0) <= 8 * n * u, (i, n)
                # the strict upper triangle is never written
                Au = A if dtype == torch.float64 else A.astype(np.float32).astype(np.float64)
                assert np.array_equal(np.triu(F, 1), np.triu(Au, 1)), (i, n)
                if dtype == torch.float64:
                    L, Lo = np.tril(F), np.tril(Fo)
                    assert np.abs(L - Lo).max() <= 64 * n * u * np.abs(Lo).max(), (i, n)
    finally:
        for b in range(3):
            dev.batched_variant(dtype, b, 0)


@pytest.mark.parametrize("nb", [16, 32])
def test_batched_register_direct_float32(dev, orc, nb):
    """The register-direct kernel kind in FP32 (bins 1 and 2, one and two workgroups per CU):
    factor error within 8 n u32 of the oracle's bar, the non-HPD column exact."""
    rng = np.random.default_rng(61 + nb)
    ns = [129, 160, 200, 255, 256, 257, 300, 384, 511, 512]
    As = [spd(n, rng) for n in ns]
    As[6][90, 90] = -1e3  # n = 300: not positive definite at column 90
    try:
        for occ in (1, 2):
            for b in (1, 2):
                dev.batched_variant(torch.float32, b, nb, occ, 1)
            say("batched fp32 register-direct", nb, occ)
            mats = [dev.to_device(A, dtype=torch.float32) for A in As]
            info = dev.potrf_vbatched(dev.Batch(mats, torch.float32)).cpu().numpy()
            for i, (A, n) in enumerate(zip(As, ns)):
                _, info_o = orc.cholesky(A, uplo=0)
                assert info[i] == info_o, (i, n, info[i], info_o)
                if info_o:
                    assert info_o == 91
                    continue
                F = dev.to_host(mats[i]).astype(np.float64)
                err = orc.chol_factor_error(A, np.asfortranarray(F), 0)
                assert err <= 8 * n * EPS32, (i, n, err)
    finally:
        for b in range(3):
            dev.batched_variant(torch.float32, b, 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_graph_replays(dev, orc, dtype):
    """The captured factor + solve, replayed three times with the inputs restored by eager
    copies in between, equals the eager potrf_vbatched + potrs_vbatched bit for bit on every
    replay; a matrix that is not HPD keeps its right-hand side unsolved."""
    rng = np.random.default_rng(47)
    ns = [int(v) for v in rng.integers(16, 513, size=40)] + [3, 64, 128, 256, 512]
    As = [spd(n, rng) for n in ns]
    bad = 5
    As[bad][min(10, ns[bad] - 1), min(10, ns[bad] - 1)] = -1e3
    nr = [int(v) for v in rng.integers(1, 9, size=len(ns))]
    Bs = [rng.standard_normal((n, r)) for n, r in zip(ns, nr)]
    A0 = [dev.to_device(A, dtype=dtype) for A in As]
    B0 = [dev.to_device(B, dtype=dtype) for B in Bs]
    # eager reference
    Ae = [dev.to_device(dev.to_host(a), dtype=dtype) for a in A0]
    Be = [dev.to_device(dev.to_host(b), dtype=dtype) for b in B0]
    fac_e = dev.Batch(Ae, dtype)
    info_e = dev.potrf_vbatched(fac_e).cpu().numpy()
    keep = [i for i in range(len(ns)) if info_e[i] == 0]
    assert bad not in keep
    dev.potrs_vbatched(dev.Batch([Ae[i] for i in keep], dtype),
                       dev.Batch([Be[i] for i in keep], dtype))
    torch.cuda.synchronize()
    # graph on its own copies
    Ag = [dev.to_device(dev.to_host(a), dtype=dtype) for a in A0]
    Bg = [dev.to_device(dev.to_host(b), dtype=dtype) for b in B0]
    g = dev.PosvGraph(dev.Batch(Ag, dtype), dev.Batch(Bg, dtype))
    try:
        for rep in range(3):
            say("graph replay", dtype, rep)
            for a, a0 in zip(Ag, A0):
                a.copy_(a0)
            for b, b0 in zip(Bg, B0):
                b.copy_(b0)
            torch.cuda.synchronize()
            info = g.launch().cpu().numpy()
            torch.cuda.synchronize()
            assert np.array_equal(info, info_e)
            for i in range(len(ns)):
                assert np.array_equal(dev.to_host(Ag[i]), dev.to_host(Ae[i])), (rep, i)
                assert np.array_equal(dev.to_host(Bg[i]), dev.to_host(Be[i])), (rep, i)
            assert np.array_equal(dev.to_host(Bg[bad]), dev.to_host(B0[bad]))
            # eager work of the program between replays
            dev.potrf_vbatched(dev.Batch([dev.to_device(spd(40, rng), dtype=dtype)], dtype))
    finally:
        g.close()
    for j, i in enumerate(keep[:10]):
        X = dev.to_host(Bg[i]).astype(np.float64)
        eta = orc.backward_error(As[i], np.asfortranarray(X), np.asfortranarray(Bs[i]))
        u = EPS64 if dtype == torch.float64 else EPS32
        assert eta <= 4 * max(ns[i], 4) * u
