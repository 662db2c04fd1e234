"""elx_dgetrs_noalloc (lu::SolveAfter without allocation: its row-permutation work space is a
fixed device array per GPU) against elx_dgetrs on the same factors and right-hand sides,
through the C ABI. Two calls in a row, the second on a smaller system, so the second reads a
scratch array the first has filled; every case is also checked against the oracle's
lu::SolveAfter (src/lapack_like/factor/LU/SolveAfter.hpp:66-92) in backward error."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53
SCRATCH_BYTES = 32 << 20


def say(*a):
    print("[case]", *a, flush=True)


def factor(dev, orc, n, seed):
    orc.rng_seed(seed)
    A = orc.gaussian(n, n)
    dA = dev.to_device(A)
    d, info = dev.getrf(dA)
    assert int(info.item()) == 0
    return A, dA, d


def solve_both(dev, dA, d, Bh, orient):
    """(elx_dgetrs result, elx_dgetrs_noalloc result) on copies of Bh."""
    from elemental_amd._lib import call
    n, nrhs = Bh.shape
    ref = dev.to_device(Bh)
    dev.getrs(dA, d, ref, orient=orient)
    got = dev.to_device(Bh)
    pA, lda = dA.data_ptr(), dA.stride(1)
    call("elx_dgetrs_noalloc", ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), orient,
         n, nrhs, pA, lda, d.data_ptr(), got.data_ptr(), got.stride(1))
    torch.cuda.synchronize()
    return dev.to_host(ref), dev.to_host(got)


@pytest.mark.parametrize("orient", [0, 1])
def test_noalloc_matches_getrs_two_calls(dev, orc, orient):
    """Products below 2^24 run on the hand-written GEMM in both entry points: bit for bit.
    The second call is smaller than the first."""
    for (n, nrhs, seed) in [(1000, 8, 41), (300, 3, 42)]:
        say("noalloc vs getrs", n, nrhs, orient)
        A, dA, d = factor(dev, orc, n, seed)
        B = orc.gaussian(n, nrhs)
        ref, got = solve_both(dev, dA, d, B, orient)
        assert np.array_equal(ref, got), "no-alloc solve must equal elx_dgetrs bit for bit"
        At = A if orient == 0 else A.T
        for j in range(nrhs):
            eta = orc.backward_error(np.asfortranarray(At), got[:, j:j + 1], B[:, j:j + 1])
            assert eta <= n * EPS, eta


def test_noalloc_large_and_limits(dev, orc):
    """A system whose updates elx_dgetrs hands to the library GEMM (4096 x 64 right-hand
    sides): equal to FP64 rounding; then a smaller one again; a size beyond the fixed scratch
    is rejected before any device work; no pivots needs no scratch at all."""
    from elemental_amd._lib import ElxError, call
    n, nrhs = 4096, 64
    say("noalloc large", n, nrhs)
    A, dA, d = factor(dev, orc, n, 43)
    rng = np.random.default_rng(7)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    ref, got = solve_both(dev, dA, d, B, 0)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-12 * scale
    say("noalloc after large", 200, 2)
    A2, dA2, d2 = factor(dev, orc, 200, 44)
    B2 = orc.gaussian(200, 2)
    ref2, got2 = solve_both(dev, dA2, d2, B2, 0)
    assert np.array_equal(ref2, got2)
    # beyond the limit: 4n rounded to 256 plus 8 n nrhs > 32 MiB
    big_n, big_rhs = 32768, 128
    assert 8 * big_n * big_rhs + 4 * big_n > SCRATCH_BYTES
    say("noalloc limit", big_n, big_rhs)
    with pytest.raises(ElxError) as e:
        call("elx_dgetrs_noalloc", ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), 0,
             big_n, big_rhs, dA.data_ptr(), big_n, d.data_ptr(), dA.data_ptr(), big_n)
    assert e.value.status == 1
    # the largest accepted system for n = 32768 is 127 right-hand sides
    assert (32768 * 4 + 255) // 256 * 256 + 8 * 32768 * 127 <= SCRATCH_BYTES
    # no permutation (the factors of the unpivoted LU): no scratch involved
    say("noalloc nopiv", 256, 4)
    orc.rng_seed(45)
    A3 = orc.gaussian(256, 256) + 256 * np.eye(256)
    dA3 = dev.to_device(np.asfortranarray(A3))
    info = dev.getrf_nopiv(dA3)
    assert int(info.item()) == 0
    B3 = orc.gaussian(256, 4)
    ref3 = dev.to_device(B3)
    dev.getrs(dA3, None, ref3)
    got3 = dev.to_device(B3)
    call("elx_dgetrs_noalloc", ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), 0, 256,
         4, dA3.data_ptr(), dA3.stride(1), None, got3.data_ptr(), got3.stride(1))
    torch.cuda.synchronize()
    assert np.array_equal(dev.to_host(ref3), dev.to_host(got3))
