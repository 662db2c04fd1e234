"""Writes tests/golden/golden.npz: known-answer vectors computed independently of the oracle.

* exact LU with partial pivoting of small integer matrices in rational arithmetic
  (fractions.Fraction): pivot sequence (Elemental's swap destinations: first row of
  largest magnitude, src/core/imports/blas/MaxInd.hpp:26-46, Panel.hpp:31-59), and L, U;
* 60-digit solutions (mpmath) of those systems;
* exact Cholesky factors of G G^T for integer lower-triangular G with power-of-two
  diagonal, where every floating-point step of the unblocked algorithm is exact;
* the first values of Elemental's deterministic Gaussian stream, recomputed from numpy's
  MT19937 (init_genrand seeding, as std::mt19937::seed) and libstdc++'s published
  normal_distribution algorithm (Marsaglia polar method over generate_canonical<double,53>,
  returning the second variate of each accepted pair and discarding the saved one, since
  Elemental builds a fresh distribution per sample: include/El/core/random/impl.hpp:185-203).

Run: python tests/golden/make_golden.py
"""
import math
import os
from fractions import Fraction

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
mpmath.mp.dps = 60


def exact_lu_partial_pivot(A):
    """Unblocked GEPP in exact arithmetic; ties go to the first (lowest) row."""
    n = len(A)
    M = [[Fraction(int(v)) for v in row] for row in A]
    dests = []
    for k in range(n):
        best, bi = Fraction(-1), -1
        for i in range(k, n):
            if abs(M[i][k]) > best:
                best, bi = abs(M[i][k]), i
        dests.append(bi)
        M[k], M[bi] = M[bi], M[k]
        if M[k][k] == 0:
            raise ValueError("singular")
        for i in range(k + 1, n):
            M[i][k] = M[i][k] / M[k][k]
            for j in range(k + 1, n):
                M[i][j] -= M[i][k] * M[k][j]
    return dests, M


def elemental_gaussian_stream(count, seed=(21 << 16)):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    raw = iter(bg.random_raw(count * 40).astype(np.uint64).tolist())

    def canonical():
        lo = next(raw)
        hi = next(raw)
        s = float(lo) + float(hi) * 4294967296.0  # exact: < 2^64 rounds once
        r = s / 18446744073709551616.0
        if r >= 1.0:
            r = math.nextafter(1.0, 0.0)
        return r

    out = []
    for _ in range(count):
        while True:
            x = 2.0 * canonical() - 1.0
            y = 2.0 * canonical() - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        out.append(y * mult)
    return np.array(out)


def main():
    rng = np.random.default_rng(20240601)
    data = {}
    # exact LU cases
    cases = []
    for n in (1, 2, 3, 4, 5, 6, 8):
        for rep in range(3):
            while True:
                A = rng.integers(-9, 10, size=(n, n))
                try:
                    d, M = exact_lu_partial_pivot(A.tolist())
                    break
                except ValueError:
                    continue
            cases.append((A, d, M))
    for idx, (A, d, M) in enumerate(cases):
        n = A.shape[0]
        F = np.array([[float(M[i][j]) for j in range(n)] for i in range(n)])
        b = rng.integers(-9, 10, size=n)
        Am = mpmath.matrix(A.tolist())
        x = mpmath.lu_solve(Am, mpmath.matrix(b.tolist()))
        data[f"lu{idx}_A"] = A.astype(np.float64)
        data[f"lu{idx}_dests"] = np.array(d, np.int32)
        data[f"lu{idx}_F"] = F
        data[f"lu{idx}_b"] = b.astype(np.float64)
        data[f"lu{idx}_x"] = np.array([float(v) for v in x])
    data["lu_count"] = np.array(len(cases))
    # exact Cholesky cases
    ch = []
    for n in (1, 2, 3, 5, 8, 12):
        G = np.tril(rng.integers(-3, 4, size=(n, n))).astype(np.float64)
        G[np.diag_indices(n)] = 2.0 ** rng.integers(0, 3, size=n)
        ch.append(G)
    for idx, G in enumerate(ch):
        data[f"chol{idx}_G"] = G
        data[f"chol{idx}_A"] = G @ G.T
    data["chol_count"] = np.array(len(ch))
    data["gaussian_first32"] = elemental_gaussian_stream(32)
    np.savez(os.path.join(HERE, "golden.npz"), **data)
    print("wrote", len(cases), "LU cases,", len(ch), "Cholesky cases")


if __name__ == "__main__":
    main()
