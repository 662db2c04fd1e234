"""The oracle's Ruiz equilibration (El::RuizEquil, src/lapack_like/equilibrate/RuizEquil.cpp:
23-60) pinned by closed forms and invariants that do not restate its loop."""
import numpy as np

from oracle import oracle as orc

TOL = (2.0 ** -53) ** 0.33


def test_rank_one_powers_of_two():
    # A = u v^T: the first column sweep leaves u_i / max|u| * sign(v_j), the row sweep then
    # the sign matrix; later sweeps see unit norms. Every scale is a power of two: exact.
    u = np.array([2.0 ** 3, -2.0 ** -5, 2.0 ** 10, 2.0 ** -2, -1.0])
    v = np.array([-2.0 ** 4, 2.0 ** -7, 2.0 ** 2])
    A = np.outer(u, v)
    Ae, dR, dC = orc.ruiz_equil(A)
    assert np.array_equal(Ae, np.sign(np.outer(u, v)))
    assert np.array_equal(dR, np.abs(u) / np.abs(u).max())
    assert np.array_equal(dC, np.abs(v) * np.abs(u).max())


def test_zero_rows_and_columns_keep_unit_scales():
    A = np.array([[0.0, 2.0, 0.0], [0.0, 0.0, 0.0], [0.0, -8.0, 0.0]])
    Ae, dR, dC = orc.ruiz_equil(A)
    assert np.array_equal(dC, [1.0, 8.0, 1.0])
    assert np.array_equal(dR, [0.25, 1.0, 1.0])
    assert np.array_equal(Ae, [[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, -1.0, 0.0]])


def test_damping_floor():
    # a column norm below eps^0.33 is raised to it (DampScaling)
    Ae, dR, dC = orc.ruiz_equil(np.array([[1e-10], [0.0]]))
    assert dC[0] == TOL
    assert dR[1] == 1.0 and Ae[1, 0] == 0.0
    assert abs(dR[0] * Ae[0, 0] * dC[0] - 1e-10) <= 1e-25


def test_reconstruction_and_unit_row_norms():
    rng = np.random.default_rng(3)
    m, n = 60, 45
    A = (2.0 ** rng.uniform(-30, 30, m))[:, None] * rng.standard_normal((m, n)) * \
        (2.0 ** rng.uniform(-30, 30, n))[None, :]
    A[:, 7] = 0.0
    Ae, dR, dC = orc.ruiz_equil(A)
    R = dR[:, None] * Ae * dC[None, :]
    assert np.max(np.abs(R - A) / np.maximum(np.abs(A), 1e-300)) <= 8 * 2.0 ** -53
    # the last step divides every row by its max norm
    assert np.allclose(np.abs(Ae).max(axis=1), 1.0, rtol=0, atol=4 * 2.0 ** -53)
    assert np.abs(Ae).max() <= 1.0 + 4 * 2.0 ** -53
    # scaling balances the matrix: column norms near 1 as well
    cm = np.abs(Ae).max(axis=0)
    assert np.all(cm[np.arange(n) != 7] > 0.5) and cm[7] == 0.0 and dC[7] == 1.0


def test_symmetric_diagonal_powers_of_four():
    # a diagonal of powers of four: the first sweep's scales are exact square roots and leave
    # the identity; later sweeps see unit norms
    dg = np.array([4.0 ** 5, 4.0 ** -3, 1.0, 4.0 ** 9])
    Ae, d = orc.symmetric_ruiz_equil(np.diag(dg))
    assert np.array_equal(Ae, np.eye(4))
    assert np.array_equal(d, np.sqrt(dg))


def test_symmetric_rank_one_powers_of_four():
    # A = u u^T with u a signed power of 16 per entry: sweep 1 scales column j by
    # sqrt(|u_j| max|u|), sweep 2 by the square root of the current column norm; for these u
    # both sweeps' scales are powers of two, so the result is exact: A_eq(i,j) =
    # sign(u_i u_j) r_i r_j with r = sqrt(|u| / max|u|) after sweep 1 and r -> r^(1/2) after 2
    u = np.array([16.0 ** 4, -16.0 ** -2, 16.0 ** 2, 1.0])
    for iters in (1, 2):
        Ae, d = orc.symmetric_ruiz_equil(np.outer(u, u), maxiter=iters)
        r = np.sqrt(np.abs(u) / np.abs(u).max())
        for _ in range(iters - 1):
            r = np.sqrt(r)
        assert np.array_equal(Ae, np.sign(np.outer(u, u)) * np.outer(r, r))
        assert np.array_equal(d[:, None] * Ae * d[None, :], np.outer(u, u))


def test_symmetric_reconstruction():
    rng = np.random.default_rng(5)
    n = 50
    G = rng.standard_normal((n, n))
    D = 2.0 ** rng.uniform(-25, 25, n)
    A = D[:, None] * (G @ G.T / n + np.eye(n)) * D[None, :]
    Ae, d = orc.symmetric_ruiz_equil(A)
    R = d[:, None] * Ae * d[None, :]
    assert np.max(np.abs(R - A) / np.abs(A)) <= 8 * 2.0 ** -53
    # each sweep takes the square root of the column-norm imbalance: after three the spread
    # (in bits) is at most a quarter of the original, and no norm exceeds 1
    cm0, cm = np.abs(A).max(axis=0), np.abs(Ae).max(axis=0)
    assert cm.max() <= 1.0 + 1e-12
    assert np.log2(cm.max() / cm.min()) <= np.log2(cm0.max() / cm0.min()) / 4


def test_geom_rank_one_closed_form():
    # A = u v^T with the spread of u below 1000 (no damping): the column pass divides column j
    # by sqrt(min|u| max|u|) |v_j|, the row pass then leaves the sign matrix; later sweeps and
    # the closing column pass see unit entries. All scales are powers of two here: exact.
    u = np.array([4.0, -0.25, 16.0, 1.0])
    v = np.array([-2.0 ** 7, 2.0 ** -9, 8.0])
    Ae, dR, dC = orc.geom_equil(np.outer(u, v))
    g = 2.0  # sqrt(0.25 * 16)
    assert np.array_equal(Ae, np.sign(np.outer(u, v)))
    assert np.array_equal(dR, np.abs(u) / g)
    assert np.array_equal(dC, np.abs(v) * g)


def test_geom_damping_and_zero_matrix():
    # one column [1, 1e-6]: sqrt(mn*mx) = 1e-3 is below sqrt(1/1000)*mx, so the damped scale
    # is used; the row pass then brings both entries to magnitude 1
    Ae, dR, dC = orc.geom_equil(np.array([[1.0], [1e-6]]))
    assert dC[0] == np.sqrt(1.0 / 1000.0)
    assert np.allclose(np.abs(Ae[:, 0]), 1.0, rtol=0, atol=4 * 2.0 ** -53)
    assert np.allclose(dR * Ae[:, 0] * dC[0], [1.0, 1e-6], rtol=4 * 2.0 ** -53, atol=0)
    Z = np.zeros((3, 2))
    Ae, dR, dC = orc.geom_equil(Z)
    assert np.array_equal(Ae, Z) and np.array_equal(dR, np.ones(3)) and np.array_equal(dC, np.ones(2))


def test_geom_reconstruction_and_unit_columns():
    rng = np.random.default_rng(9)
    m, n = 40, 30
    A = (2.0 ** rng.uniform(-30, 30, m))[:, None] * rng.standard_normal((m, n)) * \
        (2.0 ** rng.uniform(-30, 30, n))[None, :]
    A[:, 4] = 0.0
    A[7, :] = 0.0
    Ae, dR, dC = orc.geom_equil(A)
    R = dR[:, None] * Ae * dC[None, :]
    nzm = A != 0
    assert np.max(np.abs(R - A)[nzm] / np.abs(A)[nzm]) <= 32 * 2.0 ** -53
    assert np.all(Ae[~nzm] == 0)
    cm = np.abs(Ae).max(axis=0)
    assert np.allclose(cm[np.arange(n) != 4], 1.0, rtol=0, atol=2 * 2.0 ** -53) and cm[4] == 0
    # the spread of the entries shrinks by orders of magnitude
    ab = np.abs(A[nzm]); abe = np.abs(Ae[nzm])
    assert np.log2(abe.max() / abe.min()) < 0.5 * np.log2(ab.max() / ab.min())
