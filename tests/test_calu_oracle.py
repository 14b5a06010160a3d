"""CPU: the oracle's LU with tournament pivoting (oracle/calu_oracle.py) pinned against scipy's
LAPACK partial pivoting and the identity P A = L U."""
import numpy as np
import scipy.linalg as sla

from oracle import calu_oracle as C


def _rand(rng, m, n):
    return rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))


def _apply(ipiv, n):
    perm = np.arange(n)
    for i, p in enumerate(ipiv):
        perm[[i, p]] = perm[[p, i]]
    return perm


def _lapack_rows(P):
    """Rows chosen by zgetrf on the panel P (original indices, in pivot order)."""
    _, piv = sla.lu_factor(P)
    return list(_apply(piv, P.shape[0])[:P.shape[1]])


def test_single_block_tournament_is_partial_pivoting():
    rng = np.random.default_rng(1)
    P = _rand(rng, 300, 32)
    assert C.tournament(P, 512) == _lapack_rows(P)


def test_two_round_tournament_picks_lapack_pivots_of_the_block_winners():
    rng = np.random.default_rng(2)
    m, w, cp = 1000, 16, 128
    P = _rand(rng, m, w)
    wins = []
    for s in range(0, m, cp):
        blk = P[s:s + cp]
        wins += [s + r for r in _lapack_rows(blk)]
    # 8 blocks x 16 winners = 128 candidates: one second round
    want = [wins[r] for r in _lapack_rows(P[wins])]
    assert C.tournament(P, cp) == want
    # with j0: global indices shift
    assert C.tournament(P, cp, j0=7) == [r + 7 for r in want]


def test_ipiv_from_rows_moves_rows_in_order():
    rng = np.random.default_rng(3)
    n, j0 = 60, 10
    for _ in range(20):
        rows = list(rng.choice(np.arange(j0, n), size=16, replace=False))
        ipiv = C.ipiv_from_rows(rows, j0)
        perm = np.arange(n)
        for c, p in enumerate(ipiv):
            assert p >= j0 + c
            perm[[j0 + c, p]] = perm[[p, j0 + c]]
        assert list(perm[j0:j0 + 16]) == rows


def test_getrf_without_tall_panels_is_lapack():
    rng = np.random.default_rng(4)
    A = _rand(rng, 150, 150)
    lu, piv = C.getrf(A, nb=32, tall=None)
    lu2, piv2 = sla.lu_factor(A)
    assert np.array_equal(piv, piv2)
    assert np.max(np.abs(lu - lu2)) < 1e-12 * np.max(np.abs(lu2))


def test_getrf_with_tall_panels_factors_pa():
    rng = np.random.default_rng(5)
    n = 200
    A = _rand(rng, n, n)
    lu, piv = C.getrf(A, nb=16, tall=64, cp=32)
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    perm = _apply(piv, n)
    assert np.max(np.abs(A[perm] - L @ U)) < 1e-12 * np.max(np.abs(A))
    # a tall panel's pivots differ from partial pivoting's here, and |L| stays bounded
    assert not np.array_equal(piv, sla.lu_factor(A)[1])
    assert np.max(np.abs(L)) < 10
    # with every panel below the threshold: LAPACK
    assert np.array_equal(C.getrf(A, nb=16, tall=n + 1, cp=32)[1], sla.lu_factor(A)[1])
