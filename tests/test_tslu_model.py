"""The tournament-pivoting model that fixes the device LU's pivot sequence:
pinned against LAPACK where the two coincide (one leaf) and against forced
winners and a reconstruction where they do not."""
import numpy as np
import scipy.linalg as sla

import tslu_model as tm


def _permute(A, ipiv):
    PA = A.copy()
    for i, p in enumerate(ipiv):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return PA


def test_single_leaf_is_lapack_partial_pivoting():
    rng = np.random.default_rng(3)
    for n in (7, 64, 100, 128, 200, 256):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        assert np.array_equal(tm.blocked_lu_pivots(A), sla.lu_factor(A)[1])


def test_ties_go_to_the_lowest_position():
    # every candidate of column 0 has |re|+|im| = 2: LAPACK takes position 0
    M = np.array([[1 + 1j, 5], [2, 1], [-2j, 3], [1 - 1j, 2]], dtype=np.complex128)
    assert tm.gepp_select(M)[0] == 0
    assert sla.lu_factor(np.hstack([M, np.eye(4, 2)]))[1][0] == 0


def test_tournament_finds_forced_winners_across_leaves():
    # one dominant entry per column, spread over the four 256-row leaves
    rng = np.random.default_rng(5)
    m, w = 1000, 6
    P = 1e-3 * (rng.standard_normal((m, w)) + 1j * rng.standard_normal((m, w)))
    rows = [905, 17, 611, 250, 999, 388]
    for j, r in enumerate(rows):
        P[r, j] = 100.0 * (j + 1)
    assert tm.tournament(P) == rows


def test_tournament_forced_winners_through_short_groups():
    # 17 leaves: the last leaf is alone in its group of four at level 0 and
    # its node again at level 1 (17 -> 5 -> 2 -> 1 nodes)
    rng = np.random.default_rng(6)
    m, w = 4100, 5
    P = 1e-3 * (rng.standard_normal((m, w)) + 1j * rng.standard_normal((m, w)))
    rows = [4099, 3, 2048, 4096, 1023]
    for j, r in enumerate(rows):
        P[r, j] = 100.0 * (j + 1)
    assert tm.tournament(P) == rows


def test_blocked_factorization_reconstructs():
    rng = np.random.default_rng(9)
    n = 300
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    ipiv, LU = tm.blocked_lu_pivots(A, return_lu=True)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.abs(_permute(A, ipiv) - L @ U).max() < 1e-11 * np.abs(A).max()
    # multi-leaf panels: the tournament may choose other rows than GEPP
    assert np.abs(L).max() < 10
