"""CPU model of the device LU's pivot choice: right-looking blocked LU with
64-column panels whose pivot rows come from tournament pivoting (CALU/TSLU,
Grigori, Demmel and Xiang 2011) over 256-row leaves and a merge tree that
joins four nodes' nominations at a time (fewer when a level's last group is
short; a lone node is re-run by itself, which keeps its order).

Test infrastructure only: it fixes the pivot sequence the device must produce,
so the GPU test can compare LAPACK-style ipiv arrays index for index.  The
reference solves its dense systems with LAPACK's partially pivoted LU
(Julia's `\\` on a dense complex matrix); for one leaf (n <= 256) the
tournament is exactly that, and the GPU test checks it against scipy.
"""
import numpy as np

NB = 64
LEAF = 256
ARITY = 4


def _cabs1(z):
    return np.abs(z.real) + np.abs(z.imag)


def gepp_select(M):
    """Partial pivoting on the rows of M (in the given order), max |re|+|im|,
    ties to the lowest current position (LAPACK i?amax); returns the selected
    row indices of M in pivot order."""
    M = np.array(M, dtype=np.complex128)
    nv, w = M.shape
    perm = list(range(nv))
    for j in range(min(nv, w)):
        keys = _cabs1(M[perm[j:], j])
        keys = np.where(np.isnan(keys), -0.5, keys)
        p = j + int(np.argmax(keys))
        perm[j], perm[p] = perm[p], perm[j]
        pv = M[perm[j], j]
        if pv != 0:
            rows = perm[j + 1:]
            l = M[rows, j] / pv
            M[np.ix_(rows, range(j + 1, w))] -= np.outer(l, M[perm[j], j + 1:])
    return perm[:min(nv, w)]


def tournament(P):
    """Pivot rows (indices into P's rows) chosen by the tournament over the
    m x w panel P, in pivot order."""
    m = P.shape[0]
    nodes = [list(range(i, min(i + LEAF, m))) for i in range(0, m, LEAF)]
    nodes = [[c[s] for s in gepp_select(P[c])] for c in nodes]
    while len(nodes) > 1:
        nxt = []
        for i in range(0, len(nodes), ARITY):
            cand = [r for nd in nodes[i:i + ARITY] for r in nd]
            nxt.append([cand[s] for s in gepp_select(P[cand])])
        nodes = nxt
    return nodes[0]


def blocked_lu_pivots(A, return_lu=False):
    """LAPACK-convention ipiv (0-based) of the blocked LU with tournament
    pivoting, computed in FP64 (and the packed L\\U factors if asked)."""
    A = np.array(A, dtype=np.complex128)
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int64)
    for k0 in range(0, n, NB):
        w = min(NB, n - k0)
        sel = [k0 + r for r in tournament(A[k0:, k0:k0 + w])]
        # sequential interchanges: row order after each swap
        order = list(range(n))
        where = {r: r for r in range(k0, n)}
        for s, t in enumerate(sel):
            pt = where[t]
            ipiv[k0 + s] = pt
            u = order[k0 + s]
            order[k0 + s], order[pt] = t, u
            where[t], where[u] = k0 + s, pt
        A[k0:] = A[order[k0:]]
        # LU of the diagonal block without pivoting, then the block updates
        D = A[k0:k0 + w, k0:k0 + w].copy()
        for j in range(w):
            if D[j, j] != 0:
                D[j + 1:, j] /= D[j, j]
            D[j + 1:, j + 1:] -= np.outer(D[j + 1:, j], D[j, j + 1:])
        L11 = np.tril(D, -1) + np.eye(w)
        U11 = np.triu(D)
        A[k0:k0 + w, k0:k0 + w] = D
        if k0 + w < n:
            A[k0 + w:, k0:k0 + w] = np.linalg.solve(U11.T, A[k0 + w:, k0:k0 + w].T).T
            A[k0:k0 + w, k0 + w:] = np.linalg.solve(L11, A[k0:k0 + w, k0 + w:])
            A[k0 + w:, k0 + w:] -= A[k0 + w:, k0:k0 + w] @ A[k0:k0 + w, k0 + w:]
    return (ipiv, A) if return_lu else ipiv
