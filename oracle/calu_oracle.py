"""CPU restatement of LU with tournament pivoting on tall panels (test infrastructure only).

THIS MODULE IS THE CHECKER, NOT THE PRODUCT. Only ``tests/`` may import it.

The reference factors its dense BEM matrices with LAPACK's zgetrf (partial pivoting,
utils/gpu.py:145-191 through scipy.linalg.lu_factor). The product keeps partial pivoting for
panels shorter than a threshold and, for taller ones, chooses a panel's pivot rows by a
tournament (communication-avoiding LU, Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32
(2011) 1317-1350). This module states that rule:

  * ``tournament(P, cp)``: the rows of the m x w panel ``P`` in blocks of ``cp`` consecutive
    rows; each block runs Gaussian elimination with partial pivoting (largest |re| + |im|, the
    earlier candidate on a tie) on its own rows and keeps its w pivot rows in pivot order; the
    kept rows of all blocks, in block order, are the candidates of the next round, again in
    blocks of ``cp``, each elimination starting from the candidates' original entries; rounds
    end when one block remains. Its w pivot rows are the panel's pivots.
  * ``ipiv_from_rows(rows, j0)``: LAPACK-style interchanges (row j0 + c exchanged with
    ipiv[j0 + c], one after the other) that bring those rows to the top, in order.
  * ``getrf(A, nb, tall, cp)``: right-looking blocked LU with panels of nb columns; a panel of
    at least ``tall`` rows takes its pivots from ``tournament``, is permuted, and is factored
    without further pivoting (LU of the w x w top block, L21 = A21 U11^{-1}); shorter panels
    use partial pivoting. Returns the packed factor and 0-based ipiv as scipy's lu_factor.

Pinned in tests/test_calu_oracle.py: with one block the tournament's rows are scipy's partial
pivots of the panel; the rows of a two-round tournament are scipy's pivots of the stacked
winners of each block; ``tall`` above n reproduces scipy.linalg.lu_factor (pivots exactly,
factors to rounding); and P A = L U to rounding for any threshold.
"""
import numpy as np


def _cabs1(x):
    return np.abs(x.real) + np.abs(x.imag)


def gepp_rows(R, rows):
    """Pivot rows (entries of ``rows``, in pivot order) of Gaussian elimination with partial
    pivoting on the candidate rows R (k x w); ties go to the earlier candidate; at most w rows,
    fewer when k < w."""
    X = np.array(R, dtype=np.result_type(R.dtype, np.complex64), copy=True)
    rows = np.asarray(rows)
    live = np.ones(len(rows), bool)
    out = []
    w = X.shape[1]
    for k in range(w):
        if not live.any():
            break
        mag = np.where(live, _cabs1(X[:, k]), -1.0)
        i = int(np.argmax(mag))
        out.append(int(rows[i]))
        live[i] = False
        piv = X[i, k]
        if piv != 0:
            l = X[:, k] / piv
            X[live, k + 1:] -= np.outer(l[live], X[i, k + 1:])
    return out


def tournament(P, cp, j0=0):
    """Pivot rows (global indices j0 + local row) of the m x w panel P, in pivot order."""
    m, w = P.shape
    cand = [list(range(j0 + s, j0 + min(s + cp, m))) for s in range(0, m, cp)]
    while True:
        wins = [gepp_rows(P[np.asarray(c) - j0], c) for c in cand]
        if len(wins) == 1:
            return wins[0]
        flat = [r for wl in wins for r in wl]
        cand = [flat[s:s + cp] for s in range(0, len(flat), cp)]


def ipiv_from_rows(rows, j0):
    """Interchanges ipiv[j0 + c] (0-based) that move original row rows[c] to position j0 + c."""
    pos = {}                 # position -> original row, for the positions touched
    where = {}               # original row -> position, for the rows moved
    ipiv = []
    for c, want in enumerate(rows):
        col = j0 + c
        p = where.get(want, want)
        ipiv.append(p)
        vc, vp = pos.get(col, col), pos.get(p, p)
        pos[col], pos[p] = vp, vc
        where[vp], where[vc] = col, p
    return ipiv


def getrf(A, nb=32, tall=None, cp=1024):
    """(lu, ipiv) of A as scipy.linalg.lu_factor, with tournament pivoting on panels of at least
    ``tall`` rows (None: never)."""
    a = np.array(A, dtype=np.result_type(A.dtype, np.complex64), copy=True)
    n = a.shape[0]
    ipiv = np.arange(n)
    for j0 in range(0, n, nb):
        w = min(nb, n - j0)
        m = n - j0
        if tall is not None and m >= tall:
            rows = tournament(a[j0:, j0:j0 + w], cp, j0)
            piv = ipiv_from_rows(rows, j0)
            for c, p in enumerate(piv):
                ipiv[j0 + c] = p
                if p != j0 + c:
                    a[[j0 + c, p], :] = a[[p, j0 + c], :]
            # top block without pivoting, then L21 = A21 U11^{-1}
            for k in range(w):
                d = a[j0 + k, j0 + k]
                if d != 0:
                    a[j0 + k + 1:j0 + w, j0 + k] /= d
                    a[j0 + k + 1:j0 + w, j0 + k + 1:j0 + w] -= np.outer(a[j0 + k + 1:j0 + w, j0 + k],
                                                                       a[j0 + k, j0 + k + 1:j0 + w])
            U = np.triu(a[j0:j0 + w, j0:j0 + w])
            if m > w:
                dz = np.diag(U) == 0
                Us = U.copy()
                Us[dz, dz] = 1
                a[j0 + w:, j0:j0 + w] = np.linalg.solve(Us.T, a[j0 + w:, j0:j0 + w].T).T
        else:
            for k in range(j0, j0 + w):
                mag = _cabs1(a[k:, k])
                p = k + int(np.argmax(mag))
                ipiv[k] = p
                if p != k:
                    a[[k, p], :] = a[[p, k], :]
                d = a[k, k]
                if d != 0:
                    a[k + 1:, k] /= d
                    a[k + 1:, k + 1:j0 + w] -= np.outer(a[k + 1:, k], a[k, k + 1:j0 + w])
        if j0 + w < n:
            # U12 = L11^{-1} A12, trailing update
            L11 = np.tril(a[j0:j0 + w, j0:j0 + w], -1) + np.eye(w)
            a[j0:j0 + w, j0 + w:] = np.linalg.solve(L11, a[j0:j0 + w, j0 + w:])
            a[j0 + w:, j0 + w:] -= a[j0 + w:, j0:j0 + w] @ a[j0:j0 + w, j0 + w:]
    return a, ipiv
