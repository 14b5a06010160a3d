"""CPU restatement of partially pivoted ACA (test infrastructure only).

THIS MODULE IS THE CHECKER, NOT THE PRODUCT. Only ``tests/`` may import it.

``aca(A, htol, kmax)`` follows greenfun/hmatrix.py:195-336 step by step on a dense block ``A``
(the reference also takes its entries from the full dense matrix, aca_compgreen_ret.py:196-208):

  * first pivot row 0; residual row r = A[pr, :] - V[:, :k] @ U[pr, :k];
  * pivot column = argmax of |r| with used columns set to 0; stop if |pivot| < 1e-15;
  * residual column c = A[:, pc] - U[:, :k] @ V[pc, :k]; u = c / pivot, v = r;
  * |S|_F^2 += |u|^2 |v|^2 + 2 Re sum_q (U_q^H u)(V_q^H v); stop when
    |u| |v| < htol |S|_F; next pivot row = argmax |u| with used rows set to 0;
  * at most min(m, n, kmax) terms.

``confirm=True`` (the product's rule, not the reference's): a stop is accepted only when the
next cross, taken from a fresh row (an unused row from a fixed pseudo-random sequence of the
block index and step, ``fresh_row``), meets the criterion too; otherwise the pivoting goes on.
The reference's rule (``confirm=False``) stops at the first small cross, which on blocks with
paired singular values (two rods related by symmetry) can leave a residual of 1e-3 of the block
whatever htol is.

Pinned in tests/test_hmatrix_host.py by exact recovery of random low-rank matrices (rank r
matrices come back with r terms and rounding-level error) and by the error of smooth
Green-function blocks against the truncated SVD.
"""
import numpy as np


def fresh_row(used_r, block, step):
    """Unused row from the pseudo-random sequence of (block, step): start at
    ((6364136223846793005 (131 block + step + 1) + 1442695040888963407) mod 2^64) >> 33 mod m and
    walk forward to the first unused row; m when all rows are used."""
    m = len(used_r)
    h = (6364136223846793005 * (block * 131 + step + 1) + 1442695040888963407) % (1 << 64)
    start = (h >> 33) % m
    for q in range(m):
        i = (start + q) % m
        if not used_r[i]:
            return i
    return m


def aca(A, htol=1e-6, kmax=100, confirm=True, block=0):
    """Returns U (m, k), V (n, k) with A ~ U V^T, and k."""
    A = np.asarray(A)
    m, n = A.shape
    maxr = min(m, n, kmax)
    U = np.zeros((m, maxr), complex)
    V = np.zeros((n, maxr), complex)
    used_r = np.zeros(m, bool)
    used_c = np.zeros(n, bool)
    pr = 0
    frob = 0.0
    k = 0
    confirming = False
    for step in range(maxr):
        row = A[pr, :].astype(complex) - V[:, :k] @ U[pr, :k]
        ar = np.abs(row)
        ar[used_c] = 0.0
        pc = int(np.argmax(ar))
        piv = row[pc]
        if abs(piv) < 1e-15:
            break
        col = A[:, pc].astype(complex) - U[:, :k] @ V[pc, :k]
        u = col / piv
        v = row
        U[:, k] = u
        V[:, k] = v
        k += 1
        used_r[pr] = True
        used_c[pc] = True
        term = np.vdot(u, u).real * np.vdot(v, v).real
        cross = 0.0
        if k > 1:
            du = np.conj(U[:, :k - 1]).T @ u
            dv = np.conj(V[:, :k - 1]).T @ v
            cross = 2.0 * float(np.real(np.dot(du, dv)))
        frob += term + cross
        au = np.abs(u)
        au[used_r] = -1.0
        nxt = int(np.argmax(au)) if np.any(~used_r) else m
        if frob > 0 and np.sqrt(term) < htol * np.sqrt(abs(frob)):
            if not confirm or confirming:
                break
            confirming = True
            nxt = fresh_row(used_r, block, step)
        else:
            confirming = False
        if nxt >= m:
            break
        pr = nxt
    return U[:, :k], V[:, :k], k
