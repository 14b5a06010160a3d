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

Pinned in tests/test_hmatrix_host.py by exact recovery of random low-rank matrices (rank r
matrices come back with r terms and rounding-level error) and by the error of smooth
Green-function blocks against the truncated SVD.
"""
import numpy as np


def aca(A, htol=1e-6, kmax=100):
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
    for _ in range(maxr):
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
        if frob > 0 and np.sqrt(term) < htol * np.sqrt(abs(frob)):
            break
        au = np.abs(u)
        au[used_r] = 0.0
        pr = int(np.argmax(au))
    return U[:, :k], V[:, :k], k
