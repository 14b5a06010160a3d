"""The reference's dense LU dispatch (mnpbem/utils/gpu.py:145-301) on libmnpbem_hip.so.

A maintainer of the reference binds this module behind a new package tag 'mi355x' in
``lu_factor_dispatch``, ``lu_solve_dispatch``, ``lu_solve_native``, ``eye_like_lu`` and
``to_host`` (the patch is in INTEGRATION.md). The functions keep the reference's signatures,
argument meaning and return conventions:

* ``lu_factor_dispatch(A)`` -> ('mi355x', lu, perm): complex128 LU with partial pivoting on the
  device (``mnp_getrf_batched``, LAPACK's pivots), the factor column-major in a torch tensor;
  a zero pivot is not detected, as ``check_finite=False`` in the reference (utils/gpu.py:189-191),
  but reported through ``info(pkg)``.
* ``lu_solve_dispatch(pkg, b, trans=0)`` -> host numpy, ``lu_solve_native(pkg, b, trans=0)`` ->
  device tensor when ``b`` is one (else host numpy), both for ``b`` of shape (n,) or (n, r), any r:
  up to 8 columns by ``mnp_getrs_batched``; more columns through the right-side solve
  ``mnp_rsolve_batched`` (X = B A^{-1}): A^T X = B is X^T = B^T A^{-1}, A^H X = B is
  X^H = B^H A^{-1}, and A X = B with many columns is A^{-1} (the right-side solve of the identity)
  times B (``mnp_gemm_sub_batched``). scipy's ``trans`` codes 0, 1, 2 ('N', 'T', 'C').
* ``eye_like_lu(pkg, n, dtype=None)``: the identity on the package's device.
* ``to_host(x)``: numpy from a device tensor (or ``np.asarray``).
"""
import ctypes

import numpy as np
import torch

from ._lib import call

TAG = "mi355x"


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("the mi355x LU dispatch needs a ROCm GPU")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _cm(x):
    """(n, r) array-like -> column-major complex128 device tensor of shape (r, n)."""
    if isinstance(x, torch.Tensor):
        t = x.to(device=_dev(), dtype=torch.complex128)
        return t.T.contiguous() if t.dim() == 2 else t.reshape(1, -1).contiguous()
    a = np.asarray(x, dtype=complex)
    a = a.reshape(-1, 1) if a.ndim == 1 else a
    return torch.tensor(np.ascontiguousarray(a.T), device=_dev())


def lu_factor_dispatch(A, nb=32, **kwargs):
    """(utils/gpu.py:145-191) Factor A (n x n); returns (TAG, lu, perm)."""
    A = A if isinstance(A, torch.Tensor) else np.asarray(A, dtype=complex)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("lu_factor_dispatch: square matrix expected, got shape {}".format(tuple(A.shape)))
    lu = _cm(A)                                  # column-major: element (i, j) at [j, i]
    ipiv = torch.empty((1, n), dtype=torch.int32, device=lu.device)
    info = torch.empty(1, dtype=torch.int32, device=lu.device)
    perm = torch.empty((2, n), dtype=torch.int32, device=lu.device)
    s = _stream()
    call("mnp_getrf_batched", 0, n, 1, lu.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), nb, s)
    call("mnp_perm_batched", n, 1, ipiv.data_ptr(), n, perm[0].data_ptr(), perm[1].data_ptr(), n, s)
    return (TAG, lu, (perm, info, ipiv))


def info(pkg):
    """LAPACK info of the factorisation: 0, or j > 0 for an exact zero pivot in column j-1."""
    return int(pkg[2][1].item())


def _right_solve(pkg, Bt):
    """X = Bt A^{-1} for Bt column-major (tensor (n, m): m rows); returns (n, m) column-major."""
    _, lu, (perm, _, _) = pkg
    n = lu.shape[0]
    m = Bt.shape[1]
    B = Bt.clone()                               # overwritten by the solve
    X = torch.empty_like(B)
    call("mnp_rsolve_batched", 0, m, n, 1, lu.data_ptr(), n, n * n, perm[1].data_ptr(), n, B.data_ptr(), m, m * n,
         X.data_ptr(), m, m * n, 32, _stream())
    return X


def _solve_dev(pkg, b, trans=0):
    """Solution as a column-major device tensor (r, n) for b (n,) or (n, r)."""
    if isinstance(trans, str):
        trans = {"N": 0, "T": 1, "C": 2}[trans.upper()]
    _, lu, (perm, _, _) = pkg
    n = lu.shape[0]
    B = _cm(b)                                   # (r, n): column c at [c]
    r = B.shape[0]
    if B.shape[1] != n:
        raise ValueError("lu_solve: right-hand side has {} rows, the factor {}".format(B.shape[1], n))
    s = _stream()
    if trans == 0 and r <= 8:
        X = torch.empty_like(B)
        call("mnp_getrs_batched", 0, n, r, 1, lu.data_ptr(), n, n * n, perm[0].data_ptr(), n, B.data_ptr(), n,
             n * r, X.data_ptr(), n, n * r, s)
        return X
    if trans in (1, 2):
        # A^T X = B  <=>  X^T A = B^T;  A^H X = B  <=>  X^H A = B^H. As a column-major (n, r) matrix
        # B^T is B's column-major (r, n) storage read row-major, i.e. the (r x n) matrix with
        # element (c, i) = B[i, c] at [i, c]: the transpose of the stored tensor.
        Bt = B.T.contiguous()                    # column-major storage of the r x n matrix B^T
        if trans == 2:
            Bt = Bt.conj()
        Xt = _right_solve(pkg, Bt)               # column-major r x n matrix X^T (or X^H)
        X = Xt.T.contiguous()
        return X.conj() if trans == 2 else X
    # A X = B with many columns: A^{-1} (right-side solve of the identity), then A^{-1} B
    eye = torch.eye(n, dtype=torch.complex128, device=lu.device)
    Ainv = _right_solve(pkg, eye)                # column-major A^{-1}
    X = torch.zeros_like(B)
    negA = -Ainv
    call("mnp_gemm_sub_batched", 0, n, r, n, negA.data_ptr(), n, n * n, B.data_ptr(), n, n * r, X.data_ptr(), n,
         n * r, 1, s)
    return X


def _from_cm(X, like):
    """Column-major (r, n) tensor -> (n,) or (n, r) in the layout of ``like``."""
    out = X.T
    if np.ndim(like) == 1 or (isinstance(like, torch.Tensor) and like.dim() == 1):
        out = out[:, 0]
    return out


def lu_solve_native(pkg, b, trans=0, **kwargs):
    """(utils/gpu.py:254-276) Device tensor in, device tensor out; host in, host out."""
    X = _from_cm(_solve_dev(pkg, b, trans), b)
    return X if isinstance(b, torch.Tensor) else X.cpu().numpy()


def lu_solve_dispatch(pkg, b, trans=0, **kwargs):
    """(utils/gpu.py:194-251) Host numpy result, whatever b is."""
    return to_host(_from_cm(_solve_dev(pkg, b, trans), b))


def eye_like_lu(pkg, n, dtype=None):
    """(utils/gpu.py:282-296) Identity on the device of the factor."""
    return torch.eye(n, dtype=torch.complex128 if dtype is None else dtype, device=pkg[1].device)


def to_host(x):
    """(utils/gpu.py:299-305)"""
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def lu_backend(pkg):
    return pkg[0]
