"""Kernel-level parity on the GPU: matrix-core GEMM, LU, permutations, solves, GEMV.
Reference arithmetic: numpy / LAPACK (scipy) on the host."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    return torch.device("cuda", 0)


def _lib():
    from mnpbem_amd import _lib
    return _lib


def _sh():
    return torch.cuda.current_stream().cuda_stream


def _cm(mats, dtype):
    """list of (m, n) numpy matrices -> device tensor (B, n, m) holding them column-major."""
    arr = np.stack([np.asarray(M).T for M in mats])
    return torch.tensor(np.ascontiguousarray(arr), dtype=dtype, device=_dev())


def _np(t, b):
    return t[b].transpose(0, 1).cpu().numpy()


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_mfma_layout_probe():
    banner("mfma layout probe")
    out = torch.zeros(1536, dtype=torch.float64, device=_dev())
    _lib().call("mnp_probe_mfma_layout", out.data_ptr(), _sh())
    o = out.cpu().numpy()
    got = o[:512].reshape(2, 64, 4)
    l = np.arange(64)[:, None]
    r = np.arange(4)[None, :]
    # D[i][j] = (i + 16)(j + 1); register r of lane l holds row l/16 + 4r (f64) and
    # 4(l/16) + r (f32), column l % 16
    assert np.array_equal(got[0], (l // 16 + 4 * r + 16) * ((l % 16) + 1))
    assert np.array_equal(got[1], (4 * (l // 16) + r + 16) * ((l % 16) + 1))
    # 32x32x2: register r of lane l holds row (r%4) + 8(r/4) + 4(l/32), column l % 32 -- the
    # mapping the complex-float GEMM epilogue is written for
    g32 = o[512:].reshape(64, 16)
    r16 = np.arange(16)[None, :]
    rows = (r16 % 4) + 8 * (r16 // 4) + 4 * (l // 32)
    assert np.array_equal(g32, (rows + 16) * ((l % 32) + 1))


@pytest.mark.parametrize("prec,m,n,k,lda,ldb", [(0, 150, 97, 45, 150, 45), (1, 150, 97, 45, 150, 45),
                                                 (1, 149, 97, 46, 150, 46), (1, 64, 64, 8, 64, 8),
                                                 (1, 131, 70, 37, 132, 38), (0, 131, 70, 37, 132, 38)])
def test_gemm_sub(prec, m, n, k, lda, ldb):
    """Both complex-float kernels (16-byte loads for even leading dimensions, scalar loads
    otherwise) and the complex-double one, with ragged edges and padded leading dimensions."""
    banner("gemm_sub prec={} m={} n={} k={} lda={} ldb={}".format(prec, m, n, k, lda, ldb))
    rng = np.random.default_rng(1)
    B_ = 3
    dt = torch.complex128 if prec == 0 else torch.complex64
    A = [_rand(rng, lda, k) for _ in range(B_)]
    Bm = [_rand(rng, ldb, n) for _ in range(B_)]
    Cm = [_rand(rng, m, n) for _ in range(B_)]
    tA, tB, tC = _cm(A, dt), _cm(Bm, dt), _cm(Cm, dt)
    _lib().call("mnp_gemm_sub_batched", prec, m, n, k, tA.data_ptr(), lda, lda * k, tB.data_ptr(), ldb, ldb * n,
                tC.data_ptr(), m, m * n, B_, _sh())
    tol = 1e-13 if prec == 0 else 2e-5
    for b in range(B_):
        ref = Cm[b] - A[b][:m] @ Bm[b][:k]
        assert np.max(np.abs(_np(tC, b) - ref)) / np.max(np.abs(ref)) < tol


@pytest.mark.parametrize("split,tile", [(12, 11), (13, 11), (12, 21), (13, 21), (12, 22), (13, 22), (16, 11)])
@pytest.mark.parametrize("m,n,k,lda,ldb", [(200, 167, 45, 200, 46), (300, 260, 128, 300, 128), (192, 192, 16, 192, 16),
                                           (331, 270, 37, 332, 38), (171, 165, 37, 171, 37), (150, 70, 37, 150, 38)])
def test_gemm_sub_split_bf16(split, tile, m, n, k, lda, ldb):
    """Complex-float GEMM on the BF16 matrix cores with 2 or 3 operand terms split once into
    planes (12, 13); ragged edges, K not a multiple of 16, padded and odd leading dimensions,
    and a narrow C (70 columns) that the plane modes leave to the FP32 matrix cores: error
    against the exact product within the terms kept (2^-15 and 2^-22 relative to |A||B|, with
    FP32 accumulation)."""
    banner("gemm_sub split={} tile={} m={} n={} k={}".format(split, tile, m, n, k))
    rng = np.random.default_rng(2)
    B_ = 2
    A = [_rand(rng, lda, k) for _ in range(B_)]
    Bm = [_rand(rng, ldb, n) for _ in range(B_)]
    Cm = [_rand(rng, m, n) for _ in range(B_)]
    tA, tB, tC = _cm(A, torch.complex64), _cm(Bm, torch.complex64), _cm(Cm, torch.complex64)
    old, old_t = ctypes.c_int(0), ctypes.c_int(0)
    _lib().call("mnp_get_tuning", 4, ctypes.addressof(old))
    _lib().call("mnp_get_tuning", 5, ctypes.addressof(old_t))
    _lib().call("mnp_set_tuning", 4, split)
    _lib().call("mnp_set_tuning", 5, tile)
    try:
        _lib().call("mnp_gemm_sub_batched", 1, m, n, k, tA.data_ptr(), lda, lda * k, tB.data_ptr(), ldb, ldb * n,
                    tC.data_ptr(), m, m * n, B_, _sh())
        torch.cuda.synchronize()
    finally:
        _lib().call("mnp_set_tuning", 4, old.value)
        _lib().call("mnp_set_tuning", 5, old_t.value)
    for b in range(B_):
        A64 = A[b][:m].astype(np.complex64).astype(complex)
        B64 = Bm[b][:k].astype(np.complex64).astype(complex)
        ref = Cm[b].astype(np.complex64).astype(complex) - A64 @ B64
        scale = np.abs(A64) @ np.abs(B64)
        err = np.max(np.abs(_np(tC, b) - ref) / (scale + np.abs(ref)))
        assert err < (4e-5 if split % 10 == 2 else 1e-6), err
        if split % 10 == 3:           # as accurate as the FP32 kernel (within 2x)
            assert np.max(np.abs(_np(tC, b) - ref)) / np.max(np.abs(ref)) < 4e-6


def test_gemm_fp16_pairs_as_accurate_as_fp32():
    """Scaled fp16 pairs (mode 16, about 22 significant bits per operand, 3 products per real
    product) against the FP32 matrix cores (mode 0) and three bfloat16 terms (mode 13) on operands
    whose rows and columns span 12 orders of magnitude: the error of every entry relative to
    (|A||B|)_ij stays within 4x that of the FP32 kernel."""
    banner("gemm fp16 pairs vs fp32")
    rng = np.random.default_rng(21)
    m, n, k, B_ = 640, 576, 256, 2
    A = [_rand(rng, m, k) * 10.0 ** rng.uniform(-6, 6, (m, 1)) for _ in range(B_)]
    Bm = [_rand(rng, k, n) * 10.0 ** rng.uniform(-6, 6, (1, n)) for _ in range(B_)]
    Cm = [np.zeros((m, n), complex) for _ in range(B_)]
    tA, tB = _cm(A, torch.complex64), _cm(Bm, torch.complex64)
    err = {}
    old = ctypes.c_int(0)
    _lib().call("mnp_get_tuning", 4, ctypes.addressof(old))
    try:
        for mode in (0, 13, 16):
            _lib().call("mnp_set_tuning", 4, mode)
            tC = _cm(Cm, torch.complex64)
            _lib().call("mnp_gemm_sub_batched", 1, m, n, k, tA.data_ptr(), m, m * k, tB.data_ptr(), k, k * n,
                        tC.data_ptr(), m, m * n, B_, _sh())
            torch.cuda.synchronize()
            e = 0.0
            for b in range(B_):
                A64 = A[b].astype(np.complex64).astype(complex)
                B64 = Bm[b].astype(np.complex64).astype(complex)
                e = max(e, np.max(np.abs(_np(tC, b) + A64 @ B64) / (np.abs(A64) @ np.abs(B64))))
            err[mode] = e
    finally:
        _lib().call("mnp_set_tuning", 4, old.value)
    print("max error / (|A||B|): fp32 {:.3g}, 3-term bf16 {:.3g}, fp16 pairs {:.3g}".format(err[0], err[13], err[16]))
    assert err[16] < 4 * err[0]
    assert err[13] < 4 * err[0]


@pytest.mark.parametrize("m,n,k,lda,ldb", [(300, 260, 37, 301, 37), (640, 576, 128, 640, 128), (331, 270, 256, 332, 258),
                                           (500, 300, 500, 500, 501), (260, 200, 600, 260, 600), (200, 170, 64, 200, 64)])
def test_gemm_fp16_pairs_one_pass_identical(m, n, k, lda, ldb):
    """Scaled fp16 pairs: scales and planes formed in one pass over each operand (K <= 512) or in
    two (scale, then split) give bit-identical products; K = 600 takes the two-pass form in both."""
    banner("gemm fp16 pairs one pass m={} n={} k={}".format(m, n, k))
    rng = np.random.default_rng(k)
    A = [_rand(rng, lda, k) * 10.0 ** rng.uniform(-4, 4, (lda, 1)) for _ in range(2)]
    Bm = [_rand(rng, ldb, n) * 10.0 ** rng.uniform(-4, 4, (1, n)) for _ in range(2)]
    Cm = [_rand(rng, m, n) for _ in range(2)]
    tA, tB = _cm(A, torch.complex64), _cm(Bm, torch.complex64)
    old = [ctypes.c_int(0), ctypes.c_int(0)]
    _lib().call("mnp_get_tuning", 4, ctypes.addressof(old[0]))
    _lib().call("mnp_get_tuning", 11, ctypes.addressof(old[1]))
    out = {}
    try:
        _lib().call("mnp_set_tuning", 4, 16)
        for fused in (0, 1):
            _lib().call("mnp_set_tuning", 11, fused)
            tC = _cm(Cm, torch.complex64)
            _lib().call("mnp_gemm_sub_batched", 1, m, n, k, tA.data_ptr(), lda, lda * k, tB.data_ptr(), ldb, ldb * n,
                        tC.data_ptr(), m, m * n, 2, _sh())
            torch.cuda.synchronize()
            out[fused] = tC.cpu()
    finally:
        _lib().call("mnp_set_tuning", 4, old[0].value)
        _lib().call("mnp_set_tuning", 11, old[1].value)
    assert torch.equal(out[0], out[1])
    A64 = A[0][:m].astype(np.complex64).astype(complex)
    B64 = Bm[0][:k].astype(np.complex64).astype(complex)
    ref = Cm[0].astype(np.complex64).astype(complex) - A64 @ B64
    err = np.max(np.abs(out[1][0].numpy().T - ref) / (np.abs(A64) @ np.abs(B64) + np.abs(ref)))
    assert err < 1e-6, err


@pytest.mark.parametrize("split", [12, 13, 16])
@pytest.mark.parametrize("m,n,k", [(700, 650, 256), (331, 270, 37), (1000, 180, 512)])
def test_gemm_plane_tiles_identical(split, m, n, k):
    """The plane kernels with direct LDS copies (tiles 111, 21, 12, 22: 64 x 64, 128 x 64,
    64 x 128, 128 x 128) form the same term products in the same order per element as the
    64 x 64 kernel staged through registers (tile 11), for bfloat16 terms and for scaled fp16
    pairs: bit-identical results."""
    banner("gemm plane tiles 11 vs 22 split={} m={} n={} k={}".format(split, m, n, k))
    rng = np.random.default_rng(m + n + k)
    B_ = 3
    A = [_rand(rng, m, k) for _ in range(B_)]
    Bm = [_rand(rng, k, n) for _ in range(B_)]
    Cm = [_rand(rng, m, n) for _ in range(B_)]
    tA, tB = _cm(A, torch.complex64), _cm(Bm, torch.complex64)
    out = {}
    old, old_t = ctypes.c_int(0), ctypes.c_int(0)
    _lib().call("mnp_get_tuning", 4, ctypes.addressof(old))
    _lib().call("mnp_get_tuning", 5, ctypes.addressof(old_t))
    try:
        _lib().call("mnp_set_tuning", 4, split)
        for tile in (11, 111, 21, 12, 22):
            _lib().call("mnp_set_tuning", 5, tile)
            tC = _cm(Cm, torch.complex64)
            _lib().call("mnp_gemm_sub_batched", 1, m, n, k, tA.data_ptr(), m, m * k, tB.data_ptr(), k, k * n,
                        tC.data_ptr(), m, m * n, B_, _sh())
            torch.cuda.synchronize()
            out[tile] = tC.cpu()
    finally:
        _lib().call("mnp_set_tuning", 4, old.value)
        _lib().call("mnp_set_tuning", 5, old_t.value)
    for tile in (111, 21, 12, 22):
        assert torch.equal(out[11], out[tile]), tile
    A64, B64 = A[0].astype(np.complex64).astype(complex), Bm[0].astype(np.complex64).astype(complex)
    ref = Cm[0].astype(np.complex64).astype(complex) - A64 @ B64
    got = out[22][0].numpy().T
    err = np.max(np.abs(got - ref) / (np.abs(A64) @ np.abs(B64) + np.abs(ref)))
    assert err < (4e-5 if split == 12 else 1e-6), err


@pytest.mark.parametrize("n,nb", [(200, 32), (97, 32), (64, 16)])
def test_getrf_matches_lapack(n, nb):
    banner("getrf n={} nb={}".format(n, nb))
    rng = np.random.default_rng(n)
    B = 3
    A = [_rand(rng, n, n) for _ in range(B)]
    t = _cm(A, torch.complex128)
    ipiv = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(B, dtype=torch.int32, device=_dev())
    _lib().call("mnp_getrf_batched", 0, n, B, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), nb, _sh())
    assert np.all(info.cpu().numpy() == 0)
    for b in range(B):
        lu, piv = sla.lu_factor(A[b])
        assert np.array_equal(ipiv[b].cpu().numpy(), piv)
        assert np.max(np.abs(_np(t, b) - lu)) < 1e-11 * np.max(np.abs(lu))


@pytest.mark.parametrize("rec", [0, 1])
@pytest.mark.parametrize("n,nbo", [(600, 256), (777, 384), (1100, 512)])
def test_getrf_outer_block_orders_match_lapack(n, nbo, rec):
    """Outer blocks factored panel by panel (0) or by recursive halving (1; uneven halves when
    the block has an odd number of panels) on random matrices, which interchange rows at almost
    every column: LAPACK's pivots in complex128 and its factors to rounding."""
    banner("getrf n={} nbo={} recursive={}".format(n, nbo, rec))
    rng = np.random.default_rng(n + rec)
    B = 2
    A = [_rand(rng, n, n) for _ in range(B)]
    t = _cm(A, torch.complex128)
    ipiv = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(B, dtype=torch.int32, device=_dev())
    old = [ctypes.c_int(0), ctypes.c_int(0)]
    _lib().call("mnp_get_tuning", 0, ctypes.addressof(old[0]))
    _lib().call("mnp_get_tuning", 10, ctypes.addressof(old[1]))
    try:
        _lib().call("mnp_set_tuning", 0, nbo)
        _lib().call("mnp_set_tuning", 10, rec)
        _lib().call("mnp_getrf_batched", 0, n, B, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32,
                    _sh())
        torch.cuda.synchronize()
    finally:
        _lib().call("mnp_set_tuning", 0, old[0].value)
        _lib().call("mnp_set_tuning", 10, old[1].value)
    assert np.all(info.cpu().numpy() == 0)
    for b in range(B):
        lu, piv = sla.lu_factor(A[b])
        assert np.array_equal(ipiv[b].cpu().numpy(), piv)
        assert np.max(np.abs(_np(t, b) - lu)) < 1e-11 * np.max(np.abs(lu))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("rec", [0, 1])
@pytest.mark.parametrize("m,n,nbo", [(500, 777, 384), (301, 640, 256)])
def test_rsolve_outer_block_orders(m, n, nbo, rec, prec):
    """Right-side solve X = B A^{-1} with outer blocks solved panel by panel (0) or by recursive
    halving (1), B of m rows: numpy's B inv(A) (complex128 to 1e-11, complex64 to 5e-4 on these
    random matrices)."""
    banner("rsolve m={} n={} nbo={} recursive={} prec={}".format(m, n, nbo, rec, prec))
    rng = np.random.default_rng(m + n + rec)
    A = _rand(rng, n, n) + 0.1 * n * np.eye(n) * prec
    Bm = _rand(rng, m, n)
    dt = torch.complex128 if prec == 0 else torch.complex64
    t = _cm([A], dt)
    ipiv = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(1, dtype=torch.int32, device=_dev())
    pf = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    q = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    tb = _cm([Bm], dt)
    tX = torch.zeros_like(tb)
    old = [ctypes.c_int(0), ctypes.c_int(0)]
    _lib().call("mnp_get_tuning", 0, ctypes.addressof(old[0]))
    _lib().call("mnp_get_tuning", 10, ctypes.addressof(old[1]))
    try:
        _lib().call("mnp_set_tuning", 0, nbo)
        _lib().call("mnp_set_tuning", 10, rec)
        _lib().call("mnp_getrf_batched", prec, n, 1, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32,
                    _sh())
        _lib().call("mnp_perm_batched", n, 1, ipiv.data_ptr(), n, pf.data_ptr(), q.data_ptr(), n, _sh())
        _lib().call("mnp_rsolve_batched", prec, m, n, 1, t.data_ptr(), n, n * n, q.data_ptr(), n, tb.data_ptr(), m,
                    m * n, tX.data_ptr(), m, m * n, 32, _sh())
        torch.cuda.synchronize()
    finally:
        _lib().call("mnp_set_tuning", 0, old[0].value)
        _lib().call("mnp_set_tuning", 10, old[1].value)
    X = Bm @ np.linalg.inv(A)
    assert np.max(np.abs(_np(tX, 0) - X)) / np.max(np.abs(X)) < (1e-11 if prec == 0 else 5e-4)


@pytest.mark.parametrize("n", [700, 2100])
def test_getrf_single_register_panel_pivots(n):
    """Complex-float LU through the register-resident panel (every panel below 3072 rows) on
    matrices whose pivot in each column is unambiguous (a dominant entry at a random row):
    LAPACK's pivots exactly, so every interchange the panel and k_laswp make is checked, and
    P A = L U to complex64 rounding."""
    banner("getrf complex64 register panel pivots n={}".format(n))
    rng = np.random.default_rng(n)
    B = 2
    A = []
    for _ in range(B):
        M_ = _rand(rng, n, n)
        M_[rng.permutation(n), np.arange(n)] += 40.0 * np.sqrt(n)
        A.append(M_)
    t = _cm(A, torch.complex64)
    ipiv = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(B, dtype=torch.int32, device=_dev())
    _lib().call("mnp_getrf_batched", 1, n, B, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32, _sh())
    assert np.all(info.cpu().numpy() == 0)
    for b in range(B):
        _, piv = sla.lu_factor(A[b])
        assert np.array_equal(ipiv[b].cpu().numpy(), piv)
        LU = _np(t, b).astype(complex)
        L = np.tril(LU, -1) + np.eye(n)
        U = np.triu(LU)
        PA = A[b].copy()
        for i, p in enumerate(piv):
            PA[[i, p]] = PA[[p, i]]
        assert np.linalg.norm(PA - L @ U) / np.linalg.norm(A[b]) < 1e-6


def test_getrf_single_backward_error():
    banner("getrf complex64")
    rng = np.random.default_rng(5)
    n, B = 180, 2
    A = [_rand(rng, n, n) for _ in range(B)]
    t = _cm(A, torch.complex64)
    ipiv = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(B, dtype=torch.int32, device=_dev())
    _lib().call("mnp_getrf_batched", 1, n, B, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32, _sh())
    for b in range(B):
        LU = _np(t, b).astype(complex)
        L = np.tril(LU, -1) + np.eye(n)
        U = np.triu(LU)
        PA = A[b].copy()
        for i, p in enumerate(ipiv[b].cpu().numpy()):
            PA[[i, p]] = PA[[p, i]]
        assert np.linalg.norm(PA - L @ U) / np.linalg.norm(A[b]) < 1e-5


def test_singular_matrix_reports_info():
    banner("getrf singular")
    n = 40
    A = np.zeros((n, n), complex)
    A[:, :10] = np.random.default_rng(2).standard_normal((n, 10))
    t = _cm([A], torch.complex128)
    ipiv = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(1, dtype=torch.int32, device=_dev())
    _lib().call("mnp_getrf_batched", 0, n, 1, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32, _sh())
    assert info.cpu().numpy()[0] == 11


@pytest.mark.parametrize("prec", [0, 1])
def test_solves(prec):
    banner("getrs/rsolve/gemv prec={}".format(prec))
    rng = np.random.default_rng(7)
    n, B, r = 133, 2, 5
    dt = torch.complex128 if prec == 0 else torch.complex64
    A = [_rand(rng, n, n) + n * np.eye(n) * 0.1 for _ in range(B)]
    Rhs = [_rand(rng, n, r) for _ in range(B)]
    Bm = [_rand(rng, n, n) for _ in range(B)]
    L = _lib()
    t = _cm(A, dt)
    ipiv = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(B, dtype=torch.int32, device=_dev())
    L.call("mnp_getrf_batched", prec, n, B, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32, _sh())
    pf = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    q = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    L.call("mnp_perm_batched", n, B, ipiv.data_ptr(), n, pf.data_ptr(), q.data_ptr(), n, _sh())
    tr = _cm(Rhs, dt)
    tx = torch.zeros_like(tr)
    L.call("mnp_getrs_batched", prec, n, r, B, t.data_ptr(), n, n * n, pf.data_ptr(), n, tr.data_ptr(), n, n * r,
           tx.data_ptr(), n, n * r, _sh())
    tb = _cm(Bm, dt)
    tX = torch.zeros_like(tb)
    L.call("mnp_rsolve_batched", prec, n, n, B, t.data_ptr(), n, n * n, q.data_ptr(), n, tb.data_ptr(), n, n * n,
           tX.data_ptr(), n, n * n, 32, _sh())
    tA = _cm(A, dt)
    ty = torch.zeros_like(tr)
    L.call("mnp_gemv_batched", prec, n, r, B, tA.data_ptr(), n, n * n, tr.data_ptr(), n, n * r, ty.data_ptr(), n,
           n * r, _sh())
    tol = 1e-11 if prec == 0 else 5e-4
    for b in range(B):
        x = np.linalg.solve(A[b], Rhs[b])
        assert np.max(np.abs(_np(tx, b) - x)) / np.max(np.abs(x)) < tol
        X = Bm[b] @ np.linalg.inv(A[b])
        assert np.max(np.abs(_np(tX, b) - X)) / np.max(np.abs(X)) < tol
        y = A[b] @ Rhs[b]
        assert np.max(np.abs(_np(ty, b) - y)) / np.max(np.abs(y)) < (1e-13 if prec == 0 else 1e-5)


class _knobs:
    """Library knobs (key -> value) for the block, restored afterwards."""

    def __init__(self, **kv):
        self.kv = {int(k[1:]): v for k, v in kv.items()}

    def __enter__(self):
        self.old = {}
        for k, v in self.kv.items():
            o = ctypes.c_int(0)
            _lib().call("mnp_get_tuning", k, ctypes.addressof(o))
            self.old[k] = o.value
            _lib().call("mnp_set_tuning", k, v)

    def __exit__(self, *exc):
        for k, v in self.old.items():
            _lib().call("mnp_set_tuning", k, v)


def test_getrf_multiworkgroup_panel_matches_lapack():
    """Panels taller than one workgroup with partial pivoting (cooperative launch, tall-panel
    knob 9 = 0): pivots and factors as LAPACK."""
    banner("getrf n=4700 batch 2 (panel over several workgroups)")
    rng = np.random.default_rng(11)
    n, B = 4700, 2
    A = [_rand(rng, n, n) for _ in range(B)]
    t = _cm(A, torch.complex128)
    ipiv = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(B, dtype=torch.int32, device=_dev())
    with _knobs(k9=0):
        _lib().call("mnp_getrf_batched", 0, n, B, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32,
                    _sh())
    assert np.all(info.cpu().numpy() == 0)
    for b in range(B):
        lu, piv = sla.lu_factor(A[b])
        assert np.array_equal(ipiv[b].cpu().numpy(), piv)
        assert np.max(np.abs(_np(t, b) - lu)) < 1e-10 * np.max(np.abs(lu))


@pytest.mark.parametrize("n,tall", [(4700, 4096), (2000, 512)])
def test_getrf_tournament_panels_match_oracle(n, tall):
    """Tall panels by tournament pivoting (knob 9 = 1, panels of at least `tall` rows, blocks of
    256 rows for complex128): pivots equal to the oracle's tournament (oracle/calu_oracle.py)
    exactly, factors to 1e-10, and P A = L U."""
    from oracle import calu_oracle as C
    banner("getrf tournament n={} tall={} batch 2".format(n, tall))
    rng = np.random.default_rng(12)
    B = 2
    A = [_rand(rng, n, n) for _ in range(B)]
    t = _cm(A, torch.complex128)
    ipiv = torch.zeros((B, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(B, dtype=torch.int32, device=_dev())
    with _knobs(k9=1, k3=tall):
        _lib().call("mnp_getrf_batched", 0, n, B, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32,
                    _sh())
    assert np.all(info.cpu().numpy() == 0)
    for b in range(B):
        lu, piv = C.getrf(A[b], nb=32, tall=tall, cp=256)
        got = ipiv[b].cpu().numpy()
        assert not np.array_equal(piv, sla.lu_factor(A[b])[1])
        assert np.array_equal(got, piv)
        assert np.max(np.abs(_np(t, b) - lu)) < 1e-10 * np.max(np.abs(lu))


def test_getrf_tournament_panels_complex64_backward_error():
    """complex64 (blocks of 1024 rows), panels of at least 1024 rows by tournament: normwise
    backward error of P A = L U within 3x that of the same factorisation with partial pivoting
    (knob 9 = 0), bounded multipliers."""
    banner("getrf tournament complex64 n=3000")
    rng = np.random.default_rng(13)
    n = 3000
    A = _rand(rng, n, n)
    err = {}
    for mode in (1, 0):
        t = _cm([A], torch.complex64)
        ipiv = torch.zeros((1, n), dtype=torch.int32, device=_dev())
        info = torch.zeros(1, dtype=torch.int32, device=_dev())
        with _knobs(k9=mode, k3=1024):
            _lib().call("mnp_getrf_batched", 1, n, 1, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(),
                        32, _sh())
        assert int(info.item()) == 0
        lu = _np(t, 0).astype(complex)
        L = np.tril(lu, -1) + np.eye(n)
        U = np.triu(lu)
        PA = A.copy()
        for i, p in enumerate(ipiv[0].cpu().numpy()):
            PA[[i, p]] = PA[[p, i]]
        assert np.max(np.abs(L)) < 20
        err[mode] = np.linalg.norm(PA - L @ U) / np.linalg.norm(A)
    print("backward error tournament {:.3g} partial pivoting {:.3g}".format(err[1], err[0]))
    assert err[0] < 1e-4 and err[1] < 3 * err[0]


def test_getrf_tournament_singular_column_reports_info():
    """A zero column inside a tournament panel: info = its 1-based index, as LAPACK."""
    banner("getrf tournament singular n=1100")
    rng = np.random.default_rng(14)
    n = 1100
    A = _rand(rng, n, n)
    A[:, 700] = 0
    t = _cm([A], torch.complex128)
    ipiv = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(1, dtype=torch.int32, device=_dev())
    with _knobs(k9=1, k3=256):
        _lib().call("mnp_getrf_batched", 0, n, 1, t.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32,
                    _sh())
    assert int(info.item()) == 701


def test_large_single_matrix_factor_and_solve():
    """n = 17000 complex64, one matrix: multi-workgroup panels and the permutation beyond
    16384 rows; normwise backward error of the solve."""
    banner("getrf/getrs n=17000 complex64")
    n = 17000
    g = torch.Generator(device=_dev()).manual_seed(3)
    A = torch.randn((n, n), dtype=torch.complex64, device=_dev(), generator=g)   # column-major = A^T
    b = torch.randn((1, n), dtype=torch.complex64, device=_dev(), generator=g)
    LU = A.clone()
    ipiv = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    info = torch.zeros(1, dtype=torch.int32, device=_dev())
    L = _lib()
    L.call("mnp_getrf_batched", 1, n, 1, LU.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32, _sh())
    pf = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    q = torch.zeros((1, n), dtype=torch.int32, device=_dev())
    L.call("mnp_perm_batched", n, 1, ipiv.data_ptr(), n, pf.data_ptr(), q.data_ptr(), n, _sh())
    x = torch.zeros_like(b)
    L.call("mnp_getrs_batched", 1, n, 1, 1, LU.data_ptr(), n, n * n, pf.data_ptr(), n, b.data_ptr(), n, n,
           x.data_ptr(), n, n, _sh())
    assert int(info.item()) == 0
    p = pf[0].long()
    assert torch.equal(torch.sort(p).values, torch.arange(n, device=_dev()))
    # the stored array is M^T with M the factored matrix: M x = b  <=>  x^T A = b^T
    r = (x @ A) - b
    berr = float(torch.linalg.vector_norm(r) / (torch.linalg.matrix_norm(A) * torch.linalg.vector_norm(x)))
    print("backward error", berr)
    assert berr < 2e-6


def test_autotune_picks_a_candidate_and_restores_knobs():
    from mnpbem_amd import autotune
    import ctypes
    banner("lu autotune n=300 batch 4")
    before = []
    for key in (0, 1, 10):
        v = ctypes.c_int(0)
        _lib().call("mnp_get_tuning", key, ctypes.addressof(v))
        before.append(v.value)
    cfg = autotune.tune_lu(300, 4, torch.complex64, _dev())
    assert {k: cfg[k] for k in ("nb", "nbo", "lookahead", "lu_recursive")} in autotune.candidates(300)
    assert cfg["ms"] > 0
    after = []
    for key in (0, 1, 10):
        v = ctypes.c_int(0)
        _lib().call("mnp_get_tuning", key, ctypes.addressof(v))
        after.append(v.value)
    assert after == before
    # a factorisation under every candidate gives LAPACK's pivots on a random matrix
    rng = np.random.default_rng(9)
    A = _rand(rng, 300, 300)
    lu, piv = sla.lu_factor(A)
    # every shape either tuner range can pick, including outer blocks wider than the matrix
    for c in autotune.candidates(300) + autotune.candidates(2000):
        t = _cm([A], torch.complex128)
        ipiv = torch.zeros((1, 300), dtype=torch.int32, device=_dev())
        info = torch.zeros(1, dtype=torch.int32, device=_dev())
        with autotune.knobs(c):
            _lib().call("mnp_getrf_batched", 0, 300, 1, t.data_ptr(), 300, 300 * 300, ipiv.data_ptr(), 300,
                        info.data_ptr(), c["nb"], _sh())
        assert np.array_equal(ipiv[0].cpu().numpy(), piv)
        assert np.max(np.abs(_np(t, 0) - lu)) < 1e-10 * np.max(np.abs(lu))
