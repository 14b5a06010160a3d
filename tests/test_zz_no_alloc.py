"""mnp_gemm_sub_batched_noalloc: the allocation-free variant of mnp_gemm_sub_batched (operand
planes in a static device array) against the existing entry point under the same split mode,
through the C-ABI, on two calls in a row (the second smaller), plus the oversize rejection."""
import ctypes

import numpy as np
import pytest

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _lib():
    from mnpbem_amd import _lib as L
    return L


def _run(entry, terms, A, B, C, m, n, k):
    L = _lib()
    st = torch.cuda.current_stream().cuda_stream
    batch = A.shape[0]
    lda, ldb = A.shape[2], B.shape[2]
    args = (m, n, k, A.data_ptr(), lda, lda * k, B.data_ptr(), ldb, ldb * n, C.data_ptr(), m, m * n, batch, st)
    if entry == "noalloc":
        L.call("mnp_gemm_sub_batched_noalloc", terms, *args)
    else:
        old = ctypes.c_int(0)
        L.call("mnp_get_tuning", 4, ctypes.addressof(old))
        L.call("mnp_set_tuning", 4, 10 + terms)
        try:
            L.call("mnp_gemm_sub_batched", 1, *args)
        finally:
            L.call("mnp_set_tuning", 4, old.value)
    torch.cuda.synchronize()


@pytest.mark.parametrize("terms", [3, 2])
def test_noalloc_equals_allocating_entry(terms):
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(terms)
    # two calls in a row, the second smaller (the scratch keeps the first call's planes)
    for (batch, m, n, k) in ((3, 700, 650, 300), (2, 333, 250, 45)):
        banner("noalloc gemm terms={} batch={} m={} n={} k={}".format(terms, batch, m, n, k))
        A = torch.randn((batch, k, m + 2), dtype=torch.complex64, device=dev, generator=g)[:, :, :m].contiguous()
        B = torch.randn((batch, n, k), dtype=torch.complex64, device=dev, generator=g)
        C0 = torch.randn((batch, n, m), dtype=torch.complex64, device=dev, generator=g)
        C1, C2 = C0.clone(), C0.clone()
        _run("noalloc", terms, A, B, C1, m, n, k)
        _run("alloc", terms, A, B, C2, m, n, k)
        # same kernels on the same operands: bit for bit
        assert torch.equal(C1, C2)
        # and the product itself within the terms kept
        ref = C0.to(torch.complex128) - torch.matmul(A.transpose(1, 2).to(torch.complex128),
                                                     B.transpose(1, 2).to(torch.complex128)).transpose(1, 2)
        err = float((C1.to(torch.complex128) - ref).abs().max() / ref.abs().max())
        assert err < (4e-6 if terms == 3 else 1e-4)


def test_noalloc_rejects_oversize():
    L = _lib()
    nb = ctypes.c_long(0)
    L.call("mnp_noalloc_scratch_bytes", ctypes.byref(nb))
    assert nb.value == 32 << 20
    banner("noalloc oversize rejected")
    # 3 terms, k = 4096, m = n = 2048, batch 1: 2*3*256*16*(2048+2048)*2 bytes = 201 MB > 32 MiB
    with pytest.raises(ValueError, match="static scratch"):
        L.call("mnp_gemm_sub_batched_noalloc", 3, 2048, 2048, 4096, 16, 2048, 0, 16, 4096, 0, 16, 2048, 0, 1,
               None)
