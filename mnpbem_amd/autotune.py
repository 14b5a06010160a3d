"""Autotuner for the blocked LU of the BEMRet chain.

The batched LU (lu.hip) has three shape parameters that trade panel latency against GEMM
efficiency: the panel width ``nb`` (16 or 32 columns), the outer block width (the K of the big
trailing GEMM: 64 to 512) and lookahead (trailing update split over two streams). Which
combination is fastest depends on n, the batch and the precision. ``tune_lu`` times every
candidate once on random matrices of the problem's shape with n added to the diagonal (one
untimed call first) and keeps the fastest for the process. The BEM matrices are diagonally
dominant enough that partial pivoting interchanges no rows (G1, G2, H1 and Delta of an 840-face
gold rod: none), while a plain random matrix interchanges at almost every column, which would
time row interchanges the solves never make; ``knobs`` applies a choice around a factorisation and restores the
previous settings afterwards, so other callers keep theirs.
"""
import contextlib
import ctypes
import time

import torch

from ._lib import call

_CACHE = {}
_KEYS = {"nbo": 0, "lookahead": 1, "gemm_split": 4, "plane_tile": 5, "asm_rows": 7, "asm_lchunk": 8,
         "lu_recursive": 10}


def _value(key, v):
    """Library value of a knob: gemm_split 2 or 3 (bfloat16 terms per operand) selects the
    pre-split plane GEMM (library modes 12, 13), 16 the scaled fp16 pairs (library mode 16),
    anything else the FP32 matrix-core GEMM."""
    if key == "gemm_split":
        v = int(v)
        return 10 + v if v in (2, 3) else (16 if v == 16 else 0)
    return int(v)


def _get(key):
    v = ctypes.c_int(0)
    call("mnp_get_tuning", key, ctypes.addressof(v))
    return v.value


@contextlib.contextmanager
def knobs(cfg):
    """Set the library's knobs from cfg (keys 'nbo', 'lookahead', 'gemm_split', 'plane_tile',
    'asm_rows', 'asm_lchunk', 'lu_recursive') for the duration."""
    if not cfg:
        yield
        return
    old = {k: _get(i) for k, i in _KEYS.items() if k in cfg}
    try:
        for k, i in _KEYS.items():
            if k in cfg:
                call("mnp_set_tuning", i, _value(k, cfg[k]))
        yield
    finally:
        for k, v in old.items():
            call("mnp_set_tuning", _KEYS[k], v)


def candidates(n=0):
    """(nb, nbo, lookahead, lu_recursive) shapes to time. Wider outer blocks halve the passes of
    the trailing update over the matrix and deepen its K (split-BF16 GEMM at 64 x 2240^2: K 128 ->
    132, K 256 -> 158 TFLOP/s complex-equivalent) at the price of more in-block work, which the
    recursive order of the outer block (GEMMs of depth W/2, W/4, ... instead of W/nb GEMMs of
    depth nb) keeps small; for large n the 16-column panel never wins."""
    nbs = (32,) if n >= 1024 else (16, 32)
    nbos = (64, 128, 256, 384, 512) if n >= 1024 else (64, 128, 256)
    out = []
    for nb in nbs:
        for nbo in nbos + ((768, 1024) if n >= 2048 else ()):
            for rec in (0, 1):
                if nbo < nb or (nbo > 512 and not rec) or (rec and nbo < 2 * nb):
                    continue
                out += [dict(nb=nb, nbo=nbo, lookahead=la, lu_recursive=rec) for la in (0, 1)]
    return out


def tune_lu(n, batch, dtype, dev, reps=1, seed=0, gemm_split=0):
    """Fastest (nb, nbo, lookahead) for `batch` LUs of size n in `dtype` (complex64 or
    complex128) with the complex-float GEMM mode ``gemm_split``; cached per (n, batch, dtype,
    gemm_split). With gemm_split=3 the best shape is also timed on the FP32 matrix cores (same
    precision) and the faster kept. Returns the config (with gemm_split) and its time in ms."""
    key = (int(n), int(batch), str(dtype), int(gemm_split))
    if key in _CACHE:
        return _CACHE[key]
    prec = 0 if dtype == torch.complex128 else 1
    g = torch.Generator(device=dev).manual_seed(seed)
    A = torch.randn((batch, n, n), dtype=dtype, device=dev, generator=g)
    A.diagonal(dim1=-2, dim2=-1).add_(n)
    W = torch.empty_like(A)
    ipiv = torch.empty((batch, n), dtype=torch.int32, device=dev)
    info = torch.empty(batch, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    def timed(cfg):
        with knobs(cfg):
            times = []
            for rep in range(reps + 1):
                W.copy_(A)
                torch.cuda.synchronize(dev)
                t0 = time.perf_counter()
                call("mnp_getrf_batched", prec, n, batch, W.data_ptr(), n, n * n, ipiv.data_ptr(), n,
                     info.data_ptr(), cfg["nb"], st)
                torch.cuda.synchronize(dev)
                if rep:
                    times.append(time.perf_counter() - t0)
        return 1e3 * min(times)

    best = None
    tried = []
    for cfg in candidates(n):
        cfg = dict(cfg, gemm_split=gemm_split)
        if gemm_split in (2, 3, 16) and prec == 1:
            cfg["plane_tile"] = 11
        ms = timed(cfg)
        tried.append((ms, cfg))
        if best is None or ms < best["ms"]:
            best = dict(cfg, ms=round(ms, 3))
    if gemm_split in (2, 3, 16) and prec == 1:
        # the plane kernels with direct LDS copies (64 x 64, 128 x 64, 64 x 128 tiles) on the two
        # fastest shapes
        for _, cfg in sorted(tried, key=lambda e: e[0])[:2]:
            for tile in (111, 21, 12):
                c2 = dict(cfg, plane_tile=tile)
                ms = timed(c2)
                if ms < best["ms"]:
                    best = dict(c2, ms=round(ms, 3))
    if gemm_split in (3, 16) and prec == 1:
        # the FP32 matrix cores give the same precision: keep them where they are faster
        # (small matrices, where the split passes do not pay off)
        alt = dict(best, gemm_split=0)
        alt.pop("ms")
        alt.pop("plane_tile", None)
        ms = timed(alt)
        if ms < best["ms"]:
            best = dict(alt, ms=round(ms, 3))
    del A, W
    _CACHE[key] = best
    return best


_ASM_CACHE = {}


def asm_candidates(nlam):
    """Assembly shapes: rows per workgroup x wavelengths per thread (0 = all of the launch)."""
    chunks = [0] + [c for c in (32, 8, 2) if c < nlam]
    return [dict(asm_rows=r, asm_lchunk=c) for r in (64, 128, 256) for c in chunks]


def tune_assembly(gb, kmat, nlam, reps=1):
    """Fastest assembly shape for GreenBatch ``gb`` (its geometry, precision and batch) with
    wavenumbers kmat: every candidate assembles into gb's own matrices (overwritten by the real
    assembly afterwards), timed after one untimed call; cached per (n, nlam, precision,
    particles)."""
    g = gb.geo
    key = (int(g.n), int(nlam), str(gb.dtype), int(g.npart))
    if key in _ASM_CACHE:
        return _ASM_CACHE[key]
    dev = g.dev
    best = None
    for cfg in asm_candidates(nlam):
        with knobs(cfg):
            times = []
            for rep in range(reps + 1):
                torch.cuda.synchronize(dev)
                t0 = time.perf_counter()
                gb.assemble(kmat, nlam)
                torch.cuda.synchronize(dev)
                if rep:
                    times.append(time.perf_counter() - t0)
        ms = 1e3 * min(times)
        if best is None or ms < best["ms"]:
            best = dict(cfg, ms=round(ms, 3))
    _ASM_CACHE[key] = best
    return best
