"""Host-side product code against the oracle, and the C-ABI library surface. CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import bem_oracle as O
from mnpbem_amd.mesh import trisphere, trirod
from mnpbem_amd.comparticle import ComParticle
from mnpbem_amd.materials import EpsConst, EpsDrude, EpsTable
from mnpbem_amd.nearfield import NearField
from conftest import ROOT


def _cp(part, eps=None):
    eps = eps or [EpsConst(1.0), EpsDrude(10.0, 9.065, 0.0708)]
    return ComParticle(eps, [part], [[2, 1]], 1, interp="curv")


@pytest.mark.parametrize("make", [lambda: trisphere(144, 20), lambda: trirod(10.0, 40.0, [15, 15, 15]),
                                  lambda: trirod(10.0, 40.0, [12, 6, 10], triangles=True)])
def test_geometry_and_nearfield_match_oracle(make):
    cp = _cp(make())
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    assert np.allclose(p.pos, s.pos, atol=1e-13)
    assert np.allclose(p.nvec, s.nvec, atol=1e-13)
    assert np.allclose(p.area, s.area, rtol=1e-13)
    assert np.allclose(p.bradius(), s.rad, rtol=1e-13)
    nf = NearField(cp)
    g = O.Green(s, np.zeros(s.n, int), [[2, 1]], [(np.arange(s.n), np.arange(s.n), 1.0)])
    keys = sorted(g.coef)
    assert np.array_equal(nf.rows, [k[0] for k in keys]) and np.array_equal(nf.cols, [k[1] for k in keys])
    G = np.array([g.coef[k][0] for k in keys])
    F = np.array([g.coef[k][1] for k in keys])
    # self terms carry ~1e-10 conditioning noise (points 1e-7 nm from the centroid)
    assert np.max(np.abs(G - nf.g)) / np.max(np.abs(G)) < 1e-9
    assert np.max(np.abs(F - nf.f)) / np.max(np.abs(F)) < 1e-9


def test_dimer_pairs_cross_particle():
    a = trisphere(60, 10)
    b = trisphere(60, 10).shift([0, 0, 11.5])
    cp = ComParticle([EpsConst(1.0), EpsDrude(10.0, 9.065, 0.0708)], [a, b], [[2, 1], [2, 1]], 1, 2, interp="curv")
    nf = NearField(cp)
    fp = cp.face_particle
    cross = fp[nf.rows] != fp[nf.cols]
    assert cross.any()                      # gap pairs are refined
    assert np.all(nf.kind[cross] == 1)      # never as self terms


def test_materials():
    d = EpsDrude(10.0, 9.065, 0.0708)
    e, k = d(600.0)
    w = (1 / 8.0655477e-4) / 600.0
    assert np.isclose(e, 10 - 9.065 ** 2 / (w * (w + 0.0708j)))
    assert np.isclose(k, 2 * np.pi / 600 * np.sqrt(e))
    gt = EpsTable("gold.dat")
    with pytest.raises(ValueError):
        gt(100.0)
    tab = np.loadtxt(os.path.join(ROOT, "mnpbem_amd", "data", "gold.dat"), comments="%")
    lam = np.linspace(450, 900, 7)
    assert np.allclose(gt(lam)[0], O.eps_table(lam, tab)[0], rtol=1e-14)


def test_comparticle_validation():
    p = trisphere(32, 5)
    with pytest.raises(ValueError):
        ComParticle([EpsConst(1.0)], [p], [[2, 1]])
    with pytest.raises(ValueError):
        ComParticle([EpsConst(1.0), EpsConst(2.0)], [p], [[0, 1]])
    with pytest.raises(ValueError):
        ComParticle(None, [p], [[2, 1]])


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "mnpbem_hip.h")).read()
    return sorted(set(re.findall(r"^int (mnp_\w+)\(", txt, re.M)))


def test_library_exports_header():
    lib = ctypes.CDLL(os.path.join(ROOT, "mnpbem_amd", "libmnpbem_hip.so"))
    syms = _header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    from mnpbem_amd import _lib
    assert sorted(_lib.exported_symbols()) == syms
    assert lib.mnp_abi_version() == 3


def test_handles_keep_their_own_knobs_without_gpu_work():
    """Knobs belong to the handle bound to the calling thread (no device work is done)."""
    import threading
    from mnpbem_amd import _lib
    v = ctypes.c_int(0)
    _lib.call("mnp_get_tuning", 0, ctypes.addressof(v))
    default = v.value
    h = _lib.Handle()
    try:
        with h:
            _lib.call("mnp_set_tuning", 0, 512)
            _lib.call("mnp_get_tuning", 0, ctypes.addressof(v))
            assert v.value == 512
            seen = []

            def other():
                w = ctypes.c_int(0)
                _lib.call("mnp_get_tuning", 0, ctypes.addressof(w))
                seen.append(w.value)
            t = threading.Thread(target=other)
            t.start()
            t.join()
            assert seen == [128]          # another thread has its own default handle
        _lib.call("mnp_get_tuning", 0, ctypes.addressof(v))
        assert v.value == default
    finally:
        h.close()
    with pytest.raises(ValueError):
        _lib.call("mnp_handle_bind", ctypes.c_void_p(12345))      # not a handle of the library
    assert _lib.lib.mnp_release() == 0       # nothing allocated: no device call


def test_cabi_rejects_bad_arguments_without_gpu_work():
    from mnpbem_amd import _lib
    with pytest.raises(ValueError):
        _lib.call("mnp_getrf_batched", 7, 10, 1, 1, 10, 100, 1, 10, 1, 32, None)     # bad prec
    with pytest.raises(ValueError):
        _lib.call("mnp_getrs_batched", 0, 10, 9, 1, 1, 10, 100, 1, 10, 2, 10, 0, 3, 10, 0, None)  # r > 8
    with pytest.raises(ValueError):
        _lib.call("mnp_bemret_factor", 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 32, None)
    assert _lib.lib.mnp_rsolve_batched(0, 4, 4, 1, 1, 4, 16, 1, 4, 5, 4, 16, 5, 4, 16, 32, None) == -13  # B == X


def test_trace_ranges_without_gpu_work():
    """roctx ranges of host stages nest and pop without a profiler attached; a null name is
    rejected."""
    from mnpbem_amd import _lib
    with _lib.stage("gmres_iter"):
        with _lib.stage("refine"):
            pass
    assert _lib.lib.mnp_range_push(None) == -1
    assert _lib.lib.mnp_range_push(b"mnp:test") == 0
    assert _lib.lib.mnp_range_pop() == 0


def test_planewave_field_at_points_and_faces():
    """PlaneWaveRet.field (planewave_ret.py:157-240): e = pol e^{ik dir.r}, h = n_b dir x e in
    the excitation medium, zero elsewhere; for faces the chosen side selects them."""
    import mnpbem_amd as M
    cp = M.ComParticle([M.EpsConst(1.77), M.EpsDrude(10, 9.065, 0.0708)], [M.trisphere(60, 10)], [[2, 1]], 1)
    pts = np.array([[0, 0, 0.0], [0, 0, 9.0], [7.0, 1.0, -2.0]])
    pt = M.ComPoint(cp, pts)
    pw = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    f = pw.field(pt, 600.0)
    nb = np.sqrt(1.77)
    k = 2 * np.pi / 600.0 * nb
    out = pt.medium == 1
    assert np.allclose(f.e[out], np.exp(1j * k * pt.pos[out, 2])[:, None] * np.array([1.0, 0, 0]))
    assert np.allclose(f.h[out], nb * np.cross([0, 0, 1.0], f.e[out]))
    assert np.all(f.e[~out] == 0) and np.any(~out)
    fo = pw.field(cp, 600.0, inout=2)
    assert np.allclose(fo.e[:, 0], np.exp(1j * k * cp.pos[:, 2]))
    assert np.all(pw.field(cp, 600.0, inout=1).e == 0)


def test_compoint_media_and_grid_expansion():
    """ComPoint media against the oracle's rule (nearest face over all particles, outside when
    (x - x_face) . n >= 0), mindist drops points near the surface, medium filter, ordering by
    medium with the original indices; MeshField grid expansion as meshfield.py:expand."""
    import mnpbem_amd as M
    from mnpbem_amd.compoint import _expand_grid
    from oracle import bem_oracle as O
    a = M.trisphere(60, 10)
    b = M.trisphere(60, 10).shift([0, 0, 11.0])
    cp = M.ComParticle([M.EpsConst(1.33), M.EpsDrude(10, 9.065, 0.0708), M.EpsConst(2.0)], [a, b],
                       [[2, 1], [3, 1]], 1, 2)
    rng = np.random.default_rng(5)
    pts = np.column_stack([rng.uniform(-8, 8, 200), rng.uniform(-8, 8, 200), rng.uniform(-8, 19, 200)])
    pt = M.ComPoint(cp, pts)
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    want = O.point_media(s, cp.face_particle, [[2, 1], [3, 1]], pts)
    assert np.array_equal(pt.medium, want[pt.index])
    assert sorted(pt.index.tolist()) == list(range(200))
    assert np.all(np.diff(pt.medium) >= 0) and set(pt.inout) == {1, 2, 3}
    assert np.array_equal(pt.pos, pts[pt.index])
    near = M.ComPoint(cp, pts, mindist=1.5)
    dmin = np.min(np.linalg.norm(pts[:, None, :] - cp.pos[None, :, :], axis=2), axis=1)
    assert set(near.index.tolist()) == set(np.nonzero(dmin >= 1.5)[0].tolist())
    only3 = M.ComPoint(cp, pts, medium=3)
    assert np.all(only3.medium == 3) and only3.n == np.sum(want == 3)
    x, z = np.meshgrid(np.linspace(-1, 1, 5), np.linspace(-2, 2, 4))
    gx, gy, gz = _expand_grid(x, 0.0, z)
    assert gx.shape == gy.shape == gz.shape == x.shape and np.all(gy == 0)
    gx, gy, gz = _expand_grid(x, z, np.array([0.0, 1.0, 2.0]))
    assert gx.shape == (4, 5, 3) and np.array_equal(gz[1, 2], [0.0, 1.0, 2.0])


def test_singular_factor_warns():
    """Zero pivots (LAPACK info > 0) raise a SingularMatrixWarning, as scipy's LinAlgWarning."""
    from mnpbem_amd.engine import SingularMatrixWarning, warn_singular
    with pytest.warns(SingularMatrixWarning, match="column 4"):
        warn_singular("test", np.array([[0, 5], [0, 0]]))


def test_particles64_face_counts_realised():
    """The 64-particle workload draws face counts uniform in 300-3000 and the meshes realise
    them (a sphere mesh of m longitudes and t bands has exactly 2 m t faces)."""
    import bench
    rng = np.random.default_rng(20261203)
    want = []
    for _ in range(64):
        kind = ("sphere", "ellipsoid", "dimer")[int(rng.integers(3))]
        want.append((kind, int(rng.integers(300, 3001))))
        if kind != "sphere":
            rng.uniform(1, 5)
    got = bench.particles64()
    assert len(got) == 64
    for (kind, cp), (k2, nf) in zip(got, want):
        assert kind == k2
        assert abs(cp.nfaces - nf) <= max(0.03 * nf, 16), (kind, cp.nfaces, nf)
    assert {k for k, _ in got} == {"sphere", "ellipsoid", "dimer"}


def test_noalloc_gemm_rejects_oversize_without_gpu_work():
    """The allocation-free GEMM checks its plane size against the static scratch before any
    device call."""
    from mnpbem_amd import _lib
    nb = ctypes.c_long(0)
    _lib.call("mnp_noalloc_scratch_bytes", ctypes.byref(nb))
    assert nb.value == 32 << 20
    st = _lib.lib.mnp_gemm_sub_batched_noalloc(3, 2048, 2048, 4096, 16, 2048, 0, 16, 4096, 0, 16, 2048, 0, 1, None)
    assert st == 2
    assert _lib.lib.mnp_gemm_sub_batched_noalloc(4, 10, 10, 10, 16, 10, 0, 16, 10, 0, 16, 10, 0, 1, None) == -1
