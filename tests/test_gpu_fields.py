"""Surface potentials and fields of a BEMRet solution on the GPU (CompGreenRet.potential /
.field, compgreen_ret.py:911-1034) against the oracle's dense evaluation."""
import os

import numpy as np
import pytest

from conftest import ROOT, banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import bem_oracle as O                                       # noqa: E402
import mnpbem_amd as M                                                   # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)
GOLD = np.loadtxt(os.path.join(ROOT, "mnpbem_amd", "data", "gold.dat"), comments="%")


def _setup(parts, inout, eps):
    cp = M.ComParticle(eps, parts, inout, *range(1, len(parts) + 1), interp="curv")
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    fp = cp.face_particle
    closed = [(np.nonzero(fp == a)[0], np.nonzero(fp == a)[0], 1.0) for a in range(len(parts))]
    return cp, O.Green(s, fp, inout, closed)


def _check(cp, green, lam, pol, direc, tol=1e-10):
    bem = M.BEMRet(cp)
    exc = M.PlaneWaveRet(pol, direc)
    sig, _ = bem.solve(exc.potential(cp, lam))
    ks = [e(lam)[1] for e in cp.eps]
    for inout in (1, 2):
        ref = O.surface_potential_field(green, ks, lam, sig.sig1, sig.h1, sig.sig2, sig.h2, inout)
        pot = bem.potential(sig, inout)
        fld = bem.field(sig, inout)
        t = str(inout)
        got = dict(phi=pot["phi" + t], phip=pot["phi" + t + "p"], a=pot["a" + t], ap=pot["a" + t + "p"],
                   e=fld.e, h=fld.h)
        for key, want in ref.items():
            err = np.max(np.abs(got[key] - want)) / np.max(np.abs(want))
            assert err < tol, (inout, key, err)
    return sig


def test_fields_sphere_matches_oracle():
    cp, green = _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsTable("gold.dat")])
    banner("surface fields sphere n={}".format(cp.nfaces))
    for lam in (450.0, 520.0, 700.0):
        _check(cp, green, lam, [1.0, 0, 0], [0, 0, 1.0])


def test_fields_dimer_two_polarisations():
    a = M.trisphere(60, 10)
    b = M.trisphere(60, 10).shift([0, 0, 11.0])
    cp, green = _setup([a, b], [[2, 1], [2, 1]], [M.EpsConst(1.33), M.EpsDrude(*DRUDE)])
    banner("surface fields dimer n={} npol=2".format(cp.nfaces))
    lam = 650.0
    bem = M.BEMRet(cp)
    exc = M.PlaneWaveRet(np.array([[1.0, 0, 0], [0, 0, 1.0]]), np.array([[0, 0, 1.0], [1.0, 0, 0]]))
    sig, _ = bem.solve(exc.potential(cp, lam))
    ks = [e(lam)[1] for e in cp.eps]
    fld = bem.field(sig)                                   # outside by default
    pot = bem.potential(sig, 1)
    for ip in range(2):
        ref2 = O.surface_potential_field(green, ks, lam, sig.sig1[:, ip], sig.h1[:, :, ip], sig.sig2[:, ip],
                                         sig.h2[:, :, ip], 2)
        ref1 = O.surface_potential_field(green, ks, lam, sig.sig1[:, ip], sig.h1[:, :, ip], sig.sig2[:, ip],
                                         sig.h2[:, :, ip], 1)
        for got, want in ((fld.e[:, :, ip], ref2["e"]), (fld.h[:, :, ip], ref2["h"]),
                          (pot.phi1[:, ip], ref1["phi"]), (pot.a1p[:, :, ip], ref1["ap"])):
            assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-10


def test_fields_satisfy_boundary_conditions():
    """Device fields alone: normal D is continuous to solver accuracy, tangential E to the
    discretisation error."""
    cp, green = _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsTable("gold.dat")])
    lam = 520.0
    sig = _check(cp, green, lam, [1.0, 0, 0], [0, 0, 1.0])
    bem = M.BEMRet(cp)
    e1, e2 = bem.field(sig, 1).e, bem.field(sig, 2).e
    n = cp.nvec
    k = cp.eps[0](lam)[1]
    e_inc = np.exp(1j * k * cp.pos[:, 2])[:, None] * np.array([1.0, 0, 0])
    dn1 = cp.eps1(lam) * np.sum(n * e1, axis=1)
    dn2 = cp.eps2(lam) * np.sum(n * (e2 + e_inc), axis=1)
    assert np.linalg.norm(dn1 - dn2) < 1e-8 * np.linalg.norm(dn1)
    assert np.linalg.norm(np.cross(n, e1 - e2 - e_inc)) < 0.01 * np.linalg.norm(e_inc)


@pytest.mark.parametrize("parts", ["sphere", "dimer"])
def test_eval_keys_blocks_and_indices(parts):
    """CompGreenRet.eval for all seven keys (G, F, H1, H2 from the assembly; Gp, H1p, H2p
    entry-wise) against the oracle; eval_block = column slices of eval; eval(..., ind) =
    picked entries; regions that do not connect give 0 / zeros."""
    if parts == "sphere":
        cp, green = _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsTable("gold.dat")])
    else:
        a = M.trisphere(60, 10)
        b = M.trisphere(60, 10).shift([0, 0, 11.0])
        cp, green = _setup([a, b], [[2, 1], [2, 1]], [M.EpsConst(1.33), M.EpsDrude(*DRUDE)])
    lam = 580.0
    banner("eval keys {} n={}".format(parts, cp.nfaces))
    g = M.CompGreenRet(cp, cp)
    ks = [e(lam)[1] for e in cp.eps]
    n = cp.nfaces
    rng = np.random.default_rng(3)
    ind = rng.integers(0, n * n, 500)
    ind[:n // 4] = np.arange(n // 4) * (n + 1)                   # include self terms
    for r in (0, 1):
        for key in ("G", "F", "H1", "H2", "Gp", "H1p", "H2p"):
            want = green.eval(r, r, key, ks)
            got = g.eval(r, r, key, lam)
            assert got.shape == want.shape
            assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-9, (r, key)
            blk = g.eval_block(r, r, key, lam, 37, 151)
            assert np.max(np.abs(blk - got[..., 37:151])) <= 1e-12 * np.max(np.abs(got))
            pick = g.eval(r, r, key, lam, ind=ind)
            full = got if got.ndim == 2 else np.transpose(got, (0, 2, 1))
            ref = full.reshape(n * n, -1)[ind].reshape(pick.shape)
            assert np.max(np.abs(pick - ref)) <= 1e-12 * np.max(np.abs(got))
    assert g.eval(0, 1, "G", lam) == 0
    assert not np.any(g.eval_block(1, 0, "H1p", lam, 0, 10))
    with pytest.raises(AssertionError):
        g.eval_block(0, 0, "G", lam, 5, n + 1)
    with pytest.raises(ValueError):
        g.eval(0, 0, "Q", lam)


@pytest.mark.parametrize("case", ["sphere", "dimer"])
def test_point_fields_match_oracle(case):
    """Fields and potentials at points (ComPoint, CompGreenRet(pt, p)) against the oracle:
    points inside and outside, some within the refinement cutoff of a face, two polarisations
    for the dimer; media of the points as the oracle's."""
    rng = np.random.default_rng(7)
    if case == "sphere":
        cp, green = _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsTable("gold.dat")])
        d = rng.standard_normal((40, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        rr = np.concatenate([rng.uniform(10.6, 14.0, 20), rng.uniform(15.0, 60.0, 12), rng.uniform(1.0, 8.0, 8)])
        pts = d * rr[:, None]
        exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
        inout = [[2, 1]]
    else:
        a = M.trisphere(60, 10)
        b = M.trisphere(60, 10).shift([0, 0, 11.0])
        cp, green = _setup([a, b], [[2, 1], [2, 1]], [M.EpsConst(1.33), M.EpsDrude(*DRUDE)])
        pts = np.column_stack([rng.uniform(-3, 3, 30), rng.uniform(-3, 3, 30), rng.uniform(4.0, 7.0, 30)])
        pts = np.vstack([pts, [[0, 0, -12.0], [8.0, 0, 5.5], [0, 0, 0.5], [0, 0, 11.0]]])
        exc = M.PlaneWaveRet(np.array([[1.0, 0, 0], [0, 0, 1.0]]), np.array([[0, 0, 1.0], [1.0, 0, 0]]))
        inout = [[2, 1], [2, 1]]
    lam = 600.0
    banner("point fields {} npt={}".format(case, len(pts)))
    sig, _ = M.BEMRet(cp).solve(exc.potential(cp, lam))
    pt = M.ComPoint(cp, pts)
    g = M.CompGreenRet(pt, cp)
    assert g._points.n_ref > 0                       # some pairs are refined
    fld, pot = g.field(sig), g.potential(sig)
    ks = [e(lam)[1] for e in cp.eps]
    s = green.s
    assert np.array_equal(pt.medium, O.point_media(s, green.fp, inout, pt.pos))
    npol = 1 if np.ndim(sig.sig1) == 1 else sig.sig1.shape[1]
    for ip in range(npol):
        col = (lambda v: v) if npol == 1 else (lambda v: v[..., ip])
        ref = O.point_potential_field(s, green.fp, inout, pt.pos, ks, lam, col(sig.sig1), col(sig.h1),
                                      col(sig.sig2), col(sig.h2))
        for got, want in ((col(fld.e), ref["e"]), (col(fld.h), ref["h"]), (col(pot.phi1), ref["phi"]),
                          (col(pot.a1), ref["a"])):
            assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-10


def test_meshfield_grid():
    """MeshField (sim='ret'): grid + (3,) fields equal CompGreenRet(ComPoint).field at the kept
    points, NaN where mindist drops points; grid expansion of an (x, z) plane at y = 0."""
    cp, green = _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsTable("gold.dat")])
    lam = 520.0
    sig, _ = M.BEMRet(cp).solve(M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0]).potential(cp, lam))
    x, z = np.meshgrid(np.linspace(-20, 20, 17), np.linspace(-20, 20, 13))
    banner("meshfield {}".format(x.shape))
    mf = M.MeshField(cp, x, 0.0, z, mindist=0.5, sim="ret")
    e, h = mf(sig)
    assert e.shape == x.shape + (3,) and h.shape == x.shape + (3,)
    pos = np.column_stack([x.ravel(), np.zeros(x.size), z.ravel()])
    pt = M.ComPoint(cp, pos, mindist=0.5)
    ref = M.CompGreenRet(pt, cp).field(sig)
    ef = e.reshape(-1, 3)
    assert np.allclose(ef[pt.index], ref.e, rtol=0, atol=1e-12 * np.abs(ref.e).max())
    dropped = np.setdiff1d(np.arange(x.size), pt.index)
    assert np.all(np.isnan(ef[dropped]))
    r = np.linalg.norm(pos, axis=1)
    assert np.all(np.abs(r[dropped] - 10.0) < 1.0)       # only points next to the surface are dropped
    with pytest.raises(ValueError):
        M.MeshField(cp, x, 0.0, z, sim="layer")


def test_stat_point_fields():
    """Quasistatic fields at points (CompGreenStat(ComPoint, p).field = -Gp sig with G = 1/r,
    compgreen_stat.py:974-1050): against the oracle at k = 0 (1e-10), and far away against
    the field of the induced dipole p = sum a_j r_j sig_j, (3 (p.n) n - p) / R^3, to the
    quadrupole order (size / R)."""
    cp, green = _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsTable("gold.dat")])
    lam = 520.0
    sig, _ = M.BEMStat(cp).solve(M.PlaneWaveStat([1.0, 0, 0])(cp, lam))
    rng = np.random.default_rng(11)
    d = rng.standard_normal((30, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = d * np.concatenate([rng.uniform(10.5, 14.0, 15), rng.uniform(1.0, 8.0, 10), rng.uniform(20, 40, 5)])[:, None]
    banner("stat point fields npt={}".format(len(pts)))
    pt = M.ComPoint(cp, pts)
    g = M.CompGreenStat(pt, cp)
    e = g.field(sig).e
    phi = g.potential(sig).phi1
    s = green.s
    ref = O.point_potential_field(s, green.fp, [[2, 1]], pt.pos, [0.0, 0.0], lam, sig.sig, np.zeros((s.n, 3)),
                                  sig.sig, np.zeros((s.n, 3)))
    assert np.max(np.abs(e - ref["e"])) / np.max(np.abs(ref["e"])) < 1e-10
    assert np.max(np.abs(phi - ref["phi"])) / np.max(np.abs(ref["phi"])) < 1e-10
    R = 1e4
    far = M.ComPoint(cp, R * d[:8])
    ef = M.CompGreenStat(far, cp).field(sig).e
    pdip = (cp.area[:, None] * cp.pos).T @ sig.sig
    nn = far.pos / R
    edip = (3 * nn * (nn @ pdip)[:, None] - pdip[None, :]) / R ** 3
    assert np.linalg.norm(ef - edip) / np.linalg.norm(edip) < 30.0 / R
    x, z = np.meshgrid(np.linspace(-20, 20, 9), np.linspace(-20, 20, 7))
    em, hm = M.MeshField(cp, x, 0.0, z, mindist=0.5)(sig)
    assert em.shape == x.shape + (3,) and hm is None


def test_points_empty_and_single():
    """Edge cases: a ComPoint whose points are all dropped (mindist) gives empty fields, and a
    single point far away works for both Green functions."""
    cp, green = _setup([M.trisphere(60, 10)], [[2, 1]], [M.EpsConst(1.0), M.EpsDrude(*DRUDE)])
    lam = 600.0
    sig, _ = M.BEMRet(cp).solve(M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0]).potential(cp, lam))
    banner("points empty / single")
    empty = M.ComPoint(cp, cp.pos[:5] * 1.0001, mindist=1.0)
    assert empty.n == 0
    f = M.CompGreenRet(empty, cp).field(sig)
    assert f.e.shape == (0, 3) and f.h.shape == (0, 3)
    one = M.ComPoint(cp, [[0.0, 0.0, 50.0]])
    f1 = M.CompGreenRet(one, cp).field(sig)
    assert f1.e.shape == (1, 3) and np.all(np.isfinite(f1.e)) and np.abs(f1.e).max() > 0
    sst, _ = M.BEMStat(cp).solve(M.PlaneWaveStat([1.0, 0, 0])(cp, lam))
    assert M.CompGreenStat(empty, cp).field(sst).e.shape == (0, 3)


def test_bemstat_surface_field():
    """BEMStat.field (bem_stat.py:863-935, e = -H1p/H2p sig) against the oracle at k = 0, and
    normal D continuity across the surface with the incident field on both sides (the
    quasistatic BEM equation -(Lambda + F) sig = phi' is eps1 n.(e1 + e_inc) = eps2 n.(e2 +
    e_inc) with the external potential present everywhere; exact)."""
    cp, green = _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsTable("gold.dat")])
    lam = 520.0
    bem = M.BEMStat(cp)
    pol = np.array([1.0, 0, 0])
    sig, _ = bem.solve(M.PlaneWaveStat(pol)(cp, lam))
    banner("bemstat surface field")
    z = np.zeros((cp.nfaces, 3))
    for inout in (1, 2):
        e = bem.field(sig, inout).e
        ref = O.surface_potential_field(green, [0.0, 0.0], lam, sig.sig, z, sig.sig, z, inout)
        assert np.max(np.abs(e - ref["e"])) / np.max(np.abs(ref["e"])) < 1e-10
    e1, e2 = bem.field(sig, 1).e, bem.field(sig, 2).e
    n = cp.nvec
    d1 = cp.eps1(lam) * np.sum(n * (e1 + pol[None, :]), axis=1)
    d2 = cp.eps2(lam) * np.sum(n * (e2 + pol[None, :]), axis=1)
    assert np.linalg.norm(d1 - d2) < 1e-8 * np.linalg.norm(d1)
    it = M.BEMStatIter(cp)
    assert np.allclose(it.field(sig, 2).e, e2, rtol=0, atol=1e-12 * np.abs(e2).max())
