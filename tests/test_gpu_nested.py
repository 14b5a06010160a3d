"""Nested particles: a gold core inside a dielectric shell. The device assembly forms the
cross-region blocks of G1 = G(in, in) - G(out, in) and G2 = G(out, out) - G(in, out) (H1, H2
alike; bem_ret.py:407-433, materials joining the regions from compgreen_ret.py:435-505); checked
against the oracle's region-pair matrices, then the BEMRet solutions (complex128 and mixed
precision) against the oracle's chain, and BEMRetIter (dense and H-matrix operators) against
BEMRet."""
import numpy as np
import pytest
import torch

from conftest import banner
from oracle import bem_oracle as O
import mnpbem_amd as M
from mnpbem_amd.bem_iter import BEMRetIter
from mnpbem_amd.engine import GreenBatch

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
INOUT = [[2, 1], [3, 2]]          # shell surface: shell | vacuum; core surface: gold | shell


@pytest.fixture(scope="module")
def coreshell():
    shell = M.trisphere(144, 20)
    core = M.trisphere(144, 12)
    eps = [M.EpsConst(1.0), M.EpsConst(2.1), M.EpsDrude(*DRUDE)]
    cp = M.ComParticle(eps, [shell, core], INOUT, 1, 2, interp="curv")
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    fp = cp.face_particle
    closed = [(np.nonzero(fp == a)[0], np.nonzero(fp == a)[0], 1.0) for a in range(2)]
    return cp, s, fp, closed


@pytest.fixture(scope="module")
def coreshell_oracle(coreshell):
    cp, s, fp, closed = coreshell
    return O.Green(s, fp, INOUT, closed)


def _eps_funcs(cp):
    return [lambda lam, e=e: e(lam) for e in cp.eps]


def test_nested_assembly_matches_oracle(coreshell, coreshell_oracle):
    cp, s, fp, closed = coreshell
    bem = M.BEMRet(cp)
    geo = bem.g.geo
    assert geo.nested
    lams = np.array([520.0, 700.0])
    banner("nested assembly n={}".format(s.n))
    gb = GreenBatch(geo, 2, torch.complex128)
    gb.assemble(geo.media(lams)[0], 2)
    core = fp == 1
    shell = fp == 0
    for b, lam in enumerate(lams):
        ks = [e(lam)[1] for e in cp.eps]
        for name, t, R in zip(("G1", "G2", "H1", "H2"), (gb.G1, gb.G2, gb.H1, gb.H2), O.matrices(coreshell_oracle, ks)):
            got = GreenBatch.to_numpy(t, b)
            assert np.max(np.abs(got - R)) / np.max(np.abs(R)) < 1e-9, name
        # the cross blocks are there: G1 (core rows, shell columns) is minus the shell-medium
        # Green function, G2 (shell rows, core columns) likewise; the other off-diagonal blocks
        # vanish (vacuum does not reach the core, gold not the shell)
        G1, G2 = GreenBatch.to_numpy(gb.G1, b), GreenBatch.to_numpy(gb.G2, b)
        assert np.max(np.abs(G1[np.ix_(core, shell)])) > 0 and np.max(np.abs(G2[np.ix_(shell, core)])) > 0
        assert np.max(np.abs(G1[np.ix_(shell, core)])) == 0 and np.max(np.abs(G2[np.ix_(core, shell)])) == 0


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("mixed", 1e-8)])
def test_nested_bemret_matches_oracle(coreshell, precision, tol):
    cp, s, fp, closed = coreshell
    lams = np.array([480.0, 530.0, 620.0])
    banner("nested bemret {} spectrum".format(precision))
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    out = M.BEMRet(cp, batch=3, precision=precision, autotune=False).solve_spectrum(exc, lams, scattering=True)
    ref = O.solve_spectrum(s, fp, INOUT, _eps_funcs(cp), lams, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]),
                           pinf_nvec=exc.spec.nvec, pinf_area=exc.spec.area)
    assert np.max(np.abs(out["ext"][:, 0] - ref["ext"]) / np.abs(ref["ext"])) < tol
    assert np.max(np.abs(out["sca"][:, 0] - ref["sca"]) / np.abs(ref["sca"])) < tol


def test_nested_bemret_solution_and_per_wavelength_api(coreshell, coreshell_oracle):
    """sig1, h1, sig2, h2 of BEMRet.solve (the per-wavelength API) against the oracle's chain."""
    cp, s, fp, closed = coreshell
    lam = 530.0
    banner("nested bemret solution lam={}".format(lam))
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    sig, _ = M.BEMRet(cp).solve(exc.potential(cp, lam))
    ks = [e(lam)[1] for e in cp.eps]
    G1, G2, H1, H2 = O.matrices(coreshell_oracle, ks)
    e1, e2 = cp.eps1(lam), cp.eps2(lam)
    ob = O.BEMRetOracle(G1, G2, H1, H2, s.nvec, e1, e2, 2 * np.pi / lam, cross=True)
    eps_out, k_out = cp.eps[0](lam)
    io = np.asarray(INOUT)[fp]
    a2, a2p = O.planewave(s.pos, s.nvec, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), lam, eps_out,
                          np.nonzero(io[:, 1] == 1)[0])
    a1, a1p = O.planewave(s.pos, s.nvec, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), lam, eps_out,
                          np.nonzero(io[:, 0] == 1)[0])
    phi, a, alpha, De = O.excitation(s.nvec, 2 * np.pi / lam, e1, e2, a1, a1p, a2, a2p)
    s1, h1, s2, h2 = ob.solve(phi, a, alpha, De)
    for got, want in ((sig.sig1, s1), (sig.h1, h1), (sig.sig2, s2), (sig.h2, h2)):
        got = np.asarray(got).reshape(want.shape)
        assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-10


@pytest.mark.parametrize("hmatrix", [False, True])
def test_nested_bemretiter_equals_bemret(coreshell, hmatrix):
    cp, s, fp, closed = coreshell
    lam = 530.0
    banner("nested bemretiter hmatrix={}".format(hmatrix))
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    pot = exc.potential(cp, lam)
    ref, _ = M.BEMRet(cp).solve(pot)
    op = dict(tol=1e-10, maxit=400, restart=60)
    if hmatrix:
        op.update(hmatrix=True, htol=1e-10, kmax=200, cleaf=32)
    it = BEMRetIter(cp, **op)
    sig, _ = it.solve(pot)
    flag, relres, niter = it.info()
    assert flag == 0
    e_ref, e_it = exc.extinction(ref), exc.extinction(sig)
    assert abs(e_it - e_ref) / abs(e_ref) < (1e-6 if hmatrix else 1e-8)


@pytest.mark.parametrize("key", ["G", "F", "H1", "H2", "Gp", "H1p", "H2p"])
def test_nested_region_pair_blocks_match_oracle(coreshell, coreshell_oracle, key):
    """CompGreenRet.eval(i, j, key) of every region pair of the core-shell particle, the full
    matrix (assembly kernel), column blocks and picked entries (entry-wise kernel), against the
    oracle's region-pair matrices (compgreen_ret.py:507-737, :860-910). The cross pairs carry the
    shell medium between the shell's inside and the core's outside; pairs that do not connect
    are 0."""
    cp, s, fp, closed = coreshell
    lam = 560.0
    banner("nested region-pair blocks key={}".format(key))
    g = M.CompGreenRet(cp, cp)
    ks = [e(lam)[1] for e in cp.eps]
    n = s.n
    ind = np.random.default_rng(5).integers(0, n * n, 400)
    vec = key in ("Gp", "H1p", "H2p")
    for i in range(2):
        for j in range(2):
            want = coreshell_oracle.eval(i, j, key, ks)
            got = g.eval(i, j, key, lam)
            if np.isscalar(want):
                assert np.isscalar(got) and got == 0
                continue
            scale = np.max(np.abs(want))
            assert np.max(np.abs(got - want)) / scale < 1e-9, (i, j)
            blk = g.eval_block(i, j, key, lam, 100, 220)
            assert np.max(np.abs(blk - want[..., 100:220])) / scale < 1e-9, (i, j)
            picked = g.eval(i, j, key, lam, ind=ind)
            r, c = ind // n, ind % n
            want_p = want[r, :, c] if vec else want[r, c]
            assert np.max(np.abs(picked - want_p)) / scale < 1e-9, (i, j)


def _fields_from_matrices(og, ks, k0, sig, inout):
    """phi, phip, a, ap, e, h of region inout from the oracle's region-pair matrices
    (compgreen_ret.py:911-1034): both source regions contribute."""
    r = inout - 1
    H = "H1" if inout == 1 else "H2"
    Hp = "H1p" if inout == 1 else "H2p"
    srcs = ((np.asarray(sig.sig1), np.asarray(sig.h1)), (np.asarray(sig.sig2), np.asarray(sig.h2)))
    out = dict(phi=0, phip=0, a=0, ap=0, e=0, h=0)
    for src, (s, h) in enumerate(srcs):
        G = og.eval(r, src, "G", ks)
        if np.isscalar(G):
            continue
        Hm, P = og.eval(r, src, H, ks), og.eval(r, src, Hp, ks)
        out["phi"] = out["phi"] + G @ s
        out["phip"] = out["phip"] + Hm @ s
        out["a"] = out["a"] + G @ h
        out["ap"] = out["ap"] + Hm @ h
        out["e"] = out["e"] + 1j * k0 * (G @ h) - np.einsum("icj,j->ic", P, s)
        # h_i = sum_j Hp_ij x h_j
        out["h"] = out["h"] + np.stack([P[:, 1] @ h[:, 2] - P[:, 2] @ h[:, 1],
                                        P[:, 2] @ h[:, 0] - P[:, 0] @ h[:, 2],
                                        P[:, 0] @ h[:, 1] - P[:, 1] @ h[:, 0]], axis=1)
    return out


@pytest.mark.parametrize("inout", [1, 2])
def test_nested_surface_potential_and_field_match_oracle(coreshell, coreshell_oracle, inout):
    """CompGreenRet.potential / .field on the faces of the core-shell particle from its BEMRet
    solution: the sources of both regions contribute (the shell medium joins the shell's inside
    and the core's outside), against the oracle's region-pair matrices."""
    cp, s, fp, closed = coreshell
    lam = 530.0
    banner("nested surface fields inout={}".format(inout))
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    sig, bem = M.BEMRet(cp).solve(exc.potential(cp, lam))
    g = M.CompGreenRet(cp, cp)
    pot = g.potential(sig, inout)
    fld = g.field(sig, inout)
    ks = [e(lam)[1] for e in cp.eps]
    want = _fields_from_matrices(coreshell_oracle, ks, 2 * np.pi / lam, sig, inout)
    t = str(inout)
    got = dict(phi=pot["phi" + t], phip=pot["phi" + t + "p"], a=pot["a" + t], ap=pot["a" + t + "p"], e=fld.e,
               h=fld.h)
    for key in want:
        w = np.asarray(want[key])
        assert np.max(np.abs(np.asarray(got[key]).reshape(w.shape) - w)) / np.max(np.abs(w)) < 1e-9, key


def test_nested_point_fields_match_oracle(coreshell):
    """Fields and potentials at points in the core, the shell and outside (ComPoint,
    CompGreenRet(pt, p)): a point in the shell medium sees the shell faces through sig1, h1 and
    the core faces through sig2, h2; against the oracle's point fields."""
    cp, s, fp, closed = coreshell
    lam = 560.0
    rng = np.random.default_rng(9)
    d = rng.standard_normal((36, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rr = np.concatenate([rng.uniform(1.0, 4.5, 12), rng.uniform(6.8, 9.0, 12), rng.uniform(11.0, 30.0, 12)])
    pts = d * rr[:, None]
    banner("nested point fields npt={}".format(len(pts)))
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    sig, _ = M.BEMRet(cp).solve(exc.potential(cp, lam))
    pt = M.ComPoint(cp, pts)
    assert set(np.unique(pt.medium)) == {1, 2, 3}
    assert np.array_equal(pt.medium, O.point_media(s, fp, INOUT, pt.pos))
    g = M.CompGreenRet(pt, cp)
    fld, pot = g.field(sig), g.potential(sig)
    ks = [e(lam)[1] for e in cp.eps]
    ref = O.point_potential_field(s, fp, INOUT, pt.pos, ks, lam, sig.sig1, sig.h1, sig.sig2, sig.h2)
    for got, want in ((fld.e, ref["e"]), (fld.h, ref["h"]), (pot.phi1, ref["phi"]), (pot.a1, ref["a"])):
        assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-10
