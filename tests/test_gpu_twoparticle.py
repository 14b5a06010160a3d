"""CompGreenRet(p1, p2) between two different particles (compgreen_ret.py:57-190, :435-505):
rows are p1's faces, columns p2's, regions joined by material index, no closed-surface
correction and no 2 pi self terms (those belong to p1 is p2, compgreen_ret.py:1690-1750).

* Well separated particles have no refined pairs, so every entry is the closed form
  G = a_j e^{ikd}/d, F = n_i.(r_i - r_j)(ik - 1/d)/d^2 a_j e^{ikd}, Gp = (r_i - r_j)(...),
  computed here independently with k of the connecting material.
* Particles 1 nm apart have refined face pairs: the blocks equal the oracle's region-pair
  matrices of the composite surface (its p1-row / p2-column block)."""
import numpy as np
import pytest

from conftest import banner
from oracle import bem_oracle as O
import mnpbem_amd as M

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
KEYS = ["G", "F", "H1", "H2", "Gp", "H1p", "H2p"]


def _pair(gap):
    eps = [M.EpsConst(1.0), M.EpsDrude(*DRUDE)]
    s1 = M.trisphere(144, 20)
    s2 = M.trisphere(144, 16)
    s1.shift([20.0 / 2 + 16.0 / 2 + gap, 0.0, 0.0])
    p1 = M.ComParticle(eps, [s1], [[2, 1]], 1, interp="curv")
    p2 = M.ComParticle(eps, [s2], [[2, 1]], 1, interp="curv")
    return eps, p1, p2


def test_separated_particles_closed_form():
    eps, p1, p2 = _pair(12.0)
    lam = 600.0
    banner("two particles, separated, closed form")
    g = M.CompGreenRet(p1, p2)
    n1, n2 = p1.nfaces, p2.nfaces
    x = p1.pos[:, None, :] - p2.pos[None, :, :]
    d = np.sqrt(np.sum(x * x, axis=2))
    for (i, j), mat in (((0, 0), 2), ((1, 1), 1)):
        k = eps[mat - 1](lam)[1]
        ph = np.exp(1j * k * d)
        G = p2.area[None, :] / d * ph
        fac = (1j * k - 1 / d) / d ** 2 * p2.area[None, :] * ph
        Gp = np.transpose(x * fac[:, :, None], (0, 2, 1))
        F = np.einsum("id,idj->ij", p1.nvec, Gp)
        want = dict(G=G, F=F, H1=F, H2=F, Gp=Gp, H1p=Gp, H2p=Gp)
        for key in KEYS:
            got = g.eval(i, j, key, lam)
            assert got.shape == want[key].shape
            assert np.max(np.abs(got - want[key])) / np.max(np.abs(want[key])) < 1e-12, (i, j, key)
    # an inside region never meets the other particle's outside medium here
    assert g.eval(0, 1, "G", lam) == 0 and g.eval(1, 0, "H1", lam) == 0
    assert np.all(g.eval_block(0, 1, "Gp", lam, 3, 9) == 0)
    assert g.eval(1, 1, "G", lam).shape == (n1, n2)


@pytest.mark.parametrize("key", KEYS)
def test_close_particles_match_oracle(key):
    eps, p1, p2 = _pair(1.0)
    lam = 540.0
    banner("two particles 1 nm apart key={}".format(key))
    g = M.CompGreenRet(p1, p2)
    comb = M.ComParticle(eps, [p1.p[0], p2.p[0]], [[2, 1], [2, 1]])
    pc = comb.pc
    s = O.Surface(pc.verts, pc.faces, pc.verts2, pc.faces2)
    fp = comb.face_particle
    og = O.Green(s, fp, [[2, 1], [2, 1]], [])
    ks = [e(lam)[1] for e in eps]
    n1, n2 = p1.nfaces, p2.nfaces
    ind = np.random.default_rng(11).integers(0, n1 * n2, 300)
    vec = key in ("Gp", "H1p", "H2p")
    nref = 0
    for i, j in ((0, 0), (1, 1)):
        want = og.eval(i, j, key, ks)[:n1, ..., n1:]
        got = g.eval(i, j, key, lam)
        scale = np.max(np.abs(want))
        assert np.max(np.abs(got - want)) / scale < 1e-9, (i, j)
        blk = g.eval_block(i, j, key, lam, 40, 130)
        assert np.max(np.abs(blk - want[..., 40:130])) / scale < 1e-9, (i, j)
        picked = g.eval(i, j, key, lam, ind=ind)
        r, c = ind // n2, ind % n2
        assert np.max(np.abs(picked - (want[r, :, c] if vec else want[r, c]))) / scale < 1e-9, (i, j)
    # the refined pairs are there: near the gap the entries differ from the point formula
    rows, cols, kind = M.nearfield.refine_pairs(comb.pos, comb.pc.bradius(), comb.pos, comb.pc.bradius())
    assert np.any((rows < n1) & (cols >= n1))


@pytest.mark.parametrize("inout", [1, 2])
def test_two_particle_potential_and_field_match_oracle(inout):
    """Potentials and fields on p1's faces from surface charges and currents on p2 (a BEMRet
    solution of p2 alone), region inout of p1: G(r, s) and Hp(r, s) of the p1-row / p2-column
    blocks (no 2 pi self terms) against the oracle's matrices of the composite surface."""
    eps, p1, p2 = _pair(1.0)
    lam = 540.0
    banner("two particles potential/field inout={}".format(inout))
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    sig, _ = M.BEMRet(p2).solve(exc.potential(p2, lam))
    g = M.CompGreenRet(p1, p2)
    pot, fld = g.potential(sig, inout), g.field(sig, inout)
    comb = M.ComParticle(eps, [p1.p[0], p2.p[0]], [[2, 1], [2, 1]])
    pc = comb.pc
    og = O.Green(O.Surface(pc.verts, pc.faces, pc.verts2, pc.faces2), comb.face_particle, [[2, 1], [2, 1]], [])
    ks = [e(lam)[1] for e in eps]
    n1 = p1.nfaces
    r = inout - 1
    Hk, Pk = ("H1", "H1p") if inout == 1 else ("H2", "H2p")
    k0 = 2 * np.pi / lam
    want = dict(phi=0, phip=0, a=0, ap=0, e=0, h=0)
    for src, (s, h) in enumerate(((sig.sig1, sig.h1), (sig.sig2, sig.h2))):
        G = og.eval(r, src, "G", ks)
        if np.isscalar(G):
            continue
        G = G[:n1, n1:]
        Hm = og.eval(r, src, Hk, ks)[:n1, n1:]
        P = og.eval(r, src, Pk, ks)[:n1, :, n1:]
        want["phi"] = want["phi"] + G @ s
        want["phip"] = want["phip"] + Hm @ s
        want["a"] = want["a"] + G @ h
        want["ap"] = want["ap"] + Hm @ h
        want["e"] = want["e"] + 1j * k0 * (G @ h) - np.einsum("icj,j->ic", P, s)
        want["h"] = want["h"] + np.stack([P[:, 1] @ h[:, 2] - P[:, 2] @ h[:, 1], P[:, 2] @ h[:, 0] - P[:, 0] @ h[:, 2],
                                          P[:, 0] @ h[:, 1] - P[:, 1] @ h[:, 0]], axis=1)
    t = str(inout)
    got = dict(phi=pot["phi" + t], phip=pot["phi" + t + "p"], a=pot["a" + t], ap=pot["a" + t + "p"], e=fld.e,
               h=fld.h)
    for key, w in want.items():
        w = np.asarray(w)
        assert np.asarray(got[key]).shape == w.shape, key
        assert np.max(np.abs(got[key] - w)) / np.max(np.abs(w)) < 1e-9, key
