"""The oracle's nested-particle path (a core inside a shell), pinned on the CPU.

* The coated-sphere Mie series against its limits: equal core and shell media, or a vanishing
  core, give the homogeneous sphere's series (itself pinned by the reference's Mie table), and a
  shell of the outside medium gives the bare core.
* The oracle's BEM chain for a gold-Drude core (trisphere(144, 12)) in a shell of eps 2.1
  (trisphere(144, 20)) against the reference's own solution of the same mesh
  (tests/golden/reference_coreshell.npz, written by make_reference_fixtures.py): the region-pair
  matrices G1, G2, H1, H2 (diagonals and one shell and one core row), sigma1, h1, sigma2, h2 and
  the extinction. A chain with the diagonal eps instead of L = G eps G^-1, or the operator with
  eps after the Green function, misses them by factors (tested below on the oracle itself).
* The same chain against the coated-sphere Mie series (discretisation error of the 144-vertex
  spheres: within 10 %), and a shell of the outside medium against the bare core's series.
"""
import os

import numpy as np
import pytest

from oracle import bem_oracle as O
import mnpbem_amd as M

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DRUDE = (10.0, 9.065, 0.0708)
INOUT = [[2, 1], [3, 2]]


def _coreshell(eps_shell):
    eps = [M.EpsConst(1.0), M.EpsConst(eps_shell), M.EpsDrude(*DRUDE)]
    cp = M.ComParticle(eps, [M.trisphere(144, 20), M.trisphere(144, 12)], INOUT, 1, 2, interp="curv")
    p = cp.pc
    return cp, O.Surface(p.verts, p.faces, p.verts2, p.faces2)


@pytest.fixture(scope="module")
def fixture():
    return np.load(os.path.join(GOLDEN, "reference_coreshell.npz"))


def test_coated_mie_limits():
    drude = M.EpsDrude(*DRUDE)
    for lam in (480.0, 530.0, 700.0):
        ed = drude(lam)[0]
        homog = O.mie_sphere(lam, 20.0, ed, 1.0)
        assert np.allclose(O.mie_coated(lam, 12.0, 20.0, ed, ed, 1.0), homog, rtol=1e-12)
        assert np.allclose(O.mie_coated(lam, 1e-6, 20.0, 2.1, ed, 1.0), homog, rtol=1e-12)
        assert np.allclose(O.mie_coated(lam, 12.0, 20.0, ed, 1.0, 1.0), O.mie_sphere(lam, 12.0, ed, 1.0), rtol=1e-12)


def test_oracle_coreshell_matches_reference(fixture):
    cp, s = _coreshell(2.1)
    fp = cp.face_particle
    closed = [(np.nonzero(fp == a)[0], np.nonzero(fp == a)[0], 1.0) for a in range(2)]
    g = O.Green(s, fp, INOUT, closed)
    ks = [e(480.0)[1] for e in cp.eps]
    for name, m in zip(("G1", "G2", "H1", "H2"), O.matrices(g, ks)):
        d, rows = fixture["mat_diag_" + name], fixture["mat_rows_" + name]
        assert np.max(np.abs(np.diag(m) - d)) / np.max(np.abs(d)) < 1e-9, name
        assert np.max(np.abs(m[[0, 400]] - rows)) / np.max(np.abs(rows)) < 1e-9, name
    lams = fixture["lams"]
    out = O.solve_spectrum(s, fp, INOUT, [lambda lam, e=e: e(lam) for e in cp.eps], lams, np.array([1.0, 0, 0]),
                           np.array([0, 0, 1.0]), return_fields=True)
    assert np.max(np.abs(out["ext"] - fixture["ext"]) / fixture["ext"]) < 1e-9
    for b in range(len(lams)):
        sol = out["fields"][b]["sol"]
        for got, key in zip(sol, ("sig1", "h1", "sig2", "h2")):
            want = fixture[key][b]
            got = np.asarray(got).reshape(want.shape)
            assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-8, key
    # planted mistakes: the diagonal eps in place of L = G eps G^-1 misses the reference by far
    f0 = out["fields"][0]
    diag = O.BEMRetOracle(f0["G1"], f0["G2"], f0["H1"], f0["H2"], s.nvec, f0["eps1"], f0["eps2"], f0["k"],
                          cross=False)
    s1 = diag.solve(*f0["exc"])[0]
    assert np.max(np.abs(s1 - fixture["sig1"][0])) / np.max(np.abs(fixture["sig1"][0])) > 0.5


def test_oracle_operator_is_the_chain_inverse_for_nested(fixture):
    """The source-side eps operator (afun) maps the chain's solution back onto its right-hand
    side for the nested particle; the operator with eps after the Green function does not."""
    cp, s = _coreshell(2.1)
    fp = cp.face_particle
    out = O.solve_spectrum(s, fp, INOUT, [lambda lam, e=e: e(lam) for e in cp.eps], [530.0],
                           np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), return_fields=True)
    f = out["fields"][0]
    s1, h1, s2, h2 = f["sol"]
    x = np.concatenate([s1, h1.T.ravel(), s2, h2.T.ravel()])
    phi, a, alpha, De = f["exc"]
    b = O.pack(phi, a, De, alpha)
    y = O.afun(f["G1"], f["G2"], f["H1"], f["H2"], s.nvec, f["eps1"], f["eps2"], f["k"], x)
    assert np.linalg.norm(y - b) / np.linalg.norm(b) < 1e-10
    # eps after the Green function (the uniform-eps shortcut) on the same x
    n = s.n
    X = x.reshape(8, n).T
    G1v, G2v, H1v, H2v = f["G1"] @ X[:, :4], f["G2"] @ X[:, 4:], f["H1"] @ X[:, :4], f["H2"] @ X[:, 4:]
    e1, e2 = f["eps1"], f["eps2"]
    De_wrong = e1 * H1v[:, 0] - e2 * H2v[:, 0] - 1j * f["k"] * np.sum(
        s.nvec * (e1[:, None] * G1v[:, 1:4] - e2[:, None] * G2v[:, 1:4]), axis=1)
    assert np.linalg.norm(De_wrong - De) / np.linalg.norm(De) > 1e-3


def test_oracle_coreshell_against_coated_mie():
    drude = M.EpsDrude(*DRUDE)
    lams = np.array([480.0, 530.0])
    for eps_shell, tol in ((2.1, 0.10), (1.0, 0.02)):
        cp, s = _coreshell(eps_shell)
        ext = O.solve_spectrum(s, cp.face_particle, INOUT, [lambda lam, e=e: e(lam) for e in cp.eps], lams[:1],
                               np.array([1.0, 0, 0]), np.array([0, 0, 1.0]))["ext"]
        mie = O.mie_coated(lams[0], 12.0, 20.0, drude(lams[0])[0], eps_shell, 1.0)[0]
        assert abs(ext[0] - mie) / mie < tol, (eps_shell, ext[0], mie)
