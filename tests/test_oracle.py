"""Pins of the CPU oracle (oracle/bem_oracle.py) against the reference's own validation
tables and against mathematics. CPU only."""
import os
from math import factorial

import numpy as np
import pytest

from oracle import bem_oracle as O
from mnpbem_amd.mesh import trisphere, trirod
from mnpbem_amd.comparticle import ComParticle
from mnpbem_amd.materials import EpsConst
from conftest import GOLDEN, ROOT

GOLD = np.loadtxt(os.path.join(ROOT, "mnpbem_amd", "data", "gold.dat"), comments="%")
PINF = np.load(os.path.join(ROOT, "mnpbem_amd", "data", "pinfty256.npz"))


def _surface(part):
    cp = ComParticle([EpsConst(1.0), EpsConst(2.0)], [part], [[2, 1]], 1, interp="curv")
    p = cp.pc
    return O.Surface(p.verts, p.faces, p.verts2, p.faces2)


def test_lgl_and_rule18_exactness():
    x, w = O.lgl(7)
    assert abs(w.sum() - 2) < 1e-14
    # LGL with n+1 points integrates degree 2n-1 exactly
    for d in range(14):
        assert abs(np.sum(w * x ** d) - (0 if d % 2 else 2.0 / (d + 1))) < 1e-13
    xs, ys, ws = O.rule18()
    for a in range(12):
        for b in range(12 - a):
            exact = 2.0 * factorial(a) * factorial(b) / factorial(a + b + 2)
            assert abs(np.sum(ws * xs ** a * ys ** b) - exact) < 1e-15


@pytest.mark.parametrize("rule,degree,npts", [(1, 1, 1), (2, 1, 3), (3, 2, 3), (4, 3, 4), (7, 5, 7)])
def test_triangle_rules_and_subdivision_exactness(rule, degree, npts):
    """Closed-form rules integrate x^a y^b exactly up to their degree, 2 a! b! / (a+b+2)! with
    sum w = 1, and not one degree higher; subdivision keeps the degree and brings a smooth
    non-polynomial integrand closer with every level. The product's rules (mnpbem_amd/quadrature.py)
    are the same point sets."""
    from mnpbem_amd.quadrature import QuadFace
    x, y, w = O.triangle_rule(rule)
    assert len(w) == npts and abs(w.sum() - 1) < 1e-15

    def err(x, y, w, a, b):
        return abs(np.sum(w * x ** a * y ** b) - 2.0 * factorial(a) * factorial(b) / factorial(a + b + 2))
    assert max(err(x, y, w, a, d - a) for d in range(degree + 1) for a in range(d + 1)) < 1e-14
    assert max(err(x, y, w, a, degree + 1 - a) for a in range(degree + 2)) > 1e-6
    for refine in (2, 3):
        xs, ys, ws = O.subdivide(x, y, w, refine)
        assert len(ws) == npts * refine ** 2 and abs(ws.sum() - 1) < 1e-14
        assert np.all(xs >= -1e-15) and np.all(ys >= -1e-15) and np.all(xs + ys <= 1 + 1e-15)
        assert max(err(xs, ys, ws, a, d - a) for d in range(degree + 1) for a in range(d + 1)) < 1e-14
    f = lambda x, y: 1.0 / (0.2 + x + y) ** 2
    exact = 2.0 * (np.log(1.2 / 0.2) - 1.0 + 0.2 / 1.2)      # 2 * integral of f over the triangle
    e = [abs(np.sum(ws * f(xs, ys)) - exact) for xs, ys, ws in (O.subdivide(x, y, w, r) for r in (1, 2, 4))]
    assert e[2] < e[1] < e[0]
    for refine in (None, 2):
        q = QuadFace(rule=rule, refine=refine)
        o = O.subdivide(x, y, w, refine)
        a = np.lexsort((q.w, q.y, q.x))
        b = np.lexsort((o[2], o[1], o[0]))
        assert np.allclose(q.x[a], o[0][b], atol=1e-15) and np.allclose(q.y[a], o[1][b], atol=1e-15)
        assert np.allclose(q.w[a], o[2][b], atol=1e-16)


def test_polar_rules_normalised():
    (x3, y3, w3), (x4, y4, w4) = O.polar_rules((7, 5))
    assert len(w3) == 144 and len(w4) == 192
    assert abs(w3.sum() - 1) < 1e-15 and abs(w4.sum() - 4) < 1e-14
    # radial nodes reach rho = 1 + 5e-7 (the rule's offset that keeps rho > 0 at the centre)
    assert np.all(x3 >= -1e-6) and np.all(y3 >= -1e-6) and np.all(x3 + y3 <= 1 + 1e-6)
    assert np.all(np.abs(x4) <= 1 + 1e-6) and np.all(np.abs(y4) <= 1 + 1e-6)


def test_sphere_geometry_invariants():
    s = _surface(trisphere(144, 20))
    assert s.n == 284
    # curved elements reproduce the sphere area to quadrature/interpolation accuracy
    assert abs(s.area.sum() - 4 * np.pi * 100) / (4 * np.pi * 100) < 2e-4
    assert np.allclose(np.linalg.norm(s.nvec, axis=1), 1)
    assert np.all(np.sum(s.nvec * s.pos, axis=1) > 0)          # outward normals
    assert np.allclose(np.sum(s.nvec * s.t1, axis=1), 0, atol=1e-14)


def test_closed_surface_identity():
    """Fuchs-Liu: sum_i a_i F_ij / a_j = -2 pi on a closed surface once the self terms carry
    the correction (compgreen_ret.py:192-272)."""
    s = _surface(trisphere(60, 10))
    kind = O.refinement_kind(s)
    corr, F = O.closed_correction(s, kind, [(np.arange(s.n), np.arange(s.n), 1.0)])
    Fc = F + np.diag(corr)
    colsum = np.sum(s.area[:, None] * Fc / s.area[None, :], axis=0)
    assert np.allclose(colsum, -2 * np.pi, atol=1e-12)
    # and the correction is small relative to 2 pi (the integration is already accurate)
    assert np.max(np.abs(corr)) < 0.05 * 2 * np.pi


def _brute(s, i, j, sub=12):
    """g_0 and f_0 of pair (i, j) by a subdivided 28-point rule on the source element."""
    tri = s._tri_nodes(j)
    N, Nx, Ny = O._tri6(*O.rule18()[:2])
    # subdivide the reference triangle into sub^2 triangles and map the rule
    xs, ys, ws = O.rule18()
    P, W = [], []
    h = 1.0 / sub
    for a in range(sub):
        for b in range(sub - a):
            for flip in (0, 1):
                if flip and b == sub - a - 1:
                    continue
                if not flip:
                    u = (a + xs) * h
                    v = (b + ys) * h
                else:
                    u = (a + 1 - xs) * h
                    v = (b + 1 - ys) * h
                Nn, Nxx, Nyy = O._tri6(u, v)
                for nodes in tri:
                    p = Nn.T @ nodes
                    J = np.cross(Nxx.T @ nodes, Nyy.T @ nodes)
                    P.append(p)
                    W.append(0.5 * ws * h * h * np.linalg.norm(J, axis=1))
    P, W = np.vstack(P), np.concatenate(W)
    v = s.pos[i] - P
    r = np.linalg.norm(v, axis=1)
    return np.sum(W / r), -(W / r ** 3) @ v


def test_near_coefficients_match_fine_integration():
    """Near-pair n = 0 coefficients are the integrals of 1/r and -r/r^3 over the source
    element: compare with a 12x12-subdivided rule (converged to ~1e-7)."""
    s = _surface(trisphere(60, 10))
    kind = O.refinement_kind(s)
    coef = O.refined_coefficients(s, kind)
    pairs = [(i, j) for (i, j) in coef if kind[i, j] == 1][:12]
    for i, j in pairs:
        g, f = coef[(i, j)]
        gb, fb = _brute(s, i, j)
        assert abs(g[0] - gb) / abs(gb) < 2e-6
        assert np.linalg.norm(f[:, 0] - fb) / np.linalg.norm(fb) < 2e-5


def test_lu_chain_backward_error():
    rng = np.random.default_rng(0)
    n = 40
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    import scipy.linalg as sla
    lu = sla.lu_factor(A)
    x = sla.lu_solve(lu, np.eye(n))
    assert np.linalg.norm(A @ x - np.eye(n)) < 1e-12 * np.linalg.norm(A) * np.linalg.norm(x)


@pytest.fixture(scope="module")
def gold_sphere():
    part = trisphere(144, 20)
    s = _surface(part)
    ml = np.loadtxt(os.path.join(GOLDEN, "bemret_sphere_matlab.csv"), delimiter=",", skiprows=1)
    eps = [lambda e: O.eps_const(e, 1.0), lambda e: O.eps_table(e, GOLD)]
    out = O.solve_spectrum(s, np.zeros(s.n, int), [[2, 1]], eps, ml[:, 0], np.array([1.0, 0, 0]),
                           np.array([0, 0, 1.0]), PINF["nvec"], PINF["area"], return_fields=True)
    return ml, out


def test_oracle_matches_matlab_sphere(gold_sphere):
    """validation/03_bemret/sphere: trisphere(144, 20), gold, 41 wavelengths."""
    ml, out = gold_sphere
    assert np.max(np.abs(out["ext"] - ml[:, 1]) / ml[:, 1]) < 1e-9
    assert np.max(np.abs(out["sca"] - ml[:, 2]) / ml[:, 2]) < 1e-9
    assert np.max(np.abs(out["ext"] - out["sca"] - ml[:, 3]) / ml[:, 3]) < 1e-9


def test_mie_matches_matlab_table():
    """validation/03_bemret/sphere/data/mie_matlab.csv: 20 nm gold sphere in vacuum."""
    mie = np.loadtxt(os.path.join(GOLDEN, "mie_sphere_matlab.csv"), delimiter=",", skiprows=1)
    for lam, ext, sca, _ in mie:
        e, s_ = O.mie_sphere(lam, 20.0, O.eps_table(lam, GOLD)[0], 1.0)
        assert abs(e - ext) / ext < 1e-7
        assert abs(s_ - sca) / sca < 1e-7


def test_oracle_bem_close_to_mie(gold_sphere):
    """Independent physics check: the BEM cross sections of trisphere(144, 20) follow the Mie
    series within the discretisation error of a 284-face mesh: a few per cent of the peak
    (relative errors grow in the red tail, where the cross sections are small), same peak."""
    ml, out = gold_sphere
    mie = np.array([O.mie_sphere(lam, 20.0, O.eps_table(lam, GOLD)[0], 1.0) for lam in ml[:, 0]])
    assert np.max(np.abs(out["ext"] - mie[:, 0])) / mie[:, 0].max() < 0.03
    assert np.max(np.abs(out["sca"] - mie[:, 1])) / mie[:, 1].max() < 0.05
    assert np.argmax(out["ext"]) == np.argmax(mie[:, 0])
    assert np.all(np.abs(out["ext"] - mie[:, 0])[:15] / mie[:15, 0] < 0.03)


def test_energy_balance_and_operator(gold_sphere):
    ml, out = gold_sphere
    assert np.all(out["ext"] - out["sca"] > 0)     # absorption >= 0
    # the solution satisfies the exact 8N operator of bem_ret_iter (afun) to LU accuracy
    f = out["fields"][5]
    x = O.pack(*[np.asarray(v) for v in (f["sol"][0], f["sol"][1], f["sol"][2], f["sol"][3])])
    phi, a, alpha, De = f["exc"]
    b = O.pack(phi, a, De, alpha)
    y = O.afun(f["G1"], f["G2"], f["H1"], f["H2"], out["green"].s.nvec,
               f["eps1"], f["eps2"], f["k"], x)
    assert np.linalg.norm(y - b) / np.linalg.norm(b) < 1e-11


def test_oracle_matches_matlab_rod():
    """validation/03_bemret/rod: trirod(10, 40, [15, 15, 15]) (curved quads), gold, both
    polarisations, 3 of the 41 wavelengths."""
    s = _surface(trirod(10.0, 40.0, [15, 15, 15]))
    ml = np.loadtxt(os.path.join(GOLDEN, "bemret_rod_matlab.csv"), delimiter=",", skiprows=1)
    sel = [0, 20, 40]
    eps = [lambda e: O.eps_const(e, 1.0), lambda e: O.eps_table(e, GOLD)]
    for pol, dr, ce, cs in (([1.0, 0, 0], [0, 0, 1.0], 1, 2), ([0, 0, 1.0], [1.0, 0, 0], 3, 4)):
        out = O.solve_spectrum(s, np.zeros(s.n, int), [[2, 1]], eps, ml[sel, 0], np.array(pol), np.array(dr),
                               PINF["nvec"], PINF["area"])
        assert np.max(np.abs(out["ext"] - ml[sel, ce]) / ml[sel, ce]) < 1e-9
        assert np.max(np.abs(out["sca"] - ml[sel, cs]) / ml[sel, cs]) < 1e-9


def test_oracle_bemstat_matches_matlab():
    """validation/02_bemstat/sphere: quasistatic trisphere(144, 20) in gold (x-polarised)."""
    s = _surface(trisphere(144, 20))
    ml = np.loadtxt(os.path.join(GOLDEN, "bemstat_sphere_matlab.csv"), delimiter=",", skiprows=1)
    kind = O.refinement_kind(s)
    corr, F = O.closed_correction(s, kind, [(np.arange(s.n), np.arange(s.n), 1.0)])
    F = F + np.diag(corr)
    pol = np.array([1.0, 0, 0])
    for row in ml[::8]:
        lam = row[0]
        e_in = O.eps_table(lam, GOLD)[0]
        sig = O.bemstat_solve(F, np.full(s.n, e_in), np.ones(s.n), -s.nvec @ pol)
        ext, sca, ab = O.stat_cross_sections(s.pos, s.area, sig, pol, 2 * np.pi / lam)
        assert abs(ext - row[1]) / row[1] < 1e-9
        assert abs(sca - row[2]) / row[2] < 1e-9


def test_surface_fields_pinned_by_boundary_conditions(gold_sphere):
    """Pins of surface_potential_field (compgreen_ret.py:911-1034) on a BEM solution of
    trisphere(144, 20) in gold under x-polarised light along z:
      * n_i . H1p/H2p reproduces H1/H2 entry by entry (the normal part of the Cartesian
        derivative, refined and far pairs alike);
      * continuity of phi and A (the BEM equations, exact to LU accuracy), of normal D (exact),
        and of tangential E and tangential H across the surface with the incident wave added
        outside (to the discretisation error of a 284-face mesh, which falls with refinement; at
        520 nm 0.46/0.31/0.23/0.19 % for E_t and 8.3/6.5/5.7/5.3 % for H_t at 116/284/508/796 faces).
    A dropped refined term, the static -1/d in place of ik - 1/d, the sign of the 2 pi n
    self term, the orientation of the cross product or a missing i k0 A each break at least one
    of these by a factor of 3 or more."""
    ml, out = gold_sphere
    g, f = out["green"], out["fields"][5]
    lam = ml[5, 0]
    eps = [O.eps_const(lam, 1.0), O.eps_table(lam, GOLD)]
    ks = [e[1] for e in eps]
    n = g.s.nvec
    for key, hkey, r in (("H1p", "H1", 0), ("H2p", "H2", 1)):
        P, H = g.eval(r, r, key, ks), g.eval(r, r, hkey, ks)
        assert np.max(np.abs(np.einsum("ic,icj->ij", n, P) - H)) / np.max(np.abs(H)) < 1e-13
    s1, h1, s2, h2 = f["sol"]
    F1 = O.surface_potential_field(g, ks, lam, s1, h1, s2, h2, 1)
    F2 = O.surface_potential_field(g, ks, lam, s1, h1, s2, h2, 2)
    k0, km = 2 * np.pi / lam, ks[0]
    pol, dr = np.array([1.0, 0, 0]), np.array([0, 0, 1.0])
    ph = np.exp(1j * km * (g.s.pos @ dr))
    e_inc = ph[:, None] * pol[None, :]
    h_inc = (km / k0) * ph[:, None] * np.cross(dr, pol)[None, :]
    rel = lambda v: np.linalg.norm(v) / np.linalg.norm(e_inc)
    assert np.linalg.norm(F1["phi"] - F2["phi"]) < 1e-12 * np.linalg.norm(F1["phi"])
    assert np.linalg.norm(F1["a"] - F2["a"] - e_inc / (1j * k0)) < 1e-12 * np.linalg.norm(F1["a"])
    dn1 = f["eps1"] * np.sum(n * F1["e"], axis=1)
    dn2 = f["eps2"] * np.sum(n * (F2["e"] + e_inc), axis=1)
    assert np.linalg.norm(dn1 - dn2) < 1e-10 * np.linalg.norm(dn1)
    assert rel(np.cross(n, F1["e"] - F2["e"] - e_inc)) < 0.01
    assert rel(np.cross(n, F1["h"] - F2["h"] - h_inc)) < 0.1
    # the surface derivatives of the potentials are the normal parts of their gradients
    assert np.allclose(F1["phip"], np.einsum("ic,icj->ij", n, g.eval(0, 0, "H1p", ks)) @ s1, rtol=0, atol=1e-12 *
                       np.abs(F1["phip"]).max())


def test_point_fields_pinned(gold_sphere):
    """Pins of point_potential_field (fields at points, compgreen_ret.py:911-1034 for
    p1 = ComPoint): (1) far away the field tends to e^{ikR}/R times the far-field amplitude of
    spectrum_ret.py:231-392 (itself pinned by the MATLAB cross sections); (2) the near-pair
    coefficients of a point 0.7 element radii above a face are the integrals of 1/r and -r/r^3
    over that face (subdivided rule) to the accuracy of the 28-point rule; (3) media: points
    inside the sphere get medium 2, outside medium 1."""
    ml, out = gold_sphere
    g, f = out["green"], out["fields"][5]
    s = g.s
    lam = ml[5, 0]
    ks = [O.eps_const(lam, 1.0)[1], O.eps_table(lam, GOLD)[1]]
    s1, h1, s2, h2 = f["sol"]
    fp = np.zeros(s.n, int)
    dirs = np.array([[0, 0, 1.0], [1.0, 0, 0], [0.6, 0.0, 0.8], [0.3, -0.5, 0.812404]])
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    k = ks[0]
    ef, hfar = O.farfield(s.pos, s.area, dirs, lam, k, s1, h1, s2, h2, np.zeros(0, int), np.arange(s.n))
    errs = []
    for R in (1e5, 1e6):
        F = O.point_potential_field(s, fp, [[2, 1]], R * dirs, ks, lam, s1, h1, s2, h2)
        sc = R * np.exp(-1j * k * R)
        err_e = np.linalg.norm(F["e"] * sc - ef) / np.linalg.norm(ef)
        err_h = np.linalg.norm(F["h"] * sc - hfar) / np.linalg.norm(hfar)
        # the near-zone terms fall off as 1 / (k R) relative to the radiating ones
        assert err_e < 3.0 / (k.real * R) and err_h < 3.0 / (k.real * R), (R, err_e, err_h)
        errs.append(err_e)
    assert 0.05 < errs[1] / errs[0] < 0.2
    for j in (0, 17, 140):
        x = s.pos[j] + 0.7 * s.rad[j] * s.nvec[j]
        gq, fq = O.point_face_coefficients(s, x, j)
        sv = s.pos[j]
        s.pos[j] = x                                   # _brute integrates over face j seen from s.pos[i]
        try:
            gb, fb = _brute(s, j, j)
        finally:
            s.pos[j] = sv
        assert abs(gq[0] - gb) / abs(gb) < 5e-3
        assert np.linalg.norm(fq[:, 0] - fb) / np.linalg.norm(fb) < 5e-2
    pts = np.array([[0, 0, 0.0], [3.0, 2.0, -1.0], [0, 0, 15.0], [25.0, 0, 0]])
    assert list(O.point_media(s, fp, [[2, 1]], pts)) == [2, 2, 1, 1]
