"""CPU oracle for the retarded BEM solve path (test infrastructure only).

THIS MODULE IS THE CHECKER, NOT THE PRODUCT. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it. The product (``mnpbem_amd``) never imports it and fails loudly when
its HIP library is missing.

It restates, in plain FP64 numpy/scipy written face by face and pair by pair,
the reference algorithm of MNPBEM's retarded direct solver:

    element geometry      geometry/particle.py:551-613 (_norm_curv), :464-549 (_norm_flat),
                          :1520-1575 (_quad_curv), :1364-1429 (_quadpol_curv), :2848-2880 (bradius)
    quadrature tables     utils/quadrature.py:15-89 (lglnodes), :182-223 (rule 18),
                          utils/quadface.py:106-249 (polar rules)
    refinement pairs      greenfun/refine_utils.py:12-158
    refined coefficients  greenfun/greenret_refined.py:122-285 (self), :287-516 (near)
    closed surfaces       greenfun/compgreen_ret.py:192-400, greenfun/compgreen_stat.py:219-464
    Green / H matrices    greenfun/greenret_refined.py:654-810, compgreen_ret.py:435-505, :739-858
    BEMRet init/solve     bem/bem_ret.py:407-525, :1360-1429, :1708-1752
    BEMStat               bem/bem_stat.py:241-404, :752-807; simulation/planewave_stat.py:152-355
    plane wave            simulation/planewave_ret.py:257-361, :391-504
    far field             spectrum/spectrum_ret.py:231-459
    8N operator           bem/bem_ret_iter.py:1196-1232, :1362-1501
    surface fields        greenfun/compgreen_ret.py:911-1040 (Gp, H1p, H2p: greenret_refined.py:812-866)
    fields at points      geometry/compoint.py:262-297 (media), greenfun/compgreen_ret.py:911-1034,
                          refine_utils.py:89-159 and greenret_refined.py:287-420 (point-face pairs)
    nested particles      region pairs of compgreen_ret.py:435-505 in bem_ret.py:407-433, the
                          general L1 = G1 eps1 G1^-1 of bem_ret.py:450-465 and the source-side eps
                          of the operator (bem_ret_iter.py:1362-1501)

Parity pins (see tests/test_oracle.py): the reference's MATLAB validation
tables for trisphere(144, 20) in gold (ext/sca/abs over 41 wavelengths), the
Mie series for a sphere (pinned itself by the reference's Mie table), quadrature exactness, the Fuchs-Liu
closed-surface identity, brute-force integration of the refined
coefficients on fine rules, LU backward error, and ext = sca + abs >= 0; surface
fields by n.Hp = H and the boundary conditions of a BEM solution; fields at points
by the far-field limit of the pinned far-field amplitude and by continuity with the
surface fields; nested particles by the reference's own solution of a core-shell sphere
(tests/golden/reference_coreshell.npz) and by the coated-sphere Mie series.

Inputs are plain arrays: corner ``verts``/``faces`` (NaN in column 3 for
triangles) and curved ``verts2``/``faces2`` (9 columns), e.g. from
``mnpbem_amd.mesh`` generators (mesh topology only).
"""
from math import factorial

import numpy as np
import scipy.linalg as sla

TWO_PI = 2.0 * np.pi
EPSF = np.finfo(float).eps
EV2NM = 1.0 / 8.0655477e-4          # misc/units.py:8


# ------------------------------------------------------------------ materials
def eps_drude(enei, eps0, wp, gam):
    """materials/eps_drude.py:73-107"""
    w = EV2NM / enei
    e = eps0 - wp ** 2 / (w * (w + 1j * gam))
    return e, TWO_PI / enei * np.sqrt(e + 0j)


def eps_const(enei, e):
    """materials/eps_const.py:48-72"""
    e = complex(e)
    return e, TWO_PI / enei * np.sqrt(e)


def eps_table(enei, table_ev_n_k):
    """materials/eps_table.py:51-152 (not-a-knot cubic splines of n and k in wavelength)."""
    from scipy.interpolate import CubicSpline
    tab = np.asarray(table_ev_n_k, float)
    lam = EV2NM / tab[:, 0]
    o = np.argsort(lam)
    n = CubicSpline(lam[o], tab[o, 1])(enei)
    k = CubicSpline(lam[o], tab[o, 2])(enei)
    e = (n + 1j * k) ** 2
    return e, TWO_PI / enei * np.sqrt(e)


# ------------------------------------------------------------------ quadrature
def lgl(n):
    """Legendre-Gauss-Lobatto: n+1 nodes (utils/quadrature.py:15-89)."""
    x = np.cos(np.pi * np.arange(n + 1) / n)
    for _ in range(100):
        P = [np.ones_like(x), x.copy()]
        for j in range(2, n + 1):
            P.append(((2 * j - 1) * x * P[-1] - (j - 1) * P[-2]) / j)
        dx = (x * P[n] - P[n - 1]) / ((n + 1) * P[n])
        x = x - dx
        if np.max(np.abs(dx)) <= EPSF:
            break
    P = [np.ones_like(x), x.copy()]
    for j in range(2, n + 1):
        P.append(((2 * j - 1) * x * P[-1] - (j - 1) * P[-2]) / j)
    return x, 2.0 / (n * (n + 1) * P[n] ** 2)


def rule18():
    """28-point degree-11 triangle rule (utils/quadrature.py:182-223), sum w = 1."""
    o = dict(a=1 / 3, b=0.9480217181434233, c=0.02598914092828833, d=0.8114249947041546,
             e=0.09428750264792270, f=0.01072644996557060, g=0.4946367750172147, p=0.5853132347709715,
             q=0.2073433826145142, r=0.1221843885990187, s=0.4389078057004907, t=0.6779376548825902,
             u=0.04484167758913055, v=0.27722066752827925, ww=0.8588702812826364, xx=0.0,
             yy=0.1411297187173636)
    xs = "a b c c d e e f g g p q q r s s t t u u v v ww ww xx xx yy yy".split()
    ys = "a c b c e d e g f g q p q s r s u v t v t u xx yy ww yy ww xx".split()
    wv = [0.08797730116222190] + [0.008744311553736190] * 3 + [0.03808157199393533] * 3 + \
         [0.01885544805613125] * 3 + [0.07215969754474100] * 3 + [0.06932913870553720] * 3 + \
         [0.04105631542928860] * 6 + [0.007362383783300573] * 6
    w = np.array(wv)
    return np.array([o[k] for k in xs]), np.array([o[k] for k in ys]), w / w.sum()


def polar_rules(npol=(7, 5)):
    """Polar rules: triangle (x3, y3, w3; sum 1) and square (x4, y4, w4; sum 4)
    (utils/quadface.py:106-249)."""
    xr, wr = lgl(npol[0])
    xa, wa = lgl(npol[1])
    rho = 0.5 * (xr + 1 + 1e-6)
    out = []
    for nsec, half, centre, radf, norm in ((3, 60, 270, lambda p: 1 / np.abs(2 * np.sin(p)), 1.0),
                                           (4, 45, 90, lambda p: 1 / np.abs(np.sin(p)), 4.0)):
        R, Ph, W = [], [], []
        for i in range(len(rho)):
            for j in range(len(xa)):
                R.append(rho[i])
                Ph.append(np.deg2rad(centre + half * xa[j]))
                W.append(wr[i] * wa[j])
        R, Ph, W = np.array(R), np.array(Ph), np.array(W)
        rad = radf(Ph)
        step = 2 * np.pi / nsec
        xs, ys, ws = [], [], []
        for s in range(nsec):
            xs.append(np.cos(Ph + s * step) * R * rad)
            ys.append(np.sin(Ph + s * step) * R * rad)
            ws.append(W * R * rad ** 2)
        xp, yp, w = np.concatenate(xs), np.concatenate(ys), np.concatenate(ws)
        w = norm * w / w.sum()
        if nsec == 3:
            out.append(((1 - np.sqrt(3) * xp - yp) / 3, (1 + np.sqrt(3) * xp - yp) / 3, w))
        else:
            out.append((xp, yp, w))
    return out


# ------------------------------------------------------------------ elements
def _tri6(x, y):
    a, b, c = x, y, 1 - x - y
    N = np.array([a * (2 * a - 1), b * (2 * b - 1), c * (2 * c - 1), 4 * a * b, 4 * b * c, 4 * c * a])
    Nx = np.array([4 * a - 1, 0 * a, -4 * c + 1, 4 * b, -4 * b, 4 * (c - a)])
    Ny = np.array([0 * a, 4 * b - 1, -4 * c + 1, 4 * a, 4 * (c - b), -4 * a])
    return N, Nx, Ny


def _quad9(x, y):
    N = np.array([0.25 * x * (x - 1) * y * (y - 1), 0.25 * x * (x + 1) * y * (y - 1),
                  0.25 * x * (x + 1) * y * (y + 1), 0.25 * x * (x - 1) * y * (y + 1),
                  0.5 * (1 - x * x) * y * (y - 1), 0.5 * x * (x + 1) * (1 - y * y),
                  0.5 * (1 - x * x) * y * (y + 1), 0.5 * x * (x - 1) * (1 - y * y), (1 - x * x) * (1 - y * y)])
    Nx = np.array([0.25 * (2 * x - 1) * y * (y - 1), 0.25 * (2 * x + 1) * y * (y - 1),
                   0.25 * (2 * x + 1) * y * (y + 1), 0.25 * (2 * x - 1) * y * (y + 1),
                   -x * y * (y - 1), 0.5 * (2 * x + 1) * (1 - y * y), -x * y * (y + 1),
                   0.5 * (2 * x - 1) * (1 - y * y), -2 * x * (1 - y * y)])
    Ny = np.array([0.25 * x * (x - 1) * (2 * y - 1), 0.25 * x * (x + 1) * (2 * y - 1),
                   0.25 * x * (x + 1) * (2 * y + 1), 0.25 * x * (x - 1) * (2 * y + 1),
                   0.5 * (1 - x * x) * (2 * y - 1), -x * (x + 1) * y, 0.5 * (1 - x * x) * (2 * y + 1),
                   -x * (x - 1) * y, -2 * (1 - x * x) * y])
    return N, Nx, Ny


def triangle_rule(rule):
    """Triangle rules with closed-form points (utils/quadrature.py:141-181), sum w = 1, as
    barycentric orbits: rule 1 centroid; 2 vertices; 3 edge midpoints; 4 centroid (-27/48) and
    the orbit of (3/5, 1/5, 1/5) (25/48); 7 centroid (9/40) and the orbits of (1 - 2a, a, a),
    a = (6 -+ sqrt 15)/21, weights (155 -+ sqrt 15)/1200; 18 the tabulated rule."""
    if rule == 18:
        return rule18()

    def orbit(b):
        # the distinct permutations of barycentric (b0, b1, b2), as (x, y) = (b0, b1)
        out = []
        for perm in ((0, 1, 2), (1, 0, 2), (1, 2, 0), (0, 2, 1), (2, 0, 1), (2, 1, 0)):
            q = (b[perm[0]], b[perm[1]])
            if not any(abs(q[0] - o[0]) < 1e-15 and abs(q[1] - o[1]) < 1e-15 for o in out):
                out.append(q)
        return out
    pts = []
    if rule == 1:
        pts = [((1 / 3, 1 / 3), 1.0)]
    elif rule == 2:
        pts = [(q, 1 / 3) for q in orbit((1.0, 0.0, 0.0))]
    elif rule == 3:
        pts = [(q, 1 / 3) for q in orbit((0.5, 0.5, 0.0))]
    elif rule == 4:
        pts = [((1 / 3, 1 / 3), -27 / 48)] + [(q, 25 / 48) for q in orbit((0.6, 0.2, 0.2))]
    elif rule == 7:
        r = np.sqrt(15.0)
        pts = [((1 / 3, 1 / 3), 9 / 40)]
        for a, wt in (((6 - r) / 21, (155 - r) / 1200), ((6 + r) / 21, (155 + r) / 1200)):
            pts += [(q, wt) for q in orbit((1 - 2 * a, a, a))]
    else:
        raise NotImplementedError(rule)
    x = np.array([q[0] for q, _ in pts])
    y = np.array([q[1] for q, _ in pts])
    return x, y, np.array([wt for _, wt in pts])


def subdivide(x, y, w, refine):
    """A unit-triangle rule on its refine^2 congruent sub-triangles (utils/quadface.py:21-44):
    sub-triangle (i, j) upward with corners (i, j) + the unit triangle / refine, downward ones
    (i + 1, j + 1) - the unit triangle / refine; weights shared evenly."""
    if refine is None or refine <= 1:
        return x, y, w
    X, Y = [], []
    for i in range(refine):
        for j in range(refine - i):
            X.append((i + x) / refine)
            Y.append((j + y) / refine)
        for j in range(refine - i - 1):
            X.append((i + 1 - x) / refine)
            Y.append((j + 1 - y) / refine)
    return np.concatenate(X), np.concatenate(Y), np.tile(w, len(X)) / len(X)


class Surface:
    """Curved-element surface: per-face geometry and quadrature points. Element areas use the
    28-point rule; the integrals of the refined entries use ``rule`` subdivided ``refine`` times
    (the reference applies ComParticle's rule/refine options to the particles' integration
    points after their geometry is set up, geometry/comparticle.py:144-161)."""

    def __init__(self, verts, faces, verts2, faces2, npol=(7, 5), rule=18, refine=None):
        self.verts = np.asarray(verts, float)
        self.faces = np.asarray(faces, float)
        self.verts2 = np.asarray(verts2, float)
        self.faces2 = np.asarray(faces2, float)
        self.n = len(self.faces)
        self.istri = np.isnan(self.faces[:, 3])
        self.gx, self.gy, self.gw = rule18()
        self.qx, self.qy, self.qw = subdivide(*triangle_rule(rule), refine)
        (self.x3, self.y3, self.w3), (self.x4, self.y4, self.w4) = polar_rules(npol)
        self._geometry()

    def _tri_nodes(self, i):
        """6-node triangles a face integrates over (quads: two halves)."""
        f = self.faces2[i]
        if self.istri[i]:
            return [self.verts2[f[[0, 1, 2, 4, 5, 6]].astype(int)]]
        return [self.verts2[f[[0, 1, 2, 4, 5, 8]].astype(int)], self.verts2[f[[2, 3, 0, 6, 7, 8]].astype(int)]]

    def quad_points(self, i, geometry=False):
        """Integration points per 6-node triangle (particle.py:1520-1575): the element's rule, or
        the 28-point rule of its geometry (area) with ``geometry``."""
        P, W = [], []
        x, y, w = (self.gx, self.gy, self.gw) if geometry else (self.qx, self.qy, self.qw)
        N, Nx, Ny = _tri6(x, y)
        for nodes in self._tri_nodes(i):
            P.append(N.T @ nodes)
            J = np.cross(Nx.T @ nodes, Ny.T @ nodes)
            W.append(0.5 * w * np.sqrt(np.sum(J * J, axis=1)))
        return np.vstack(P), np.concatenate(W)

    def polar_points(self, i):
        """Polar rule about the element centre (particle.py:1364-1429)."""
        f = self.faces2[i]
        if self.istri[i]:
            nodes = self.verts2[f[[0, 1, 2, 4, 5, 6]].astype(int)]
            N, Nx, Ny = _tri6(self.x3, self.y3)
            J = np.cross(Nx.T @ nodes, Ny.T @ nodes)
            return N.T @ nodes, self.w3 * 0.5 * np.sqrt(np.sum(J * J, axis=1))
        nodes = self.verts2[f[:9].astype(int)]
        N, Nx, Ny = _quad9(self.x4, self.y4)
        J = np.cross(Nx.T @ nodes, Ny.T @ nodes)
        return N.T @ nodes, self.w4 * np.sqrt(np.sum(J * J, axis=1))

    def _geometry(self):
        """particle.py:551-613 and :2848-2880"""
        n = self.n
        self.pos = np.zeros((n, 3))
        self.nvec = np.zeros((n, 3))
        self.t1 = np.zeros((n, 3))
        self.t2 = np.zeros((n, 3))
        self.area = np.zeros(n)
        self.rad = np.zeros(n)
        for i in range(n):
            self.area[i] = self.quad_points(i, geometry=True)[1].sum()
            f = self.faces2[i]
            if self.istri[i]:
                nodes = self.verts2[f[[0, 1, 2, 4, 5, 6]].astype(int)]
                N, Nx, Ny = _tri6(np.array([1 / 3]), np.array([1 / 3]))
                c, dx, dy = (N.T @ nodes)[0], (Nx.T @ nodes)[0], (Ny.T @ nodes)[0]
            else:
                nodes = self.verts2[f[[0, 1, 2, 4, 5, 8]].astype(int)]
                c = nodes[5]
                dx = nodes[0] - nodes[2]
                dy = -nodes[1] - nodes[2] + 2 * nodes[3] + 2 * nodes[4] - 2 * nodes[5]
            nn = np.cross(dy, dx)
            nn = nn / max(np.linalg.norm(nn), 1e-14)
            t1 = dy / max(np.linalg.norm(dy), 1e-14)
            self.pos[i], self.nvec[i], self.t1[i], self.t2[i] = c, nn, t1, np.cross(nn, t1)
            corners = self.faces[i][: 3 if self.istri[i] else 4].astype(int)
            self.rad[i] = max(np.linalg.norm(self.pos[i] - self.verts[v]) for v in corners)


# ------------------------------------------------------------------ refinement
def refinement_kind(s, rel=3.0, absc=0.0):
    """Dense (N, N) int matrix: 2 self, 1 near, 0 far (greenfun/refine_utils.py:29-158)."""
    p = s.pos
    sq = np.sum(p * p, axis=1)
    d2 = sq[:, None] + sq[None, :] - 2.0 * (p @ p.T)
    d2[d2 < 1e-10] = 0.0
    d = np.sqrt(d2)
    dr = d - s.rad[None, :]
    kind = np.zeros(d.shape, dtype=np.int8)
    kind[((dr < absc) | (dr / s.rad[:, None] < rel)) & (d != 0)] = 1
    kind[d == 0] = 2
    return kind


def refined_coefficients(s, kind, order=5):
    """{(i, j): (g[order+1], f[3, order+1])} (greenret_refined.py:122-516)."""
    out = {}
    for j in range(s.n):
        obs = np.nonzero(kind[:, j])[0]
        if len(obs) == 0:
            continue
        qp, qw = s.quad_points(j)
        pp, pw = s.polar_points(j)
        for i in obs:
            g = np.zeros(order + 1)
            f = np.zeros((3, order + 1))
            if kind[i, j] == 2:
                v = s.pos[i] - pp
                r = np.maximum(np.sqrt(np.sum(v * v, axis=1)), EPSF)
                nr, a1, a2 = v @ s.nvec[i], v @ s.t1[i], v @ s.t2[i]
                rr = np.maximum(r, 1e-4 * r.max())
                for n in range(order + 1):
                    g[n] = np.sum(pw * r ** (n - 1)) / factorial(n)
                    c1 = np.sum(pw * (n - 1) * nr * r ** (n - 3)) / factorial(n)
                    c2 = np.sum(pw * (n - 1) * a1 * r ** n / rr ** 3) / factorial(n)
                    c3 = np.sum(pw * (n - 1) * a2 * r ** n / rr ** 3) / factorial(n)
                    f[:, n] = s.nvec[i] * c1 + s.t1[i] * c2 + s.t2[i] * c3
            else:
                v = s.pos[i] - qp
                r = np.maximum(np.sqrt(np.sum(v * v, axis=1)), EPSF)
                r0 = np.linalg.norm(s.pos[j] - s.pos[i])
                dl = r - r0
                for n in range(order + 1):
                    g[n] = np.sum(qw * dl ** n / r) / factorial(n)
                    if n == 0:
                        sc = -1.0 / r ** 3
                    else:
                        sc = -dl ** n / (r ** 3 * factorial(n)) + dl ** (n - 1) / (r ** 2 * factorial(n - 1))
                    f[:, n] = (qw * sc) @ v
            out[(i, j)] = (g, f)
    return out


def stat_F(s, kind):
    """Refined quasistatic surface derivative F (compgreen_stat.py:219-464)."""
    r = s.pos[:, None, :] - s.pos[None, :, :]
    d = np.maximum(np.sqrt(np.sum(r * r, axis=2)), EPSF)
    F = -np.einsum("id,ijd->ij", s.nvec, r) / d ** 3 * s.area[None, :]
    np.fill_diagonal(F, 0.0)
    for j in range(s.n):
        obs = np.nonzero(kind[:, j])[0]
        if len(obs) == 0:
            continue
        qp, qw = s.quad_points(j)
        pp, pw = s.polar_points(j)
        for i in obs:
            pts, w = (pp, pw) if kind[i, j] == 2 else (qp, qw)
            v = s.pos[i] - pts
            rr = np.maximum(np.sqrt(np.sum(v * v, axis=1)), EPSF)
            F[i, j] = -np.sum(w * (v @ s.nvec[i]) / rr ** 3)
    return F


def closed_correction(s, kind, closed_sets):
    """-2 pi dir - sum_i a_i F_ij / a_j per face of each closed particle
    (compgreen_ret.py:192-272, :299-356)."""
    F = stat_F(s, kind)
    corr = np.zeros(s.n)
    for ind, loc, direction in closed_sets:
        fj = np.sum(s.area[loc][:, None] * F[np.ix_(loc, ind)] / s.area[ind][None, :], axis=0)
        corr[ind] = -TWO_PI * direction - fj
    return corr, F


class Green:
    """Refined retarded Green function of one surface, split into particles."""

    def __init__(self, s, face_particle, inout, closed_sets, order=5):
        self.s = s
        self.order = order
        self.fp = np.asarray(face_particle)
        self.inout = np.asarray(inout)
        self.kind = refinement_kind(s)
        self.coef = refined_coefficients(s, self.kind, order)
        self.corr, self.Fstat = closed_correction(s, self.kind, closed_sets)
        for i in range(s.n):
            if (i, i) in self.coef and self.corr[i] != 0.0:
                g, f = self.coef[(i, i)]
                f = f.copy()
                f[:, 0] += self.corr[i] * s.nvec[i]
                self.coef[(i, i)] = (g, f)

    def con(self, r1, r2):
        """Material connecting region r1 of the observer particle with region r2
        of the source particle, 0 = none (compgreen_ret.py:435-505)."""
        a = self.inout[:, r1][:, None]
        b = self.inout[:, r2][None, :]
        return np.where(a == b, a, 0)

    def block(self, k, key):
        """G, F or the Cartesian derivative Gp (n, 3, n) for all faces at wavenumber k
        (greenret_refined.py:654-866)."""
        s = self.s
        x = s.pos[:, None, :] - s.pos[None, :, :]
        d = np.maximum(np.sqrt(np.sum(x * x, axis=2)), EPSF)
        if key == "Gp":
            ikn = (1j * k) ** np.arange(self.order + 1)
            M = x * ((1j * k - 1 / d) / d ** 2 * s.area[None, :])[:, :, None]
            for (i, j), (g, f) in self.coef.items():
                M[i, j] = f @ ikn
            return np.transpose(M * np.exp(1j * k * d)[:, :, None], (0, 2, 1))
        if key == "G":
            M = s.area[None, :] / d + 0j
        else:
            M = np.einsum("id,ijd->ij", s.nvec, x) * (1j * k - 1 / d) / d ** 2 * s.area[None, :]
        ikn = (1j * k) ** np.arange(self.order + 1)
        for (i, j), (g, f) in self.coef.items():
            M[i, j] = g @ ikn if key == "G" else s.nvec[i] @ (f @ ikn)
        return M * np.exp(1j * k * d)

    def eval(self, r1, r2, key, ks):
        """Region pair (r1, r2) matrix; ks = wavenumber per material (1-based index - 1).
        Keys G, F, H1, H2 -> (n, n); Gp, H1p, H2p -> (n, 3, n) (H1p/H2p = Gp +- 2 pi n_i on
        the self terms, greenret_refined.py:846-863)."""
        con = self.con(r1, r2)
        vec = key in ("Gp", "H1p", "H2p")
        n = self.s.n
        out = np.zeros((n, 3, n) if vec else (n, n), complex)
        any_block = False
        for a in range(con.shape[0]):
            for b in range(con.shape[1]):
                if con[a, b] == 0:
                    continue
                any_block = True
                base = "Gp" if vec else ("G" if key == "G" else "F")
                M = self.block(ks[con[a, b] - 1], base)
                ia, ib = np.nonzero(self.fp == a)[0], np.nonzero(self.fp == b)[0]
                if vec:
                    blk = M[np.ix_(ia, np.arange(3), ib)]
                    if key in ("H1p", "H2p") and a == b:
                        sg = TWO_PI if key == "H1p" else -TWO_PI
                        for q, i in enumerate(ia):
                            blk[q, :, q] += sg * self.s.nvec[i]
                    out[np.ix_(ia, np.arange(3), ib)] = blk
                    continue
                blk = M[np.ix_(ia, ib)]
                if key in ("H1", "H2") and a == b:
                    blk = blk + (TWO_PI if key == "H1" else -TWO_PI) * np.eye(len(ia))
                out[np.ix_(ia, ib)] = blk
        return out if any_block else 0


# ------------------------------------------------------------------ BEMRet
def matrices(green, ks):
    """G1, G2, H1, H2 (bem_ret.py:407-433)."""
    def sub(a, b):
        return a if (np.isscalar(b) and b == 0) else a - b
    G1 = sub(green.eval(0, 0, "G", ks), green.eval(1, 0, "G", ks))
    G2 = sub(green.eval(1, 1, "G", ks), green.eval(0, 1, "G", ks))
    H1 = sub(green.eval(0, 0, "H1", ks), green.eval(1, 0, "H1", ks))
    H2 = sub(green.eval(1, 1, "H2", ks), green.eval(0, 1, "H2", ks))
    return G1, G2, H1, H2


class BEMRetOracle:
    """Dense LU chain of bem_ret.py:438-523 (CPU path: explicit inverses). ``cross``: the
    inside of one particle and the outside of another share a medium (nested particles,
    con{1,2} != 0); with non-uniform eps the chain then uses L1 = G1 eps1 G1^-1 and
    L2 = G2 eps2 G2^-1 (Eq. 22 of Garcia de Abajo and Howie, bem_ret.py:450-465) instead of
    the diagonal eps."""

    def __init__(self, G1, G2, H1, H2, nvec, eps1, eps2, k, cross=False):
        self.k, self.nvec = k, nvec
        n = len(nvec)
        self.G1_lu = sla.lu_factor(G1, check_finite=False)
        self.G2_lu = sla.lu_factor(G2, check_finite=False)
        G1i = sla.lu_solve(self.G1_lu, np.eye(n), check_finite=False)
        G2i = sla.lu_solve(self.G2_lu, np.eye(n), check_finite=False)
        e1, e2 = np.asarray(eps1), np.asarray(eps2)
        uniform = np.allclose(e1, e1[0]) and np.allclose(e2, e2[0])
        if uniform:
            self.L1, self.L2 = e1[0], e2[0]
        elif cross:
            self.L1 = G1 @ (e1[:, None] * G1i)
            self.L2 = G2 @ (e2[:, None] * G2i)
        else:
            self.L1, self.L2 = np.diag(e1), np.diag(e2)
        self.Sigma1 = H1 @ G1i
        Sigma2 = H2 @ G2i
        self.Delta_lu = sla.lu_factor(self.Sigma1 - Sigma2, check_finite=False)
        Deltai = sla.lu_solve(self.Delta_lu, np.eye(n), check_finite=False)
        L = self.L1 - self.L2
        nn = nvec @ nvec.T
        if np.isscalar(L):
            Sigma = self.Sigma1 * self.L1 - Sigma2 * self.L2 + k ** 2 * L * (Deltai * nn) * L
        else:
            Sigma = self.Sigma1 @ self.L1 - Sigma2 @ self.L2 + k ** 2 * ((L @ Deltai) * nn) @ L
        self.Sigma_lu = sla.lu_factor(Sigma, check_finite=False)

    def _mul(self, L, x):
        return L * x if np.isscalar(L) else L @ x

    def solve(self, phi, a, alpha, De):
        """Eqs. (19)-(20) of Garcia de Abajo & Howie (bem_ret.py:1708-1752)."""
        k, nv = self.k, self.nvec
        ls = lambda lu, b: sla.lu_solve(lu, b, check_finite=False)
        L1phi, L1a = self._mul(self.L1, phi), self._mul(self.L1, a)
        alpha = alpha - self.Sigma1 @ a + 1j * k * nv * L1phi[:, None]
        De = De - self.Sigma1 @ L1phi + 1j * k * np.sum(nv * L1a, axis=1)
        Ld = self.L1 - self.L2
        inner = np.sum(nv * self._mul(Ld, ls(self.Delta_lu, alpha)), axis=1)
        sig2 = ls(self.Sigma_lu, De + 1j * k * inner)
        h2 = ls(self.Delta_lu, 1j * k * nv * self._mul(Ld, sig2)[:, None] + alpha)
        return (ls(self.G1_lu, sig2 + phi), ls(self.G1_lu, h2 + a),
                ls(self.G2_lu, sig2), ls(self.G2_lu, h2))


def excitation(nvec, k0, eps1, eps2, a1, a1p, a2, a2p):
    """phi, a, alpha, De for phi1 = phi2 = 0 (bem_ret.py:1360-1429, Eqs. 10-11, 15, 18)."""
    phi = np.zeros(len(nvec), complex)
    a = a2 - a1
    alpha = a2p - a1p
    De = -1j * k0 * (eps2 * np.sum(nvec * a2, axis=1) - eps1 * np.sum(nvec * a1, axis=1))
    return phi, a, alpha, De


def planewave(pos, nvec, pol, direc, enei, eps_medium, excited):
    """a, a' = e^{ik r.dir}/(ik0) pol, ik (n.dir) a on ``excited`` faces
    (planewave_ret.py:257-361)."""
    k0 = TWO_PI / enei
    k = k0 * np.sqrt(complex(eps_medium))
    a = np.zeros((len(pos), 3), complex)
    ap = np.zeros((len(pos), 3), complex)
    ph = np.exp(1j * k * (pos[excited] @ direc)) / (1j * k0)
    a[excited] = ph[:, None] * pol
    ap[excited] = (1j * k * (nvec[excited] @ direc))[:, None] * ph[:, None] * pol
    return a, ap


def farfield(pos, area, dirs, enei, k, sig1, h1, sig2, h2, ind1, ind2):
    """e(dir), h(dir) (spectrum_ret.py:231-392)."""
    k0 = TWO_PI / enei
    ph = np.exp(-1j * k * (dirs @ pos.T)) * area[None, :]
    e = np.zeros((len(dirs), 3), complex)
    hf = np.zeros((len(dirs), 3), complex)
    for ind, s, h in ((ind1, sig1, h1), (ind2, sig2, h2)):
        if len(ind) == 0:
            continue
        P = ph[:, ind]
        Ph = P @ h[ind]
        e += 1j * k0 * Ph - 1j * k * dirs * (P @ s[ind])[:, None]
        hf += 1j * k * np.cross(dirs, Ph)
    return e, hf


def extinction(pol, direc, enei, k, eps_medium, pos, area, sol, ind1, ind2):
    """planewave_ret.py:437-504"""
    e, _ = farfield(pos, area, direc[None, :], enei, k, *sol, ind1, ind2)
    return float(np.real(4 * np.pi / k * np.imag(np.sum(np.conj(pol) * e[0]))))


def scattering(enei, k, eps_medium, pos, area, sol, ind1, ind2, pinf_nvec, pinf_area):
    """planewave_ret.py:391-435, spectrum_ret.py:395-459"""
    e, h = farfield(pos, area, pinf_nvec, enei, k, *sol, ind1, ind2)
    dsca = 0.5 * np.real(np.sum(pinf_nvec * np.cross(e, np.conj(h)), axis=1))
    return float(np.dot(pinf_area, dsca) / (0.5 * np.real(np.sqrt(complex(eps_medium)))))


# ------------------------------------------------------------------ BEMStat
def stat_G(s, kind):
    """Refined quasistatic G = a_j / d (compgreen_stat.py:219-464): self and near entries
    integrate 1/r over the source element."""
    r = s.pos[:, None, :] - s.pos[None, :, :]
    d = np.maximum(np.sqrt(np.sum(r * r, axis=2)), EPSF)
    G = s.area[None, :] / d
    np.fill_diagonal(G, 0.0)
    for j in range(s.n):
        obs = np.nonzero(kind[:, j])[0]
        if len(obs) == 0:
            continue
        qp, qw = s.quad_points(j)
        pp, pw = s.polar_points(j)
        for i in obs:
            pts, w = (pp, pw) if kind[i, j] == 2 else (qp, qw)
            G[i, j] = np.sum(w / np.maximum(np.linalg.norm(s.pos[i] - pts, axis=1), EPSF))
    return G


def bemstat_solve(F, eps1, eps2, phip):
    """sigma = -(diag Lambda + F)^{-1} phi' (bem_stat.py:241-404, :752-807)."""
    lam = 2 * np.pi * (eps1 + eps2) / (eps1 - eps2)
    M = -(np.diag(lam) + F)
    return sla.lu_solve(sla.lu_factor(M, check_finite=False), phip, check_finite=False)


def stat_cross_sections(pos, area, sig, pol, k):
    """Dipole absorption / scattering / extinction (planewave_stat.py:225-355)."""
    dip = (area[:, None] * pos).T @ sig
    ab = float(np.real(4 * np.pi * k * np.imag(np.dot(pol, dip))))
    sc = float(np.real(8 * np.pi / 3 * np.real(k) ** 4 * np.sum(np.abs(dip) ** 2)))
    return sc + ab, sc, ab


# ------------------------------------------------------------------ 8N operator
def afun(G1, G2, H1, H2, nvec, eps1, eps2, k, x):
    """Exact operator on x = [sig1, h1, sig2, h2] packed column-major
    (bem_ret_iter.py:1196-1232, :1362-1501): returns [phi, a, De, alpha]. The dielectric
    function multiplies the sources before the Green function and its derivative,
    alpha = H1 h1 - H2 h2 - i k n (G1 (eps1 sig1) - G2 (eps2 sig2)),
    De = H1 (eps1 sig1) - H2 (eps2 sig2) - i k n . (G1 (eps1 h1) - G2 (eps2 h2)), which is the
    chain's L = G eps G^-1 applied to G sig, and equals eps (G sig) when eps is uniform."""
    n = len(nvec)
    X = x.reshape(8, n).T
    v1, v2 = X[:, 0:4], X[:, 4:8]
    e1 = np.broadcast_to(np.asarray(eps1), (n,))[:, None]
    e2 = np.broadcast_to(np.asarray(eps2), (n,))[:, None]
    G1v, G2v, H1v, H2v = G1 @ v1, G2 @ v2, H1 @ v1, H2 @ v2
    G1e, G2e, H1e, H2e = G1 @ (e1 * v1), G2 @ (e2 * v2), H1 @ (e1 * v1), H2 @ (e2 * v2)
    phi = G1v[:, 0] - G2v[:, 0]
    a = G1v[:, 1:4] - G2v[:, 1:4]
    alpha = H1v[:, 1:4] - H2v[:, 1:4] - 1j * k * nvec * (G1e[:, 0] - G2e[:, 0])[:, None]
    De = H1e[:, 0] - H2e[:, 0] - 1j * k * np.sum(nvec * (G1e[:, 1:4] - G2e[:, 1:4]), axis=1)
    return np.concatenate([phi, a.T.ravel(), De, alpha.T.ravel()])


def pack(phi, a, De, alpha):
    return np.concatenate([phi, a.T.ravel(), De, alpha.T.ravel()])


# ------------------------------------------------------------------ driver
def solve_spectrum(surface, face_particle, inout, eps_funcs, enei_list, pol, direc,
                   pinf_nvec=None, pinf_area=None, medium=1, closed_sets=None, return_fields=False):
    """ext (and sca) per wavelength for a composite particle: the loop of
    validation/03_bemret/sphere/run_python.py:28-45."""
    fp = np.asarray(face_particle)
    inout = np.asarray(inout)
    if closed_sets is None:
        closed_sets = [(np.nonzero(fp == a)[0], np.nonzero(fp == a)[0], 1.0) for a in range(inout.shape[0])]
    green = Green(surface, fp, inout, closed_sets)
    io_faces = inout[fp]
    exc2 = np.nonzero(io_faces[:, 1] == medium)[0]
    exc1 = np.nonzero(io_faces[:, 0] == medium)[0]
    ext, sca, fields = [], [], []
    for enei in enei_list:
        eps_k = [f(enei) for f in eps_funcs]
        ks = [ek[1] for ek in eps_k]
        G1, G2, H1, H2 = matrices(green, ks)
        e1 = np.array([eps_k[m - 1][0] for m in inout[:, 0]])[fp]
        e2 = np.array([eps_k[m - 1][0] for m in inout[:, 1]])[fp]
        k0 = TWO_PI / enei
        cross = bool(np.any(inout[:, 0][:, None] == inout[:, 1][None, :]))
        bem = BEMRetOracle(G1, G2, H1, H2, surface.nvec, e1, e2, k0, cross=cross)
        epsm, km = eps_k[medium - 1]
        a2, a2p = planewave(surface.pos, surface.nvec, pol, direc, enei, epsm, exc2)
        a1, a1p = planewave(surface.pos, surface.nvec, pol, direc, enei, epsm, exc1)
        phi, a, alpha, De = excitation(surface.nvec, k0, e1, e2, a1, a1p, a2, a2p)
        sol = bem.solve(phi, a, alpha, De)
        ext.append(extinction(pol, direc, enei, km, epsm, surface.pos, surface.area, sol, exc1, exc2))
        if pinf_nvec is not None:
            sca.append(scattering(enei, km, epsm, surface.pos, surface.area, sol, exc1, exc2, pinf_nvec, pinf_area))
        if return_fields:
            fields.append(dict(G1=G1, G2=G2, H1=H1, H2=H2, sol=sol, exc=(phi, a, alpha, De),
                               eps1=e1, eps2=e2, k=k0))
    out = dict(ext=np.array(ext), sca=np.array(sca) if sca else None, green=green)
    if return_fields:
        out["fields"] = fields
    return out


def mie_sphere(enei, diameter, eps_in, eps_out, lmax=20):
    """Extinction and scattering cross sections (nm^2) of a sphere from the Mie series
    (Bohren & Huffman, Absorption and Scattering of Light by Small Particles, ch. 4):
    x = k a, m = n_in / n_out, a_n = (m psi_n(mx) psi_n'(x) - psi_n(x) psi_n'(mx)) /
    (m psi_n(mx) xi_n'(x) - xi_n(x) psi_n'(mx)), b_n likewise with m moved to the other terms,
    C_ext = 2 pi / k^2 sum (2n+1) Re(a_n + b_n), C_sca = 2 pi / k^2 sum (2n+1)(|a_n|^2 + |b_n|^2),
    with the Riccati-Bessel functions psi_n(z) = z j_n(z), xi_n(z) = z h_n^(1)(z). lmax = 20
    terms as the toolbox's miesolver default (mie/miesolver options)."""
    from scipy.special import spherical_jn, spherical_yn
    nb = np.sqrt(complex(eps_out))
    k = 2 * np.pi / enei * nb.real
    m = np.sqrt(complex(eps_in)) / nb
    x = k * 0.5 * diameter
    n = np.arange(1, lmax + 1)

    def psi(z):
        return z * spherical_jn(n, z), spherical_jn(n, z) + z * spherical_jn(n, z, derivative=True)

    def xi(z):
        h = spherical_jn(n, z) + 1j * spherical_yn(n, z)
        dh = spherical_jn(n, z, derivative=True) + 1j * spherical_yn(n, z, derivative=True)
        return z * h, h + z * dh

    px, dpx = psi(x)
    pmx, dpmx = psi(m * x)
    xx, dxx = xi(x)
    an = (m * pmx * dpx - px * dpmx) / (m * pmx * dxx - xx * dpmx)
    bn = (pmx * dpx - m * px * dpmx) / (pmx * dxx - m * xx * dpmx)
    ext = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * np.real(an + bn))
    sca = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * (np.abs(an) ** 2 + np.abs(bn) ** 2))
    return float(ext), float(sca)


def mie_coated(enei, d_core, d_shell, eps_core, eps_shell, eps_out, lmax=20):
    """Extinction and scattering cross sections (nm^2) of a coated sphere (core diameter
    d_core, shell diameter d_shell) from the Aden-Kerker series as written by Bohren and Huffman
    (ch. 8.1): x = k a, y = k b, m1, m2 = core and shell index over the medium's,
    A_n = (m2 psi_n(m2 x) psi_n'(m1 x) - m1 psi_n'(m2 x) psi_n(m1 x)) /
          (m2 chi_n(m2 x) psi_n'(m1 x) - m1 chi_n'(m2 x) psi_n(m1 x)),
    B_n = (m2 psi_n(m1 x) psi_n'(m2 x) - m1 psi_n(m2 x) psi_n'(m1 x)) /
          (m2 chi_n'(m2 x) psi_n(m1 x) - m1 psi_n'(m1 x) chi_n(m2 x)),
    a_n = (psi(y) [psi'(m2 y) - A chi'(m2 y)] - m2 psi'(y) [psi(m2 y) - A chi(m2 y)]) /
          (xi(y) [psi'(m2 y) - A chi'(m2 y)] - m2 xi'(y) [psi(m2 y) - A chi(m2 y)]),
    b_n likewise with B and m2 on the other terms; chi_n(z) = -z y_n(z). Reduces to mie_sphere
    when the core and shell media agree or the core vanishes. A check of the nested-particle
    path (no counterpart in the reference)."""
    from scipy.special import spherical_jn, spherical_yn
    nb = np.sqrt(complex(eps_out))
    k = 2 * np.pi / enei * nb.real
    m1, m2 = np.sqrt(complex(eps_core)) / nb, np.sqrt(complex(eps_shell)) / nb
    x, y = k * 0.5 * d_core, k * 0.5 * d_shell
    n = np.arange(1, lmax + 1)

    def psi(z):
        return z * spherical_jn(n, z), spherical_jn(n, z) + z * spherical_jn(n, z, derivative=True)

    def chi(z):
        return -z * spherical_yn(n, z), -(spherical_yn(n, z) + z * spherical_yn(n, z, derivative=True))

    def xi(z):
        h = spherical_jn(n, z) + 1j * spherical_yn(n, z)
        dh = spherical_jn(n, z, derivative=True) + 1j * spherical_yn(n, z, derivative=True)
        return z * h, h + z * dh

    p1x, dp1x = psi(m1 * x)
    p2x, dp2x = psi(m2 * x)
    c2x, dc2x = chi(m2 * x)
    A = (m2 * p2x * dp1x - m1 * dp2x * p1x) / (m2 * c2x * dp1x - m1 * dc2x * p1x)
    B = (m2 * p1x * dp2x - m1 * p2x * dp1x) / (m2 * dc2x * p1x - m1 * dp1x * c2x)
    py, dpy = psi(y)
    xy, dxy = xi(y)
    p2y, dp2y = psi(m2 * y)
    c2y, dc2y = chi(m2 * y)
    an = (py * (dp2y - A * dc2y) - m2 * dpy * (p2y - A * c2y)) / (xy * (dp2y - A * dc2y) - m2 * dxy * (p2y - A * c2y))
    bn = (m2 * py * (dp2y - B * dc2y) - dpy * (p2y - B * c2y)) / (m2 * xy * (dp2y - B * dc2y) - dxy * (p2y - B * c2y))
    ext = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * np.real(an + bn))
    sca = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * (np.abs(an) ** 2 + np.abs(bn) ** 2))
    return float(ext), float(sca)


def surface_potential_field(green, ks, enei, sig1, h1, sig2, h2, inout):
    """Potentials and fields on the faces from a solution (CompGreenRet.potential and .field,
    greenfun/compgreen_ret.py:912-1040): region inout (1 inside, 2 outside), sources of both
    regions. Returns phi, phip (n), a, ap, e, h (n, 3)."""
    r = inout - 1
    k0 = TWO_PI / enei
    hkey, pkey = ("H1", "H1p") if inout == 1 else ("H2", "H2p")

    def mm(M, x):
        if isinstance(M, int):
            return 0
        return np.tensordot(M, x, axes=([M.ndim - 1], [0]))

    G1, G2 = green.eval(r, 0, "G", ks), green.eval(r, 1, "G", ks)
    H1, H2 = green.eval(r, 0, hkey, ks), green.eval(r, 1, hkey, ks)
    P1, P2 = green.eval(r, 0, pkey, ks), green.eval(r, 1, pkey, ks)
    phi = mm(G1, sig1) + mm(G2, sig2)
    phip = mm(H1, sig1) + mm(H2, sig2)
    a = mm(G1, h1) + mm(G2, h2)
    ap = mm(H1, h1) + mm(H2, h2)
    e = 1j * k0 * a - mm(P1, sig1) - mm(P2, sig2)

    def cross(P, h):
        if isinstance(P, int):
            return 0
        return np.stack([mm(P[:, 1, :], h[:, 2]) - mm(P[:, 2, :], h[:, 1]),
                         mm(P[:, 2, :], h[:, 0]) - mm(P[:, 0, :], h[:, 2]),
                         mm(P[:, 0, :], h[:, 1]) - mm(P[:, 1, :], h[:, 0])], axis=1)

    hf = cross(P1, h1) + cross(P2, h2)
    return dict(phi=phi, phip=phip, a=a, ap=ap, e=e, h=hf)


def point_media(s, fp, inout, pts):
    """Medium of each point: the side of the nearest face (outside when (x - x_face) . n >= 0),
    compoint.py:262-297 with the nearest face over all particles."""
    d = np.sqrt(np.sum((pts[:, None, :] - s.pos[None, :, :]) ** 2, axis=2))
    j = np.argmin(d, axis=1)
    side = np.sum((pts - s.pos[j]) * s.nvec[j], axis=1) >= 0
    io = np.asarray(inout)
    return np.where(side, io[fp[j], 1], io[fp[j], 0])


def point_face_coefficients(s, x, j, order=5):
    """Taylor coefficients in ik of G and Gp between observer x and face j from the face's
    quadrature points around the centroid distance r0 (greenret_refined.py:287-420):
    g_n = sum_q w_q (r_q - r0)^n / (n! r_q), f_n the gradient of the same sum in x."""
    qp, qw = s.quad_points(j)
    v = x - qp
    r = np.maximum(np.sqrt(np.sum(v * v, axis=1)), EPSF)
    dl = r - np.linalg.norm(x - s.pos[j])
    g = np.zeros(order + 1)
    f = np.zeros((3, order + 1))
    for n in range(order + 1):
        g[n] = np.sum(qw * dl ** n / r) / factorial(n)
        sc = -1.0 / r ** 3 if n == 0 else -dl ** n / (r ** 3 * factorial(n)) + dl ** (n - 1) / (r ** 2 * factorial(n - 1))
        f[:, n] = (qw * sc) @ v
    return g, f


def point_potential_field(s, fp, inout, pts, ks, enei, sig1, h1, sig2, h2, order=5, rel=3.0, absc=0.0):
    """Potentials and fields at points from a solution: CompGreenRet(ComPoint, ComParticle)
    .potential/.field (compgreen_ret.py:911-1034). A point in medium m sees faces whose inside
    medium is m through sig1, h1 and faces whose outside medium is m through sig2, h2. Pairs
    with (d - rad_j) / rad_j < rel (refine_utils.py:89-159, radius of the source face) are
    Taylor-integrated over the source face's quadrature points (greenret_refined.py:287-420).
    Returns phi (np,), a, e, h (np, 3)."""
    med = point_media(s, fp, inout, pts)
    io = np.asarray(inout)
    npt = len(pts)
    k0 = TWO_PI / enei
    phi = np.zeros(npt, complex)
    a = np.zeros((npt, 3), complex)
    e = np.zeros((npt, 3), complex)
    hf = np.zeros((npt, 3), complex)
    ikn = lambda k: (1j * k) ** np.arange(order + 1)
    for i in range(npt):
        m = med[i]
        k = ks[m - 1]
        for j in range(s.n):
            if io[fp[j], 0] == m:
                sv, hv = sig1[j], h1[j]
            elif io[fp[j], 1] == m:
                sv, hv = sig2[j], h2[j]
            else:
                continue
            x = pts[i] - s.pos[j]
            d = np.linalg.norm(x)
            if d == 0:
                continue
            ph = np.exp(1j * k * d)
            dr = d - s.rad[j]
            if dr < absc or dr / s.rad[j] < rel:
                g, f = point_face_coefficients(s, pts[i], j, order)
                G = (g @ ikn(k)) * ph
                Gp = (f @ ikn(k)) * ph
            else:
                G = s.area[j] / d * ph
                Gp = x * ((1j * k - 1 / d) / d ** 2 * s.area[j]) * ph
            phi[i] += G * sv
            a[i] += G * hv
            e[i] -= Gp * sv
            hf[i] += np.cross(Gp, hv)
    e += 1j * k0 * a
    return dict(phi=phi, a=a, e=e, h=hf, medium=med)
