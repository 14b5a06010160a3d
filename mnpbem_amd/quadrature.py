"""Quadrature rules on the unit triangle and the unit square.

* ``lgl_nodes``: Legendre-Gauss-Lobatto nodes/weights by Newton iteration on
  x P_n(x) - P_{n-1}(x) from Chebyshev-Lobatto starting points (the method of
  the reference's utils/quadrature.py:15-89).
* ``triangle_rule18``: the published 28-point, degree-11 triangle rule
  (Burkardt's triangle rule table, rule 18; reference utils/quadrature.py:182-223),
  weights normalised to sum 1.
* ``triangle_rule(rule)``: the reference's other rule numbers whose points follow from closed
  forms (utils/quadrature.py:141-181): 1 (centroid, degree 1), 2 (vertices, degree 1),
  3 (edge midpoints, degree 2), 4 (centroid and three interior points, degree 3), 7 (Radon's
  seven-point rule, degree 5, with the sqrt(15) orbits); 18 as above.
* ``subdivide``: a rule repeated on the refine x refine congruent sub-triangles of the unit
  triangle (upward and downward ones), weights divided evenly (the reference's ``refine``
  option, utils/quadface.py:21-44).
* ``polar_triangle`` / ``polar_square``: polar rules around the element centre
  used for the singular self-term integrals (reference utils/quadface.py:106-249):
  LGL nodes in radius and angle over one 60 (45) degree sector, replicated
  over 3 (4) sectors; weights sum to 1 (triangle) or 4 (square [-1,1]^2).
"""
import numpy as np


def lgl_nodes(n):
    """n+1 Legendre-Gauss-Lobatto nodes on [-1, 1] and their weights."""
    if n < 1:
        raise ValueError("lgl_nodes needs n >= 1")
    x = np.cos(np.pi * np.arange(n + 1) / n)
    prev = np.full_like(x, 2.0)
    P = np.zeros((n + 1, n + 1))
    while np.max(np.abs(x - prev)) > np.finfo(float).eps:
        prev = x.copy()
        P[:, 0] = 1.0
        P[:, 1] = x
        for j in range(2, n + 1):
            P[:, j] = ((2 * j - 1) * x * P[:, j - 1] - (j - 1) * P[:, j - 2]) / j
        x = prev - (x * P[:, n] - P[:, n - 1]) / ((n + 1) * P[:, n])
    w = 2.0 / (n * (n + 1) * P[:, n] ** 2)
    return x, w


def triangle_rule18():
    """28-point symmetric rule on the triangle (0,0),(1,0),(0,1); sum(w) = 1."""
    # orbit generators (published table values)
    a = 1.0 / 3.0
    b, c = 0.9480217181434233, 0.02598914092828833
    d, e = 0.8114249947041546, 0.09428750264792270
    f, g = 0.01072644996557060, 0.4946367750172147
    p, q = 0.5853132347709715, 0.2073433826145142
    r, s = 0.1221843885990187, 0.4389078057004907
    t, u, v = 0.6779376548825902, 0.04484167758913055, 0.27722066752827925
    ww, xx, yy = 0.8588702812826364, 0.0, 0.1411297187173636
    w1, w2, w3, w4 = 0.08797730116222190, 0.008744311553736190, 0.03808157199393533, 0.01885544805613125
    w5, w6, w7, w8 = 0.07215969754474100, 0.06932913870553720, 0.04105631542928860, 0.007362383783300573
    pts = [
        (a, a, w1),
        (b, c, w2), (c, b, w2), (c, c, w2),
        (d, e, w3), (e, d, w3), (e, e, w3),
        (f, g, w4), (g, f, w4), (g, g, w4),
        (p, q, w5), (q, p, w5), (q, q, w5),
        (r, s, w6), (s, r, w6), (s, s, w6),
        (t, u, w7), (t, v, w7), (u, t, w7), (u, v, w7), (v, t, w7), (v, u, w7),
        (ww, xx, w8), (ww, yy, w8), (xx, ww, w8), (xx, yy, w8), (yy, ww, w8), (yy, xx, w8),
    ]
    arr = np.asarray(pts)
    w = arr[:, 2] / arr[:, 2].sum()
    return arr[:, 0].copy(), arr[:, 1].copy(), w


def triangle_rule(rule=18):
    """Points (x, y) and weights (sum 1) of triangle rule number ``rule`` on (0,0),(1,0),(0,1)."""
    if rule == 18:
        return triangle_rule18()
    if rule == 1:
        x, y, w = [1 / 3], [1 / 3], [1.0]
    elif rule == 2:
        x, y, w = [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1 / 3] * 3
    elif rule == 3:
        x, y, w = [0.5, 0.5, 0.0], [0.0, 0.5, 0.5], [1 / 3] * 3
    elif rule == 4:
        # centroid with weight -27/48 and the orbit of (3/5, 1/5, 1/5) with 25/48 each
        x, y, w = [1 / 3, 0.6, 0.2, 0.2], [1 / 3, 0.2, 0.6, 0.2], [-27 / 48] + [25 / 48] * 3
    elif rule == 7:
        # Radon: centroid (9/40) and the orbits of barycentric (b, a, a) with
        # a = (6 -+ sqrt 15) / 21, b = 1 - 2a, weights (155 -+ sqrt 15) / 1200
        r = np.sqrt(15.0)
        a1, a2 = (6 - r) / 21, (6 + r) / 21
        b1, b2 = 1 - 2 * a1, 1 - 2 * a2
        w1, w2 = (155 - r) / 1200, (155 + r) / 1200
        x = [1 / 3, b1, a1, a1, b2, a2, a2]
        y = [1 / 3, a1, b1, a1, a2, b2, a2]
        w = [9 / 40, w1, w1, w1, w2, w2, w2]
    else:
        raise NotImplementedError("triangle rule {} is not provided (rules 1, 2, 3, 4, 7, 18)".format(rule))
    return np.asarray(x, float), np.asarray(y, float), np.asarray(w, float)


def subdivide(x, y, w, refine):
    """The rule (x, y, w) on each of the refine^2 sub-triangles of the unit triangle: upward
    triangles (i, j) + (x, y) and downward ones (i + 1, j + 1) - (x, y), scaled by 1 / refine;
    weights w / refine^2."""
    refine = int(refine)
    if refine <= 1:
        return x, y, w
    xs, ys = [], []
    for i in range(refine):
        for j in range(refine - i):
            xs.append(i + x)
            ys.append(j + y)
            if j < refine - 1 - i:
                xs.append(i + 1 - x)
                ys.append(j + 1 - y)
    k = len(xs)
    return np.concatenate(xs) / refine, np.concatenate(ys) / refine, np.tile(w, k) / k


def _polar_grid(npol, half_angle_deg, centre_deg):
    n_rad, n_ang = npol
    xr, wr = lgl_nodes(n_rad)
    xa, wa = lgl_nodes(n_ang)
    rho = 0.5 * (xr + 1.0 + 1e-6)
    phi = np.deg2rad(centre_deg + half_angle_deg * xa)
    rho2, phi2 = np.meshgrid(rho, phi, indexing="ij")
    w2 = np.outer(wr, wa).ravel()
    return rho2.ravel(), phi2.ravel(), w2


def polar_triangle(npol=(7, 5)):
    """Polar rule on the unit triangle around its centroid; sum(w) = 1."""
    if np.isscalar(npol):
        npol = (int(npol), int(npol))
    rho, phi, w0 = _polar_grid(npol, 60.0, 270.0)
    rad = 1.0 / np.abs(2.0 * np.sin(phi))
    sector = np.deg2rad(120.0)
    phis = np.concatenate([phi, phi + sector, phi + 2 * sector])
    radial = np.tile(rho * rad, 3)
    xp = np.cos(phis) * radial
    yp = np.sin(phis) * radial
    x = (1.0 - np.sqrt(3.0) * xp - yp) / 3.0
    y = (1.0 + np.sqrt(3.0) * xp - yp) / 3.0
    w = np.tile(w0 * rho * rad ** 2, 3)
    return x, y, w / w.sum()


def polar_square(npol=(7, 5)):
    """Polar rule on [-1, 1]^2 around the origin; sum(w) = 4."""
    if np.isscalar(npol):
        npol = (int(npol), int(npol))
    rho, phi, w0 = _polar_grid(npol, 45.0, 90.0)
    rad = 1.0 / np.abs(np.sin(phi))
    phis = np.concatenate([phi + q * 0.5 * np.pi for q in range(4)])
    radial = np.tile(rho * rad, 4)
    w = np.tile(w0 * rho * rad ** 2, 4)
    return np.cos(phis) * radial, np.sin(phis) * radial, 4.0 * w / w.sum()


class QuadFace:
    """Bundle of the rules a boundary element uses (reference utils/quadface.py:48-104)."""

    def __init__(self, rule=18, npol=(7, 5), refine=None):
        if np.isscalar(npol):
            npol = (int(npol), int(npol))
        self.rule, self.refine = int(rule), refine
        self.npol = tuple(npol)
        self.x, self.y, self.w = triangle_rule(self.rule)
        if refine is not None and int(refine) > 1:
            self.x, self.y, self.w = subdivide(self.x, self.y, self.w, refine)
        self.x3, self.y3, self.w3 = polar_triangle(self.npol)
        self.x4, self.y4, self.w4 = polar_square(self.npol)
