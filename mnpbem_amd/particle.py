"""Discretised particle surfaces: flat or curved triangles and quadrilaterals.

Host-side setup code (wavelength independent). A face is a row of ``faces``
(corner indices, NaN in column 3 for a triangle). Curved elements carry
``verts2``/``faces2``: ``faces2`` rows are
``[v0, v1, v2, nan, m01, m12, m20, nan, nan]`` for 6-node triangles and
``[v0, v1, v2, v3, m01, m12, m23, m30, centre]`` for 9-node quadrilaterals
(reference geometry/particle.py:249-397 and :1745-1831).

Per-face geometry (centroid ``pos``, unit normal ``nvec``, tangents
``tvec1``/``tvec2``, ``area``) follows geometry/particle.py:464-613; the
boundary-element and polar quadrature points follow :1276-1575.
"""
import copy

import numpy as np

from .quadrature import QuadFace

_EPS_NORM = 1e-14


# --------------------------------------------------------------------------
# shape functions on the reference elements
# --------------------------------------------------------------------------
def tri6(x, y):
    """6-node triangle: value, d/dx, d/dy at (x, y); nodes (1,0),(0,1),(0,0), mid-edges."""
    x = np.atleast_1d(np.asarray(x, float))
    y = np.atleast_1d(np.asarray(y, float))
    l1, l2, l3 = x, y, 1.0 - x - y
    N = np.stack([l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), l3 * (2 * l3 - 1),
                  4 * l1 * l2, 4 * l2 * l3, 4 * l3 * l1], axis=1)
    z = np.zeros_like(x)
    dx = np.stack([4 * l1 - 1, z, 1 - 4 * l3, 4 * l2, -4 * l2, 4 * (l3 - l1)], axis=1)
    dy = np.stack([z, 4 * l2 - 1, 1 - 4 * l3, 4 * l1, 4 * (l3 - l2), -4 * l1], axis=1)
    return N, dx, dy


def quad9(x, y):
    """9-node quadrilateral on [-1,1]^2: corners (-1,-1),(1,-1),(1,1),(-1,1), mid-edges, centre."""
    x = np.atleast_1d(np.asarray(x, float))
    y = np.atleast_1d(np.asarray(y, float))
    ym, yp, y0 = 0.5 * y * (y - 1), 0.5 * y * (y + 1), 1 - y ** 2
    xm, xp, x0 = 0.5 * x * (x - 1), 0.5 * x * (x + 1), 1 - x ** 2
    dym, dyp, dy0 = y - 0.5, y + 0.5, -2 * y
    dxm, dxp, dx0 = x - 0.5, x + 0.5, -2 * x
    N = np.stack([xm * ym, xp * ym, xp * yp, xm * yp, x0 * ym, xp * y0, x0 * yp, xm * y0, x0 * y0], axis=1)
    dx = np.stack([dxm * ym, dxp * ym, dxp * yp, dxm * yp, dx0 * ym, dxp * y0, dx0 * yp, dxm * y0, dx0 * y0], axis=1)
    dy = np.stack([xm * dym, xp * dym, xp * dyp, xm * dyp, x0 * dym, xp * dy0, x0 * dyp, xm * dy0, x0 * dy0], axis=1)
    return N, dx, dy


def quad4(x, y):
    """Bilinear quadrilateral on [-1,1]^2."""
    x = np.atleast_1d(np.asarray(x, float))
    y = np.atleast_1d(np.asarray(y, float))
    N = 0.25 * np.stack([(1 - x) * (1 - y), (1 + x) * (1 - y), (1 + x) * (1 + y), (1 - x) * (1 + y)], axis=1)
    dx = 0.25 * np.stack([-(1 - y), 1 - y, 1 + y, -(1 + y)], axis=1)
    dy = 0.25 * np.stack([-(1 - x), -(1 + x), 1 + x, 1 - x], axis=1)
    return N, dx, dy


def _unit(v):
    n = np.sqrt(np.sum(v * v, axis=-1, keepdims=True))
    return v / np.maximum(n, _EPS_NORM)


def _map(shape, nodes):
    """Contract shape values (m, c) with element nodes (k, c, 3) -> (k, m, 3)."""
    return np.einsum("mc,kcd->kmd", shape, nodes)


# --------------------------------------------------------------------------
# Particle
# --------------------------------------------------------------------------
class Particle:
    """Faces and vertices of a discretised particle surface."""

    def __init__(self, verts=None, faces=None, interp="flat", norm=True, rule=18, npol=(7, 5), refine=None):
        self.interp = interp
        self.verts2 = None
        self.faces2 = None
        self.quad_rule = QuadFace(rule=rule, npol=npol, refine=refine)
        if verts is None or len(verts) == 0:
            self.verts = np.zeros((0, 3))
            self.faces = np.zeros((0, 4))
            self._set_empty()
            return
        verts = np.asarray(verts, dtype=float)
        if verts.ndim != 2 or verts.shape[1] != 3:
            raise ValueError("Particle: 'verts' must have shape (N, 3); got {}.".format(verts.shape))
        if faces is None:
            self.verts = verts
            self.faces = np.zeros((0, 4))
        else:
            faces = np.asarray(faces, dtype=float)
            if faces.ndim != 2:
                raise ValueError("Particle: 'faces' must be 2D; got {}D.".format(faces.ndim))
            if faces.shape[1] == 3:
                self.verts = verts
                self.faces = np.hstack([faces, np.full((len(faces), 1), np.nan)])
            elif faces.shape[1] == 4:
                self.verts = verts
                self.faces = faces
            else:
                # curved input: corner vertices are the subset referenced by columns 0..3
                self.verts2 = verts
                self.faces2 = faces
                corners = faces[:, :4]
                ok = ~np.isnan(corners)
                used, inv = np.unique(corners[ok].astype(np.int64), return_inverse=True)
                self.verts = verts[used]
                f = np.full(corners.shape, np.nan)
                f[ok] = inv
                self.faces = f
        if norm and len(self.faces) > 0:
            self._norm()
        else:
            self._set_empty()

    # ---------------- basic properties ----------------
    def _set_empty(self):
        self.pos = np.zeros((0, 3))
        self.vec = [np.zeros((0, 3))] * 3
        self.area = np.zeros(0)

    @property
    def nvec(self):
        return self.vec[2]

    @property
    def tvec1(self):
        return self.vec[0]

    @property
    def tvec2(self):
        return self.vec[1]

    @property
    def nfaces(self):
        return self.faces.shape[0]

    n = nfaces

    @property
    def nverts(self):
        return self.verts.shape[0]

    def index34(self, ind=None):
        f = self.faces if ind is None else self.faces[np.asarray(ind)]
        tri = np.isnan(f[:, 3])
        return np.nonzero(tri)[0], np.nonzero(~tri)[0]

    # ---------------- element geometry ----------------
    def _norm(self):
        if self.interp == "flat":
            self._norm_flat()
        else:
            self._norm_curv()

    def _norm_flat(self):
        """Centroids, areas and bases of flat elements (particle.py:464-549)."""
        nf = self.nfaces
        i3, i4 = self.index34()
        pos = np.zeros((nf, 3))
        f = self.faces
        if len(i3):
            c = f[i3, :3].astype(np.int64)
            pos[i3] = self.verts[c].sum(axis=1) / 3.0
        if len(i4):
            c = f[i4, :4].astype(np.int64)
            pos[i4] = self.verts[c].sum(axis=1) / 4.0
        # quads are split into (v0,v1,v2) and (v2,v3,v0); the larger half sets the basis
        tri = [f[:, :3]]
        if len(i4):
            tri.append(f[i4][:, [2, 3, 0]])
        tri = np.vstack(tri).astype(np.int64)
        a, b, c = self.verts[tri[:, 0]], self.verts[tri[:, 1]], self.verts[tri[:, 2]]
        e1, e2 = a - b, c - b
        nv = np.cross(e1, e2)
        ar = 0.5 * np.sqrt(np.sum(nv * nv, axis=1))
        t1, nn = _unit(e1), _unit(nv)
        t2 = np.cross(nn, t1)
        area = ar[:nf].copy()
        T1, T2, NN = t1[:nf].copy(), t2[:nf].copy(), nn[:nf].copy()
        if len(i4):
            second = nf + np.arange(len(i4))
            area[i4] += ar[second]
            bigger = ar[second] > ar[i4]
            sel, src = i4[bigger], second[bigger]
            T1[sel], T2[sel], NN[sel] = t1[src], t2[src], nn[src]
        self.pos, self.area, self.vec = pos, area, [T1, T2, NN]

    def _norm_curv(self):
        """Centroids and bases of curved elements (particle.py:551-613);
        areas are the sums of the boundary-element quadrature weights."""
        nf = self.nfaces
        _, w, iface = self.quad()
        self.area = np.bincount(iface, weights=w, minlength=nf)
        i3, i4 = self.index34()
        pos = np.zeros((nf, 3))
        d1 = np.zeros((nf, 3))
        d2 = np.zeros((nf, 3))
        f2 = self.faces2
        if len(i3):
            nodes = self.verts2[f2[i3][:, [0, 1, 2, 4, 5, 6]].astype(np.int64)]
            N, dx, dy = tri6(1.0 / 3.0, 1.0 / 3.0)
            pos[i3] = _map(N, nodes)[:, 0]
            d1[i3] = _map(dy, nodes)[:, 0]
            d2[i3] = _map(dx, nodes)[:, 0]
        if len(i4):
            # centre node; derivatives of the first split triangle (v0,v1,v2,m01,m12,centre)
            nodes = self.verts2[f2[i4][:, [0, 1, 2, 4, 5, 8]].astype(np.int64)]
            pos[i4] = nodes[:, 5]
            cx = np.array([1.0, 0, -1, 0, 0, 0])
            cy = np.array([0.0, -1, -1, 2, 2, -2])
            d1[i4] = np.einsum("c,kcd->kd", cy, nodes)
            d2[i4] = np.einsum("c,kcd->kd", cx, nodes)
        nn = np.cross(d1, d2)
        nn = nn / np.maximum(np.linalg.norm(nn, axis=1, keepdims=True), _EPS_NORM)
        t1 = d1 / np.maximum(np.linalg.norm(d1, axis=1, keepdims=True), _EPS_NORM)
        self.pos, self.vec = pos, [t1, np.cross(nn, t1), nn]

    # ---------------- quadrature ----------------
    def _split_triangles(self, ind):
        """Triangles a face list decomposes into, with their owner positions in ``ind``."""
        i3, i4 = self.index34(ind)
        ind = np.asarray(ind)
        if self.interp == "flat":
            first = self.faces[ind][:, :3]
            second = self.faces[ind[i4]][:, [2, 3, 0]]
        else:
            f2 = self.faces2[ind]
            first = np.zeros((len(ind), 6))
            first[i3] = f2[i3][:, [0, 1, 2, 4, 5, 6]]
            first[i4] = f2[i4][:, [0, 1, 2, 4, 5, 8]]
            second = self.faces2[ind[i4]][:, [2, 3, 0, 6, 7, 8]]
        tri = np.vstack([first, second]).astype(np.int64)
        owner = np.concatenate([np.arange(len(ind)), i4])
        return tri, owner

    def quad(self, ind=None):
        """Boundary-element quadrature (28 points per triangle; quads split in two).

        Returns ``(pos, w, iface)``: points (M, 3), weights (M,) with
        sum over a face = its area, and the owning position in ``ind`` per point
        (particle.py:1431-1575)."""
        ind = np.arange(self.nfaces) if ind is None else np.asarray(ind)
        q = self.quad_rule
        tri, owner = self._split_triangles(ind)
        m = len(q.w)
        if self.interp == "flat":
            N = np.stack([q.x, q.y, 1 - q.x - q.y], axis=1)
            nodes = self.verts[tri]
            a = nodes[:, 0] - nodes[:, 2]
            b = nodes[:, 1] - nodes[:, 2]
            jac = np.linalg.norm(np.cross(a, b), axis=1)[:, None] * np.ones((1, m))
        else:
            N, dx, dy = tri6(q.x, q.y)
            nodes = self.verts2[tri]
            jac = np.linalg.norm(np.cross(_map(dx, nodes), _map(dy, nodes)), axis=2)
        pts = _map(N, nodes).reshape(-1, 3)
        w = (0.5 * q.w[None, :] * jac).ravel()
        return pts, w, np.repeat(owner, m)

    quad_integration = quad

    def quadpol(self, ind=None):
        """Polar quadrature around each element centre (144 points per
        triangle, 192 per quadrilateral). Returns ``(pos, w, row)`` with
        triangles first, then quadrilaterals; ``row`` is the position in ``ind``
        (particle.py:1276-1429)."""
        ind = np.arange(self.nfaces) if ind is None else np.asarray(ind)
        q = self.quad_rule
        i3, i4 = self.index34(ind)
        P, W, R = [], [], []
        if len(i3):
            if self.interp == "flat":
                N = np.stack([q.x3, q.y3, 1 - q.x3 - q.y3], axis=1)
                nodes = self.verts[self.faces[ind[i3], :3].astype(np.int64)]
                jac = np.repeat(self.area[ind[i3]][:, None], len(q.w3), axis=1)
            else:
                N, dx, dy = tri6(q.x3, q.y3)
                nodes = self.verts2[self.faces2[ind[i3]][:, [0, 1, 2, 4, 5, 6]].astype(np.int64)]
                jac = 0.5 * np.linalg.norm(np.cross(_map(dx, nodes), _map(dy, nodes)), axis=2)
            P.append(_map(N, nodes).reshape(-1, 3))
            W.append((q.w3[None, :] * jac).ravel())
            R.append(np.repeat(i3, len(q.w3)))
        if len(i4):
            if self.interp == "flat":
                N, dx, dy = quad4(q.x4, q.y4)
                nodes = self.verts[self.faces[ind[i4], :4].astype(np.int64)]
            else:
                N, dx, dy = quad9(q.x4, q.y4)
                nodes = self.verts2[self.faces2[ind[i4]][:, :9].astype(np.int64)]
            jac = np.linalg.norm(np.cross(_map(dx, nodes), _map(dy, nodes)), axis=2)
            P.append(_map(N, nodes).reshape(-1, 3))
            W.append((q.w4[None, :] * jac).ravel())
            R.append(np.repeat(i4, len(q.w4)))
        if not P:
            return np.zeros((0, 3)), np.zeros(0), np.zeros(0, dtype=np.int64)
        return np.vstack(P), np.concatenate(W), np.concatenate(R)

    def bradius(self):
        """Radius of the sphere about each centroid enclosing the element corners
        (particle.py:2848-2880)."""
        r = np.zeros(self.nfaces)
        i3, i4 = self.index34()
        for idx, nc in ((i3, 3), (i4, 4)):
            if len(idx) == 0:
                continue
            for c in range(nc):
                v = self.verts[self.faces[idx, c].astype(np.int64)]
                r[idx] = np.maximum(r[idx], np.linalg.norm(self.pos[idx] - v, axis=1))
        return r

    # ---------------- transformations ----------------
    def shift(self, vec):
        vec = np.asarray(vec, float)
        self.verts = self.verts + vec
        if self.verts2 is not None:
            self.verts2 = self.verts2 + vec
        self._norm()
        return self

    def scale(self, s):
        s = np.asarray(s, float)
        self.verts = self.verts * s
        if self.verts2 is not None:
            self.verts2 = self.verts2 * s
        self._norm()
        return self

    def flipfaces(self):
        """Copy with reversed face orientation (particle.py:868-908)."""
        new = copy.deepcopy(self)
        i3, i4 = new.index34()
        if len(i3):
            new.faces[i3, :3] = new.faces[i3, :3][:, ::-1]
        if len(i4):
            new.faces[i4, :4] = new.faces[i4, :4][:, ::-1]
        if new.faces2 is not None:
            if len(i3):
                new.faces2[np.ix_(i3, [0, 1, 2, 4, 5, 6])] = new.faces2[i3][:, [2, 1, 0, 5, 4, 6]]
            if len(i4):
                new.faces2[np.ix_(i4, list(range(8)))] = new.faces2[i4][:, [3, 2, 1, 0, 6, 5, 4, 7]]
        new._norm()
        return new

    def flip(self, dirs=0):
        """Copy mirrored along axis/axes ``dirs`` with orientation restored (particle.py:843-866)."""
        new = copy.deepcopy(self)
        for d in np.atleast_1d(dirs):
            new.verts[:, d] = -new.verts[:, d]
            if new.verts2 is not None:
                new.verts2[:, d] = -new.verts2[:, d]
        return new.flipfaces()

    def curved(self):
        """Switch to curved interpolation (adds flat mid-points when missing)."""
        if self.verts2 is None:
            add_flat_midpoints(self)
        self.interp = "curv"
        self._norm()
        return self

    def flat(self):
        self.interp = "flat"
        self._norm()
        return self

    def vertcat(self, *others):
        """Concatenated surface (particle.py:1005-1044)."""
        verts, faces = [self.verts], [self.faces]
        verts2 = [self.verts2] if self.verts2 is not None else None
        faces2 = [self.faces2] if self.faces2 is not None else None
        nv, nv2 = len(self.verts), (len(self.verts2) if self.verts2 is not None else 0)
        for o in others:
            verts.append(o.verts)
            faces.append(o.faces + nv)
            nv += len(o.verts)
            if verts2 is not None and o.verts2 is not None:
                verts2.append(o.verts2)
                faces2.append(o.faces2 + nv2)
                nv2 += len(o.verts2)
        new = Particle(np.vstack(verts), np.vstack(faces), interp=self.interp, norm=False)
        new.quad_rule = self.quad_rule
        if verts2 is not None:
            new.verts2, new.faces2 = np.vstack(verts2), np.vstack(faces2)
        new._norm()
        return new

    __add__ = vertcat

    def select(self, index):
        """Split into (selected faces, the rest) (particle.py:910-999)."""
        index = np.asarray(index, dtype=np.int64)
        rest = np.setdiff1d(np.arange(self.nfaces), index)
        return self._compress(index), self._compress(rest)

    def _compress(self, index):
        if len(index) == 0:
            return Particle(np.zeros((0, 3)), np.zeros((0, 4)))
        faces = self.faces[index].copy()
        ok = ~np.isnan(faces)
        used = np.unique(faces[ok].astype(np.int64))
        remap = np.zeros(self.nverts, dtype=np.int64)
        remap[used] = np.arange(len(used))
        faces[ok] = remap[faces[ok].astype(np.int64)]
        new = Particle(self.verts[used], faces, interp=self.interp, norm=False)
        new.quad_rule = self.quad_rule
        if self.verts2 is not None:
            f2 = self.faces2[index].copy()
            ok2 = ~np.isnan(f2)
            used2 = np.unique(f2[ok2].astype(np.int64))
            remap2 = np.zeros(len(self.verts2), dtype=np.int64)
            remap2[used2] = np.arange(len(used2))
            f2[ok2] = remap2[f2[ok2].astype(np.int64)]
            new.verts2, new.faces2 = self.verts2[used2], f2
        new._norm()
        return new

    def clean(self, cutoff=1e-10):
        """Merge coincident corner vertices, turn quads with a repeated corner
        into triangles and drop faces of negligible area (particle.py:2331-2398).
        ``verts2``/``faces2`` are left as they are, as in the reference."""
        rounded = np.round(self.verts, 8)
        uniq, inv = np.unique(rounded, axis=0, return_inverse=True)
        inv = inv.ravel()
        if len(uniq) != len(self.verts):
            f = self.faces.copy()
            ok = ~np.isnan(f)
            f[ok] = inv[f[ok].astype(np.int64)]
            self.verts, self.faces = uniq, f
        for i in np.nonzero(~np.isnan(self.faces[:, 3]))[0]:
            v = self.faces[i].astype(np.int64)
            _, first = np.unique(v, return_index=True)
            if len(first) == 3:
                keep = np.sort(first)
                self.faces[i] = [v[keep[0]], v[keep[1]], v[keep[2]], np.nan]
        self._norm()
        good = np.nonzero(self.area > cutoff * np.mean(self.area))[0]
        if len(good) < self.nfaces:
            kept, _ = self.select(good)
            self.verts, self.faces, self.verts2, self.faces2 = kept.verts, kept.faces, kept.verts2, kept.faces2
            self._norm()
        return self

    def __repr__(self):
        return "Particle(nverts={}, nfaces={}, interp='{}')".format(self.nverts, self.nfaces, self.interp)


def add_flat_midpoints(p):
    """Attach straight-edge mid-points (and quad centres) as ``verts2``/``faces2``
    (reference geometry/mesh_generators.py:133-190, particle midpoints 'flat')."""
    i3, i4 = p.index34()
    f = p.faces
    edges, owner_col = [], []
    for idx, nc in ((i3, 3), (i4, 4)):
        for c in range(nc):
            e = np.stack([f[idx, c], f[idx, (c + 1) % nc]], axis=1).astype(np.int64)
            edges.append(np.sort(e, axis=1))
            owner_col.append((idx, 4 + c))
    alle = np.vstack(edges) if edges else np.zeros((0, 2), np.int64)
    uniq, inv = np.unique(alle, axis=0, return_inverse=True)
    inv = inv.ravel()
    nv = p.nverts
    verts2 = np.vstack([p.verts, 0.5 * (p.verts[uniq[:, 0]] + p.verts[uniq[:, 1]])])
    faces2 = np.hstack([f, np.full((p.nfaces, 5), np.nan)])
    off = 0
    for idx, col in owner_col:
        faces2[idx, col] = nv + inv[off:off + len(idx)]
        off += len(idx)
    if len(i4):
        c4 = f[i4, :4].astype(np.int64)
        faces2[i4, 8] = len(verts2) + np.arange(len(i4))
        verts2 = np.vstack([verts2, 0.25 * p.verts[c4].sum(axis=1)])
    p.verts2, p.faces2 = verts2, faces2
    return p
