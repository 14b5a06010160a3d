"""Points in the dielectric environment of a ComParticle (geometry/compoint.py:102-330) and the
retarded Green function between such points and the particle's faces, for potentials and
fields at points (CompGreenRet(p1 = ComPoint, p2 = ComParticle), greenfun/compgreen_ret.py:
911-1034).

A point belongs to the medium on its side of the nearest face: outside when (x - x_face) . n
>= 0, else inside; points nearer than ``mindist`` to the surface are dropped (medium 0), and
``medium`` keeps only points in the media listed. The reference takes the nearest face of each
particle in turn, so with several particles the last one decides; here the nearest face over
all particles decides (the two agree for one particle and for points outside all particles).

Point-face pairs closer than the refinement cutoff (RelCutoff element radii of the source face,
or AbsCutoff) are integrated over the source face's quadrature points as Taylor coefficients in
ik (greenret_refined.py:287-420, the observer being the point); the other pairs use the
centroid formula. Entries and sums run on the GPU (``mnp_point_eval``).
"""
import numpy as np
import torch
from scipy.spatial import cKDTree

from .nearfield import ORDER, _grouped_points, _near_terms


class ComPoint:
    """Points grouped by medium. ``pos`` (n, 3) are the kept points ordered by medium (group
    after group), ``inout`` the medium of each group, ``ind`` the original indices of each
    group's points, ``index`` (n,) the original index of every kept point."""

    def __init__(self, p, pos, mindist=None, medium=None):
        pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
        if pos.shape[1] != 3:
            raise ValueError("ComPoint: positions must have shape (n, 3)")
        self.pin = p
        self.eps = p.eps
        d, j = cKDTree(p.pos).query(pos)
        part = p.face_particle[j]
        side = np.sum((pos - p.pos[j]) * p.nvec[j], axis=1) >= 0
        io = np.asarray(p.inout)
        med = np.where(side, io[part, 1], io[part, 0]).astype(int)
        if mindist is not None:
            md = np.broadcast_to(np.abs(np.asarray(mindist, float)), (io.shape[0],)) if np.ndim(mindist) else \
                np.full(io.shape[0], abs(float(mindist)))
            med[d < md[part]] = 0
        keep_media = np.unique(med[med > 0])
        if medium is not None:
            keep_media = np.intersect1d(keep_media, np.atleast_1d(medium))
        self.inout = keep_media.astype(int)
        self.ind = [np.nonzero(med == m)[0] for m in self.inout]
        self.index = np.concatenate(self.ind) if self.ind else np.zeros(0, int)
        self.pos = pos[self.index]
        self.medium = med[self.index]
        self.n = len(self.index)
        self.nvec = np.zeros((self.n, 3))
        self.area = np.ones(self.n)

    def __repr__(self):
        return "ComPoint(n={}, media={})".format(self.n, list(self.inout))


def point_refine_pairs(pt, p, rel_cutoff=3.0, abs_cutoff=0.0):
    """(rows, cols) of point-face pairs to refine, sorted by row then column: (d - rad_j) <
    abs_cutoff or (d - rad_j) / rad_j < rel_cutoff with rad_j the source face's radius
    (refine_utils.py:89-159; points carry no radius). Pairs at d = 0 are excluded."""
    rad = p.bradius()
    reach = float(np.max(rad, initial=0.0)) * (1 + rel_cutoff) + abs_cutoff + 1e-9
    cand = cKDTree(p.pos).query_ball_point(pt.pos, reach)
    counts = np.array([len(c) for c in cand], dtype=np.int64)
    rows = np.repeat(np.arange(pt.n), counts)
    cols = np.concatenate([np.asarray(c, dtype=np.int64) for c in cand]) if rows.size else np.zeros(0, np.int64)
    d = np.sqrt(np.sum((pt.pos[rows] - p.pos[cols]) ** 2, axis=1))
    dr = d - rad[cols]
    keep = ((dr < abs_cutoff) | (dr / rad[cols] < rel_cutoff)) & (d > 0)
    rows, cols = rows[keep], cols[keep]
    order = np.lexsort((cols, rows))
    return rows[order], cols[order]


def point_coefficients(pt, p, rows, cols, order=ORDER):
    """Taylor coefficients g (m, order+1), f (m, 3, order+1) of the refined point-face pairs,
    from the source faces' quadrature points (greenret_refined.py:287-420)."""
    g = np.zeros((len(rows), order + 1))
    f = np.zeros((len(rows), 3, order + 1))
    if not len(rows):
        return g, f
    src = np.unique(cols)
    slot = np.full(p.nfaces, -1, dtype=np.int64)
    slot[src] = np.arange(len(src))
    for sel, pts, w in _grouped_points(p, src, polar=False):
        grp = np.zeros(len(src), dtype=bool)
        grp[sel] = True
        local = np.full(len(src), -1, dtype=np.int64)
        local[sel] = np.arange(len(sel))
        pr = np.nonzero(grp[slot[cols]])[0]
        li = local[slot[cols[pr]]]
        gg, ff = _near_terms(pt.pos[rows[pr]], p.pos[cols[pr]], pts[li], w[li], order)
        g[pr] = gg
        f[pr] = ff
    return g, f


class PointGreen:
    """Device data of the Green function between a ComPoint and a ComParticle."""

    def __init__(self, pt, p, dev, rel_cutoff=3.0, abs_cutoff=0.0, order=ORDER):
        # a point in medium m sees every face that touches m, through sig1, h1 where m is the
        # face's inside and sig2, h2 where it is its outside; for nested particles a point in the
        # shell medium so sees the shell's inside and the core's outside (compgreen_ret.py:435-505)
        self.pt, self.p, self.dev, self.order = pt, p, dev, order
        rows, cols = point_refine_pairs(pt, p, rel_cutoff, abs_cutoff)
        g, f = point_coefficients(pt, p, rows, cols, order)
        rowptr = np.searchsorted(rows, np.arange(pt.n + 1)).astype(np.int32)
        T = lambda a, dt=None: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        self.ppos = T(pt.pos, torch.float64)
        self.pmed = T(pt.medium.astype(np.int32))
        self.fpos = T(p.pos, torch.float64)
        self.farea = T(p.area, torch.float64)
        self.fpart = T(p.face_particle.astype(np.int32))
        self.inout = T(np.asarray(p.inout, dtype=np.int32))
        self.rowptr, self.rcol = T(rowptr), T(cols.astype(np.int32))
        self.rg, self.rf = T(g, torch.float64), T(f, torch.float64)
        self.n_ref = len(rows)

    def eval(self, sig, enei, static=False):
        """(npt, 14, npol) or (npt, 14) complex: phi, phip, a, ap, e, h at the points.
        ``static``: the quasistatic Green function 1/r of every face with the one charge
        ``sig.sig`` (CompGreenStat(ComPoint, p), compgreen_stat.py:974-1050): k = 0, every
        point connected to every face, no currents; then e = -Gp sig."""
        from ._lib import call, PointGeomC
        from .engine import stream_handle
        from .hmatrix import C_addr
        p, n = self.p, self.p.nfaces
        enei = float(enei)
        if self.pt.n == 0:                     # no points kept (all dropped by mindist or medium)
            npol = 1 if np.ndim(sig.sig if static else sig.sig1) == 1 else np.shape(sig.sig if static else sig.sig1)[1]
            return np.zeros((0, 14) if npol == 1 else (0, 14, npol), complex)
        if static:
            kmat = torch.zeros(1, dtype=torch.complex128, device=self.dev)
            pmed = torch.ones_like(self.pmed)
            inout = torch.ones_like(self.inout)
            s1 = s2 = np.asarray(sig.sig, complex)
            h1 = h2 = np.zeros(s1.shape[:1] + (3,) + s1.shape[1:], complex)
        else:
            kmat = torch.tensor(np.array([complex(e(enei)[1]) for e in p.eps]), dtype=torch.complex128,
                                device=self.dev)
            pmed, inout = self.pmed, self.inout
            s1, h1 = np.asarray(sig.sig1, complex), np.asarray(sig.h1, complex)
            s2, h2 = np.asarray(sig.sig2, complex), np.asarray(sig.h2, complex)
        g = PointGeomC(self.pt.n, n, len(p.inout), self.order, *[t.data_ptr() for t in (
            self.ppos, pmed, self.fpos, self.farea, self.fpart, inout, kmat, self.rowptr, self.rcol,
            self.rg, self.rf)])
        single = s1.ndim == 1
        s1, s2 = s1.reshape(n, -1), s2.reshape(n, -1)
        h1, h2 = h1.reshape(n, 3, -1), h2.reshape(n, 3, -1)
        npol = s1.shape[1]
        T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.complex128, device=self.dev)
        ts1, ts2 = T(s1.T), T(s2.T)
        th1, th2 = T(np.transpose(h1, (2, 0, 1))), T(np.transpose(h2, (2, 0, 1)))
        out = torch.empty((npol, self.pt.n, 14), dtype=torch.complex128, device=self.dev)
        st = stream_handle()
        for q in range(npol):
            call("mnp_point_eval", C_addr(g), 2 * np.pi / enei, ts1[q].data_ptr(), th1[q].data_ptr(),
                 ts2[q].data_ptr(), th2[q].data_ptr(), out[q].data_ptr(), st)
        o = np.transpose(out.cpu().numpy(), (1, 2, 0))
        del kmat, pmed, inout
        return o[:, :, 0] if single else o


def _expand_grid(x, y, z):
    """Grid arrays of one shape from x, y, z as the reference's meshfield (meshfield.py:
    expand): equal shapes stay; two equal shapes and a scalar third broadcast; two equal
    shapes and a 1-D third add a trailing axis over the third."""
    arrs = [np.asarray(v, dtype=np.float64) for v in (x, y, z)]
    shapes = [a.shape for a in arrs]
    if len(set(shapes)) == 1:
        return arrs
    for i, j, k in ((0, 1, 2), (0, 2, 1), (1, 2, 0)):
        if shapes[i] == shapes[j]:
            base, third = arrs[i], arrs[k]
            out = [None, None, None]
            if third.size == 1:
                out[i], out[j] = arrs[i], arrs[j]
                out[k] = np.zeros_like(base) + float(third.ravel()[0])
            else:
                n2 = third.size
                tile = lambda a: np.repeat(a[..., None], n2, axis=-1)
                out[i], out[j] = tile(arrs[i]), tile(arrs[j])
                out[k] = np.broadcast_to(third.ravel(), base.shape + (n2,)).copy()
            return out
    return np.broadcast_arrays(*arrs)


class MeshField:
    """Scattered fields of a BEM solution on a grid (simulation/meshfield.py:18-520): grid
    points become a ComPoint (``mindist`` drops points near the surface; their values are NaN),
    fields come from ``CompGreenRet(ComPoint, p).field`` (``sim='ret'``) or
    ``CompGreenStat(ComPoint, p).field`` (``sim='stat'``, the reference's default) on the
    device, shaped as the grid + (3,). The incident field is the excitation's and is not added."""

    def __init__(self, p, x, y, z=None, nmax=None, mindist=None, sim="stat", **options):
        if sim not in ("stat", "ret"):
            raise ValueError("MeshField: sim must be 'stat' or 'ret'")
        from .greenfun import CompGreenRet, CompGreenStat
        self.x, self.y, self.z = _expand_grid(x, y, 0.0 if z is None else z)
        self.p, self.nmax, self.sim = p, nmax, sim
        pos = np.column_stack([self.x.ravel(), self.y.ravel(), self.z.ravel()])
        self.pt = ComPoint(p, pos, mindist=mindist)
        self.g = (CompGreenRet if sim == "ret" else CompGreenStat)(self.pt, p, **options)
        self._npos = len(pos)

    def _grid(self, v):
        v = np.asarray(v)
        out = np.full((self._npos,) + v.shape[1:], np.nan + 0j, dtype=complex)
        out[self.pt.index] = v
        return out.reshape(self.x.shape + v.shape[1:])

    def field(self, sig, inout=2):
        """(e, h): scattered fields on the grid, shape grid + (3,) [+ (npol,)]; h is None for
        sim='stat'."""
        f = self.g.field(sig, inout)
        return self._grid(f.e), (self._grid(f.h) if "h" in f else None)

    __call__ = field

    def potential(self, sig, inout=2):
        """phi and a on the grid (phi: grid [+ (npol,)], a: grid + (3,) [+ (npol,)]; a is None
        for sim='stat')."""
        pot = self.g.potential(sig, inout)
        t = str(inout)
        return self._grid(pot["phi" + t]), (self._grid(pot["a" + t]) if ("a" + t) in pot else None)
