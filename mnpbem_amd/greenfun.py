"""Composite Green functions and the field container that crosses the API.

``CompGreenRet(p, p)`` mirrors greenfun/compgreen_ret.py:14-858: wavelength-independent
set-up (refinement pairs, Taylor coefficients, closed-surface correction) runs once on the
host; ``eval(i, j, key, enei)`` assembles the requested region-pair matrix on the GPU.
"""
import numpy as np
import torch

from .engine import DeviceGeometry, GreenBatch, connectivity, device
from .nearfield import NearField


class CompStruct:
    """Attribute container (greenfun/compgreen_stat.py:1154-1420): particle ``p``,
    wavelength ``enei`` and named fields."""

    def __init__(self, p, enei, **fields):
        self.p = p
        self.enei = enei
        self._fields = {}
        for k, v in fields.items():
            self.set_field(k, v)

    def set_field(self, key, val):
        self._fields[key] = val
        object.__setattr__(self, key, val)

    def set(self, **fields):
        for k, v in fields.items():
            self.set_field(k, v)
        return self

    def get(self, key, default=None):
        if key == "enei":
            return self.enei
        return self._fields.get(key, default)

    def __getitem__(self, key):
        if key == "enei":
            return self.enei
        return self._fields[key]

    def __contains__(self, key):
        return key in self._fields

    def keys(self):
        return self._fields.keys()

    def __repr__(self):
        return "CompStruct(enei={}, fields={})".format(self.enei, list(self._fields))


class CompGreenRet:
    """Retarded Green function of a composite particle with itself.

    Options (reference defaults): ``RelCutoff=3``, ``AbsCutoff=0``, ``order=5``,
    ``deriv='cart'``."""

    def __init__(self, p1, p2, **op):
        if op.get("deriv", "cart") != "cart":
            raise NotImplementedError("only the Cartesian surface derivative (deriv='cart') is provided")
        from .compoint import ComPoint
        self._points = None
        self._split = None
        if isinstance(p1, ComPoint):
            # observer points against a particle: potentials and fields at points
            from .compoint import PointGreen
            from .nearfield import ORDER
            self.p1, self.p2, self.deriv = p1, p2, "cart"
            self._points = PointGreen(p1, p2, device(), op.get("RelCutoff", 3.0), op.get("AbsCutoff", 0.0),
                                      op.get("order", ORDER))
            self.nearfield, self._geo = None, None
            return
        if p1 is not p2:
            # two different particles: rows are p1's faces (observer centroids and normals), columns
            # p2's (sources), regions joined by material index (compgreen_ret.py:57-190, :435-505).
            # The entries are those of the composite of both, whose p1-row / p2-column block is
            # returned: the same far-field formula, the refinement of face pairs closer than
            # RelCutoff face radii, and no closed-surface correction (the reference applies it only
            # for p1 is p2).
            from .comparticle import ComParticle
            if not isinstance(p2, ComParticle) or not isinstance(p1, ComParticle):
                raise NotImplementedError("CompGreenRet(p1, p2) takes a ComParticle or ComPoint p1 and a "
                                          "ComParticle p2")
            if len(p1.eps) != len(p2.eps):
                raise ValueError("CompGreenRet: p1 and p2 must share the table of dielectric functions")
            comb = ComParticle(p1.eps, list(p1.p) + list(p2.p), np.vstack([p1.inout, p2.inout]))
            self._split = (p1.nfaces, p2.nfaces, len(p1.p), len(p2.p))
            src = comb
        else:
            src = p1
        self.p1, self.p2 = p1, p2
        self.deriv = "cart"
        self.nearfield = NearField(src, rel_cutoff=op.get("RelCutoff", 3.0), abs_cutoff=op.get("AbsCutoff", 0.0),
                                   order=op.get("order", 5), device=device())
        self._src = src
        if self._split is None:
            self.con = [[c for c in row] for row in connectivity(p1.inout)]
        else:
            full = connectivity(src.inout)
            a = self._split[2]
            self.con = [[full[r1][r2][:a, a:] for r2 in range(2)] for r1 in range(2)]
        self._geo = None

    @property
    def geo(self):
        if self._points is not None:
            raise NotImplementedError("for observer points the device path provides potential and field")
        if self._geo is None:
            self._geo = DeviceGeometry(self._src, self.nearfield, device())
        return self._geo

    _KEYS = ("G", "F", "H1", "H2", "Gp", "H1p", "H2p")

    def _struct(self, enei, pair=None):
        """Geometry descriptor for entry-wise kernels at one wavelength: (struct, keep-alive).
        ``pair`` = (i, j): both connectivity tables are region pair (i, j)'s, so that either
        side evaluates that pair alone."""
        from ._lib import GreenGeomC
        geo = self.geo
        kmat = geo.media([float(enei)])[0][0].contiguous()
        rowptr, rcol, rg, rf = geo.refined_csr()
        cin, cout = (geo.con_in, geo.con_out) if pair is None else (self._pair(*pair),) * 2
        g = GreenGeomC(geo.n, geo.npart, geo.order, *[t.data_ptr() for t in (
            geo.pos, geo.nvec, geo.area, geo.fpart, cin, cout, kmat, rowptr, rcol, rg, rf)])
        return g, kmat

    def _pair(self, i, j):
        """Connectivity table of region pair (i, j) alone (material per particle pair, 0 = not
        connected) on the device: the tables of the assembly without the subtracted
        cross-region terms of nested particles (compgreen_ret.py:435-505)."""
        if not hasattr(self, "_pairs"):
            self._pairs = {}
        if (i, j) not in self._pairs:
            t = np.asarray(self.con[i][j])
            if self._split is not None:
                # p1's particles first, then p2's: only the p1-row / p2-column block connects
                a, b = self._split[2], self._split[3]
                full = np.zeros((a + b, a + b), dtype=t.dtype)
                full[:a, a:] = t
                t = full
            self._pairs[(i, j)] = torch.tensor(np.ascontiguousarray(t).ravel().astype(np.int32),
                                               device=self.geo.dev)
        return self._pairs[(i, j)]

    def _check(self, i, j, key):
        if self._points is not None:
            raise NotImplementedError("for observer points the device path provides potential and field")
        if key not in self._KEYS:
            raise ValueError("Unknown key: {}".format(key))
        if not np.any(self.con[i][j]):
            return False
        return True

    def _nested(self):
        return self._split is None and bool(np.any(self.con[1][0]) or np.any(self.con[0][1]))

    def _block(self, m):
        """The p1-row / p2-column block of a matrix over the composite (two particles), else m."""
        if self._split is None:
            return m
        n1 = self._split[0]
        return m[:n1, ..., n1:]

    def eval(self, i, j, key, enei, ind=None):
        """Region-pair matrix, or 0 when the regions do not connect (compgreen_ret.py:507-553).
        Keys 'G', 'F', 'H1', 'H2' give (n, n) complex128 host arrays, 'Gp', 'H1p', 'H2p' give
        (n, 3, n). With ``ind`` (row-major linear indices i * n + j) only those entries:
        (len(ind),) or (len(ind), 3)."""
        if ind is not None:
            return self._pick(i, j, key, enei, ind)
        if not self._check(i, j, key):
            return 0
        if key in ("Gp", "H1p", "H2p"):
            return self.eval_block(i, j, key, enei, 0, self._dims()[1])
        # the assembly with region pair (i, j)'s table on both sides: G1 = G2 = G(i, j), and
        # H1 = F + 2 pi, H2 = F - 2 pi on the diagonal where the pair connects a face to itself
        geo = self.geo
        gb = GreenBatch(geo, 1, torch.complex128)
        kmat = geo.media([enei])[0]
        tables = geo.con_in, geo.con_out
        geo.con_in = geo.con_out = self._pair(i, j)
        try:
            gb.assemble(kmat, 1)
        finally:
            geo.con_in, geo.con_out = tables
        if key == "G":
            return self._block(GreenBatch.to_numpy(gb.G1, 0))
        if key == "F":
            return self._block(0.5 * (GreenBatch.to_numpy(gb.H1, 0) + GreenBatch.to_numpy(gb.H2, 0)))
        return self._block(GreenBatch.to_numpy(gb.H1 if key == "H1" else gb.H2, 0))

    def eval_block(self, i, j, key, enei, col_start, col_stop):
        """Columns [col_start, col_stop) of ``eval(i, j, key, enei)`` without the full matrix
        (compgreen_ret.py:555-737): (n, ncol), or (n, 3, ncol) for the derivative keys; zeros
        when the regions do not connect."""
        from ._lib import call
        from .engine import stream_handle
        from .hmatrix import C_addr
        c0, c1 = int(col_start), int(col_stop)
        nr, nc = self._dims()
        if not (0 <= c0 <= c1 <= nc):
            raise AssertionError("[error] eval_block col range ({}, {}) out of bounds for p2.n={}".format(c0, c1, nc))
        vec = key in ("Gp", "H1p", "H2p")
        shape = (nr, 3, c1 - c0) if vec else (nr, c1 - c0)
        if not self._check(i, j, key):
            return np.zeros(shape, complex)
        n, off = self.geo.n, self._col0()
        g, keep = self._struct(enei, (i, j))
        out = torch.empty((n, 3, c1 - c0) if vec else (n, c1 - c0), dtype=torch.complex128, device=self.geo.dev)
        call("mnp_green_block", C_addr(g), i, self._KEYS.index(key), off + c0, c1 - c0, out.data_ptr(),
             stream_handle())
        res = out[:nr].cpu().numpy()
        del keep
        return res

    def _dims(self):
        """Rows (p1 faces) and columns (p2 faces) of a region-pair matrix."""
        return (self._split[0], self._split[1]) if self._split is not None else (self.geo.n, self.geo.n)

    def _col0(self):
        """First column of p2's faces in the composite geometry."""
        return self._split[0] if self._split is not None else 0

    def _pick(self, i, j, key, enei, ind):
        """Entries at row-major linear indices (compgreen_ret.py:860-910)."""
        from ._lib import call
        from .engine import stream_handle
        from .hmatrix import C_addr
        ind = np.asarray(ind, dtype=np.int64).ravel()
        nr, nc = self._dims()
        vec = key in ("Gp", "H1p", "H2p")
        shape = (len(ind), 3) if vec else (len(ind),)
        if len(ind) and (ind.min() < 0 or ind.max() >= nr * nc):
            raise IndexError("linear index out of range for a {} x {} matrix".format(nr, nc))
        if not self._check(i, j, key) or len(ind) == 0:
            return np.zeros(shape, complex)
        if self._split is not None:
            ind = (ind // nc) * self.geo.n + self._col0() + ind % nc
        g, keep = self._struct(enei, (i, j))
        d_ind = torch.tensor(ind, device=self.geo.dev)
        out = torch.empty(shape, dtype=torch.complex128, device=self.geo.dev)
        call("mnp_green_pick", C_addr(g), i, self._KEYS.index(key), d_ind.data_ptr(), len(ind), out.data_ptr(),
             stream_handle())
        res = out.cpu().numpy()
        del keep
        return res

    # ------------------------------------------------------------ potentials and fields
    def _surface_eval(self, sig, inout):
        """Per polarisation, the 14 complex columns (phi, phip, a, ap, e, h) of region ``inout``
        on p1's faces, from the kernel that evaluates G, Gp and H entry-wise (no matrices):
        phi = G(r, 1) sig1 + G(r, 2) sig2, e = i k0 (G(r, 1) h1 + G(r, 2) h2) - Hp(r, 1) sig1 -
        Hp(r, 2) sig2 and so on, with r = inout and the self-term sign of H, Hp set by inout
        (compgreen_ret.py:911-1034). Each source region is one kernel pass with that region
        pair's table; for non-nested particles only the pair (r, r) connects."""
        from ._lib import call
        from .engine import stream_handle
        from .hmatrix import C_addr
        if inout not in (1, 2):
            raise ValueError("inout must be 1 (inside) or 2 (outside)")
        geo, enei = self.geo, float(sig.enei)
        nr, nc = self._dims()
        n, off = geo.n, self._col0()
        dev = geo.dev
        st = stream_handle()
        r = inout - 1
        total, single = None, None
        for src, (s, h) in enumerate(((sig.sig1, sig.h1), (sig.sig2, sig.h2))):
            s = np.asarray(s, complex)
            h = np.asarray(h, complex)
            if s.shape[0] != nc or h.shape[:2] != (nc, 3):
                raise ValueError("surface charges must have n = {} rows and currents shape (n, 3[, npol])"
                                 .format(nc))
            single = s.ndim == 1
            s2 = s.reshape(nc, -1)
            h3 = h.reshape(nc, 3, -1)
            npol = s2.shape[1]
            if h3.shape[2] != npol:
                raise ValueError("charges and currents have different numbers of polarisations")
            if not np.any(self.con[r][src]):
                continue
            # sources on p2's faces, placed at their columns of the composite (two particles)
            sd = torch.zeros((npol, n), dtype=torch.complex128, device=dev)
            hd = torch.zeros((npol, n, 3), dtype=torch.complex128, device=dev)
            sd[:, off:off + nc] = torch.tensor(np.ascontiguousarray(s2.T), dtype=torch.complex128, device=dev)
            hd[:, off:off + nc] = torch.tensor(np.ascontiguousarray(np.transpose(h3, (2, 0, 1))),
                                               dtype=torch.complex128, device=dev)
            g, kmat = self._struct(enei, (r, src))
            out = torch.empty((npol, n, 14), dtype=torch.complex128, device=dev)
            for q in range(npol):
                call("mnp_surface_eval", C_addr(g), r, 2 * np.pi / enei, sd[q].data_ptr(), hd[q].data_ptr(),
                     out[q].data_ptr(), st)
            o = np.transpose(out[:, :nr].cpu().numpy(), (1, 2, 0))     # (nr, 14, npol)
            del kmat
            total = o if total is None else total + o
        if total is None:
            total = np.zeros((nr, 14, npol), complex)
        if single:
            total = total[:, :, 0]
        return total

    def potential(self, sig, inout=1):
        """Potentials and surface derivatives inside (inout=1) or outside (inout=2) of the
        particle from a BEMRet solution (compgreen_ret.py:975-1034): CompStruct with
        phi1, phi1p, a1, a1p (or phi2, ...). For p1 = ComPoint the points have one region each
        (their own medium), so ``inout`` only names the fields (phi1 ... for 1, phi2 ... else)."""
        o = self._points.eval(sig, sig.enei) if self._points is not None else self._surface_eval(sig, inout)
        t = str(inout)
        return CompStruct(self.p1, sig.enei, **{"phi" + t: o[:, 0], "phi" + t + "p": o[:, 1],
                                                 "a" + t: o[:, 2:5], "a" + t + "p": o[:, 5:8]})

    def field(self, sig, inout=1):
        """Electric and magnetic fields inside (inout=1) or outside (inout=2) of the particle
        surface from a BEMRet solution (compgreen_ret.py:911-973): CompStruct with e, h
        (n, 3[, npol]). For p1 = ComPoint: the scattered fields at the points (``inout`` is not
        used: each point sees the side of the surface that faces its medium)."""
        o = self._points.eval(sig, sig.enei) if self._points is not None else self._surface_eval(sig, inout)
        return CompStruct(self.p1, sig.enei, e=o[:, 8:11], h=o[:, 11:14])


class CompGreenStat:
    """Quasistatic Green function G = a_j / d and surface derivative
    F = -n_i.(r_i - r_j) a_j / d^3 of a composite particle with itself, with refined self
    and near entries and the closed-surface correction (greenfun/compgreen_stat.py:70-528).

    They are the k -> 0 limit of the retarded assembly (phase 1, only the n = 0 Taylor
    coefficients), so the device assembly runs at k = 0 with every particle pair connected.
    ``G`` and ``F`` are host numpy (n, n); ``device_F()`` keeps F + 2 pi I on the GPU.

    As in the reference, the closed-surface correction is applied only when the particle
    was given closed arguments (compgreen_stat.py:154-163); ``diag_shift`` is then 0,
    otherwise it removes the correction that the retarded set-up folds in by default."""

    def __init__(self, p1, p2, **op):
        from .compoint import ComPoint
        self._points = None
        if isinstance(p1, ComPoint):
            # observer points: potentials and fields at points (k = 0 Green function)
            from .compoint import PointGreen
            from .nearfield import ORDER
            self.p1, self.p2 = p1, p2
            self._points = PointGreen(p1, p2, device(), op.get("RelCutoff", 3.0), op.get("AbsCutoff", 0.0),
                                      op.get("order", ORDER))
            return
        if p1 is not p2:
            raise NotImplementedError("CompGreenStat on the device covers p1 is p2 (the BEM solve path) and "
                                      "p1 = ComPoint with p2 a ComParticle (fields at points)")
        self.p1 = self.p2 = p1
        self.green = op.get("green") or CompGreenRet(p1, p1, **{k: v for k, v in op.items()
                                                                if k in ("RelCutoff", "AbsCutoff", "order")})
        self._dev = None
        corr = self.green.nearfield.closed_correction
        given = any(c is not None for c in p1.closed)
        self.diag_shift = np.zeros_like(corr) if given else -corr

    def _assemble(self):
        if self._dev is None:
            geo = self.green.geo
            gb = GreenBatch(geo, 1, torch.complex128)
            kmat = torch.zeros((1, geo.nmat), dtype=torch.complex128, device=geo.dev)
            con_in, con_out = geo.con_in, geo.con_out
            geo.con_in = geo.con_out = torch.full_like(con_in, 1)
            try:
                gb.assemble(kmat, 1)
            finally:
                geo.con_in, geo.con_out = con_in, con_out
            self._dev = gb
        return self._dev

    def device_F(self):
        """(G1 tensor, H1 tensor): G and F + 2 pi I (column-major, batch of one)."""
        gb = self._assemble()
        return gb.G1, gb.H1

    @property
    def G(self):
        return GreenBatch.to_numpy(self._assemble().G1, 0).real.copy()

    @property
    def F(self):
        m = GreenBatch.to_numpy(self._assemble().H1, 0).real.copy()
        return m + np.diag(self.diag_shift - 2 * np.pi)

    def potential(self, sig, inout=1):
        """phi = G sig, phip = H sig (compgreen_stat.py:1011-1050); for p1 = ComPoint on the
        device (phip = 0 for bare points), else from the host matrices."""
        t = "1" if inout == 1 else "2"
        if self._points is not None:
            o = self._points.eval(sig, sig.enei, static=True)
            return CompStruct(self.p1, sig.enei, **{"phi" + t: o[:, 0], "phi" + t + "p": o[:, 1]})
        s = np.asarray(sig.sig)
        H = self.F + (2 * np.pi if inout == 1 else -2 * np.pi) * np.eye(s.shape[0])
        return CompStruct(self.p1, sig.enei, **{"phi" + t: self.G @ s, "phi" + t + "p": H @ s})

    def field(self, sig, inout=1):
        """e = -Hp sig (compgreen_stat.py:974-1009): at points (p1 = ComPoint, Hp = Gp), or on the
        faces inside (inout=1, H1p = Gp + 2 pi n on the self terms) or outside (inout=2, H2p =
        Gp - 2 pi n), from the k = 0 entries on the device (``mnp_surface_eval`` without
        currents)."""
        if self._points is not None:
            o = self._points.eval(sig, sig.enei, static=True)
            return CompStruct(self.p1, sig.enei, e=o[:, 8:11])
        if np.any(self.diag_shift != 0):
            raise NotImplementedError("surface fields need the closed-surface correction of the particle "
                                      "(give ComParticle its closed arguments)")
        from ._lib import call, GreenGeomC
        from .engine import stream_handle
        from .hmatrix import C_addr
        if inout not in (1, 2):
            raise ValueError("inout must be 1 (inside) or 2 (outside)")
        geo = self.green.geo
        n = geo.n
        s = np.asarray(sig.sig, complex)
        single = s.ndim == 1
        s2 = s.reshape(n, -1)
        kmat = torch.zeros(geo.nmat, dtype=torch.complex128, device=geo.dev)
        ones = torch.ones_like(geo.con_in)
        rowptr, rcol, rg, rf = geo.refined_csr()
        g = GreenGeomC(n, geo.npart, geo.order, *[t.data_ptr() for t in (
            geo.pos, geo.nvec, geo.area, geo.fpart, ones, ones, kmat, rowptr, rcol, rg, rf)])
        sd = torch.tensor(np.ascontiguousarray(s2.T), dtype=torch.complex128, device=geo.dev)
        hz = torch.zeros((n, 3), dtype=torch.complex128, device=geo.dev)
        out = torch.empty((s2.shape[1], n, 14), dtype=torch.complex128, device=geo.dev)
        for q in range(s2.shape[1]):
            call("mnp_surface_eval", C_addr(g), inout - 1, 2 * np.pi / float(sig.enei), sd[q].data_ptr(),
                 hz.data_ptr(), out[q].data_ptr(), stream_handle())
        e = np.transpose(out.cpu().numpy()[:, :, 8:11], (1, 2, 0))
        del kmat, ones
        return CompStruct(self.p1, sig.enei, e=e[:, :, 0] if single else e)

