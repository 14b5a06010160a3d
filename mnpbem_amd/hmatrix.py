"""Cluster tree, block partition and H-matrices on the device.

Host part (set-up, O(n log n)):
  * ``ClusterTree`` (greenfun/clustertree.py:13-380): faces are split first by particle, then
    by bisection of the longest edge of the centroid bounding box, down to leaves of at most
    ``cleaf`` faces; every node keeps the bounding sphere (mid, rad) of its centroids and a
    contiguous index range in cluster order.
  * ``block_partition``: the block tree of two clusters is admissible (low rank) when both lie
    in one particle and ``fadmiss(rad1, rad2, dist)`` holds; leaf pairs that are not are dense
    (clustertree.py:304-376). ``make_kaware_fadmiss`` (greenfun/hmatrix.py:1284-1311) adds the
    wavelength constraint for retarded kernels.

Device part (libmnpbem_hip.so): every matrix entry is evaluated on the fly from geometry
(including the refined near-field entries), dense blocks are filled directly, admissible
blocks are compressed by partially pivoted ACA (hmatrix.py:159-336) with one workgroup per
block, and the matrix-vector product streams dense blocks (complex128) and low-rank factors
(complex64) with one workgroup per block.
"""
import numpy as np


class ClusterTree:
    """Binary cluster tree of face centroids.

    Attributes follow the reference: ``son`` (nodes, 2) child indices (-1 for leaves),
    ``cind`` (nodes, 2) inclusive index range in cluster order, ``mid``/``rad`` bounding
    spheres, ``ipart`` particle of the node (-1 when it spans several), ``ind`` (n, 2):
    column 0 maps cluster position -> face, column 1 face -> cluster position."""

    def __init__(self, pos, cleaf=32, ipart=None):
        self.pos = np.asarray(pos, dtype=np.float64)
        self.n = len(self.pos)
        self.cleaf = int(cleaf)
        self._ip = np.zeros(self.n, np.int64) if ipart is None else np.asarray(ipart, np.int64)
        son, cind, mid, rad, order = [], [], [], [], []

        def sphere(idx):
            p = self.pos[idx]
            lo, hi = p.min(axis=0), p.max(axis=0)
            return 0.5 * (lo + hi), 0.5 * float(np.linalg.norm(hi - lo))

        def add(start, idx):
            m, r = sphere(idx)
            son.append([-1, -1])
            cind.append([start, start + len(idx) - 1])
            mid.append(m)
            rad.append(r)
            return len(son) - 1

        def split(idx):
            ip = self._ip[idx]
            if np.unique(ip).size > 1:
                mask = ip == ip[0]
                return idx[mask], idx[~mask], True
            p = self.pos[idx]
            lo, hi = p.min(axis=0), p.max(axis=0)
            k = int(np.argmax(hi - lo))
            mask = p[:, k] < lo[k] + 0.5 * (hi[k] - lo[k])
            a, b = idx[mask], idx[~mask]
            if len(a) == 0:
                a, b = idx[:1], idx[1:]
            elif len(b) == 0:
                a, b = idx[:-1], idx[-1:]
            return a, b, False

        def bisect(node, idx, start):
            a, b, psplit = split(idx)
            n1 = add(start, a)
            n2 = add(start + len(a), b)
            son[node] = [n1, n2]
            for child, sub, st in ((n1, a, start), (n2, b, start + len(a))):
                if len(sub) > self.cleaf or psplit:
                    bisect(child, sub, st)
                else:
                    order.append((st, sub))

        root = add(0, np.arange(self.n))
        if self.n > self.cleaf or np.unique(self._ip).size > 1:
            bisect(root, np.arange(self.n), 0)
        else:
            order.append((0, np.arange(self.n)))
        self.son = np.array(son, np.int64)
        self.cind = np.array(cind, np.int64)
        self.mid = np.array(mid, np.float64)
        self.rad = np.array(rad, np.float64)
        c2p = np.empty(self.n, np.int64)
        for st, sub in order:
            c2p[st:st + len(sub)] = sub
        self.ind = np.empty((self.n, 2), np.int64)
        self.ind[:, 0] = c2p
        self.ind[c2p, 1] = np.arange(self.n)
        first = self._ip[c2p[self.cind[:, 0]]]
        last = self._ip[c2p[self.cind[:, 1]]]
        # a node lies in one particle when all of its faces do (particles are split first,
        # so first == last identifies it)
        self.ipart = np.where(first == last, first, -1)

    @property
    def nodes(self):
        return len(self.son)

    def is_leaf(self, i):
        return self.son[i, 0] == -1

    def size(self, i):
        return int(self.cind[i, 1] - self.cind[i, 0] + 1)


def make_kaware_fadmiss(k, eta0=2.5):
    """eta0 * min(rad) < dist, and for k * min(rad) > 1 also dist >= one wavelength."""
    if k is None or not np.isfinite(abs(k)) or abs(k) < 1e-12:
        return lambda r1, r2, d: eta0 * min(r1, r2) < d
    ka = abs(complex(k))
    lam = 2 * np.pi / ka

    def fadmiss(r1, r2, d):
        rmin = min(r1, r2)
        if eta0 * rmin >= d:
            return False
        if ka * rmin > 1.0 and d < lam:
            return False
        return True
    return fadmiss


def block_partition(tree, fadmiss=None, eta=2.5, max_lowrank=None):
    """List of (node1, node2, kind): kind 1 = admissible (low rank), 2 = dense leaf pair.

    ``max_lowrank``: admissible pairs with more rows or columns than this are split further
    (the device ACA keeps its pivot masks in LDS); None = no limit, as the reference."""
    if fadmiss is None:
        fadmiss = lambda r1, r2, d: eta * min(r1, r2) < d
    out = []
    stack = [(0, 0)]
    while stack:
        i1, i2 = stack.pop()
        if tree.is_leaf(i1) and tree.is_leaf(i2):
            out.append((i1, i2, 2))
            continue
        if tree.ipart[i1] != -1 and tree.ipart[i2] != -1:
            d = float(np.linalg.norm(tree.mid[i1] - tree.mid[i2]))
            small = max_lowrank is None or max(tree.size(i1), tree.size(i2)) <= max_lowrank
            if small and fadmiss(tree.rad[i1], tree.rad[i2], d):
                out.append((i1, i2, 1))
                continue
        c1 = [i1] if tree.is_leaf(i1) else list(tree.son[i1])
        c2 = [i2] if tree.is_leaf(i2) else list(tree.son[i2])
        for a in reversed(c1):
            for b in reversed(c2):
                stack.append((a, b))
    return out


def coverage(tree, blocks):
    """Number of times each (cluster row, cluster column) is covered (must be 1 everywhere;
    dense check for tests on small trees)."""
    c = np.zeros((tree.n, tree.n), np.int64)
    for i1, i2, _ in blocks:
        r0, r1 = tree.cind[i1]
        c0, c1 = tree.cind[i2]
        c[r0:r1 + 1, c0:c1 + 1] += 1
    return c


# ------------------------------------------------------------------------------ device part
ACA_MAXDIM = 8192
KINDS = ("G1", "G2", "H1", "H2")


class HMatrixRet:
    """G1, G2, H1, H2 of one wavelength as H-matrices on the device.

    Set-up: the block partition of ``tree`` (admissibility ``fadmiss``, default the k-aware
    rule with the vacuum wavenumber 2 pi / enei as aca_compgreen_ret.py:164-170), dense blocks
    evaluated entry by entry in complex128, admissible blocks compressed by ACA
    (``htol``, ``kmax``) in complex128 and stored as complex64 factors.
    ``stats`` reports block counts, ranks, compression and bytes per 8-column product."""

    def __init__(self, geo, tree, enei, htol=1e-6, kmax=100, eta=2.5, fadmiss=None, static=False,
                 kinds=(0, 1, 2, 3)):
        """``static=True``: the quasistatic kernels (k = 0, every particle pair connected; G1 = G2
        = G, H1 = F + 2 pi, H2 = F - 2 pi) as ACACompGreenStat (aca_compgreen_stat.py), with
        ``enei`` ignored. ``kinds``: which of G1, G2, H1, H2 get low-rank factors (the others
        keep rank 0 in the admissible blocks, so only their dense blocks are valid)."""
        import torch
        from . import engine
        from ._lib import call, GreenGeomC, HMatC
        self.geo, self.tree, self.enei = geo, tree, float(enei)
        self.htol, self.kmax = float(htol), int(kmax)
        self.static, self.kinds = bool(static), tuple(int(q) for q in kinds)
        if fadmiss is None:
            fadmiss = make_kaware_fadmiss(0.0 if static else 2 * np.pi / enei, eta)
        blocks = block_partition(tree, fadmiss, eta, max_lowrank=ACA_MAXDIM)
        d = geo.dev
        n = geo.n
        cind = tree.cind

        def rect(i1, i2):
            return (int(cind[i1, 0]), int(cind[i2, 0]), tree.size(i1), tree.size(i2))

        dense = np.array([rect(a, b) for a, b, k in blocks if k == 2], np.int32).reshape(-1, 4)
        low = np.array([rect(a, b) for a, b, k in blocks if k == 1], np.int32).reshape(-1, 4)
        self.nd, self.nlr = len(dense), len(low)
        dsz = dense[:, 2].astype(np.int64) * dense[:, 3]
        doff = np.concatenate([[0], np.cumsum(dsz)]).astype(np.int64)
        self.dstride = int(doff[-1])
        kr = np.minimum(np.minimum(low[:, 2], low[:, 3]), self.kmax).astype(np.int64)
        uws = np.concatenate([[0], np.cumsum(low[:, 2] * kr)]).astype(np.int64)
        vws = np.concatenate([[0], np.cumsum(low[:, 3] * kr)]).astype(np.int64)
        T = lambda a: torch.tensor(np.ascontiguousarray(a), device=d)
        self.c2p = T(tree.ind[:, 0].astype(np.int32))
        self.dblk = T(dense if self.nd else np.zeros((1, 4), np.int32))
        self.doff = T(doff[:-1] if self.nd else np.zeros(1, np.int64))
        self.lblk = T(low if self.nlr else np.zeros((1, 4), np.int32))
        self.D = torch.empty(max(4 * self.dstride, 1), dtype=torch.complex128, device=d)
        self.rank = torch.zeros((max(self.nlr, 1), 4), dtype=torch.int32, device=d)
        kmat, self.L1, self.L2, self.kvac, _, _ = geo.media([enei])
        self.kmat = kmat[0].contiguous()
        con_in, con_out = geo.con_in, geo.con_out
        if self.static:
            self.kmat = torch.zeros_like(self.kmat)
            con_in = con_out = torch.ones_like(geo.con_in)
        rowptr, rcol, rg, rf = geo.refined_csr()
        self._keep = (rowptr, rcol, rg, rf, con_in, con_out)
        g = GreenGeomC(n, geo.npart, geo.order, *[t.data_ptr() for t in (
            geo.pos, geo.nvec, geo.area, geo.fpart, con_in, con_out, self.kmat, rowptr, rcol, rg, rf)])
        self._g = g
        st = engine.stream_handle()
        gp = C_addr(g)
        call("mnp_hmat_fill_dense", gp, self.c2p.data_ptr(), self.nd, self.dblk.data_ptr(), self.doff.data_ptr(),
             self.D.data_ptr(), self.dstride, st)
        uoff = np.zeros((4, max(self.nlr, 1)), np.int64)
        voff = np.zeros((4, max(self.nlr, 1)), np.int64)
        ranks = np.zeros((max(self.nlr, 1), 4), np.int32)
        self.U = self.V = None
        if self.nlr:
            uws_t, vws_t = T(uws[:-1]), T(vws[:-1])
            Uw = torch.empty(int(uws[-1]), dtype=torch.complex128, device=d)
            Vw = torch.empty(int(vws[-1]), dtype=torch.complex128, device=d)
            Us, Vs, ubase, vbase = [], [], [0], [0]
            for kind in range(4):
                if kind not in self.kinds:
                    uoff[kind] = ubase[-1]
                    voff[kind] = vbase[-1]
                    continue
                # ACA of this kind, then its complex64 copy (the workspace is reused by the next kind)
                call("mnp_hmat_aca", gp, self.c2p.data_ptr(), kind, self.nlr, self.lblk.data_ptr(),
                     uws_t.data_ptr(), vws_t.data_ptr(), Uw.data_ptr(), Vw.data_ptr(), self.rank.data_ptr(),
                     self.htol, self.kmax, st)
                rk = self.rank[:, kind].cpu().numpy().astype(np.int64)
                ranks[:, kind] = rk
                usz, vsz = low[:, 2] * rk, low[:, 3] * rk
                uo = np.concatenate([[0], np.cumsum(usz)[:-1]]).astype(np.int64)
                vo = np.concatenate([[0], np.cumsum(vsz)[:-1]]).astype(np.int64)
                Uq = torch.empty(max(int(usz.sum()), 1), dtype=torch.complex64, device=d)
                Vq = torch.empty(max(int(vsz.sum()), 1), dtype=torch.complex64, device=d)
                uo_t, vo_t = T(uo), T(vo)
                call("mnp_hmat_compact", self.nlr, self.lblk.data_ptr(), kind, self.rank.data_ptr(),
                     uws_t.data_ptr(), vws_t.data_ptr(), Uw.data_ptr(), Vw.data_ptr(), uo_t.data_ptr(),
                     vo_t.data_ptr(), Uq.data_ptr(), Vq.data_ptr(), st)
                uoff[kind] = ubase[-1] + uo
                voff[kind] = vbase[-1] + vo
                ubase.append(ubase[-1] + Uq.numel())
                vbase.append(vbase[-1] + Vq.numel())
                Us.append(Uq)
                Vs.append(Vq)
            self.U = torch.cat(Us)
            self.V = torch.cat(Vs)
            self.uoff, self.voff = T(uoff), T(voff)
            torch.cuda.synchronize(d)
            del Uw, Vw
        else:
            self.uoff, self.voff = T(uoff), T(voff)
            self.U = torch.empty(1, dtype=torch.complex64, device=d)
            self.V = torch.empty(1, dtype=torch.complex64, device=d)
        self.ranks = ranks[: self.nlr]
        rmax = ranks[: self.nlr].max(axis=1) if self.nlr else np.zeros(1, np.int64)
        toff = np.concatenate([[0], np.cumsum(rmax.astype(np.int64) * 8)]).astype(np.int64)
        self.tmp = torch.empty(max(int(toff[-1]), 1), dtype=torch.complex128, device=d)
        self.toff = T(toff[:-1] if self.nlr else np.zeros(1, np.int64))
        self.desc = HMatC(n, self.nd, self.nlr, self.c2p.data_ptr(), self.dblk.data_ptr(), self.doff.data_ptr(),
                          self.D.data_ptr(), self.dstride, self.lblk.data_ptr(), self.uoff.data_ptr(),
                          self.voff.data_ptr(), self.U.data_ptr(), self.V.data_ptr(), self.rank.data_ptr(),
                          self.tmp.data_ptr(), self.toff.data_ptr())
        self._dense, self._low = dense, low
        self.stats = self._stats(n, dense, low, ranks[: self.nlr])

    @staticmethod
    def _stats(n, dense, low, ranks):
        dmn = int((dense[:, 2].astype(np.int64) * dense[:, 3]).sum()) if len(dense) else 0
        out = dict(n=n, dense_blocks=len(dense), lowrank_blocks=len(low), dense_entries=dmn)
        if len(low):
            mn = low[:, 2].astype(np.int64) + low[:, 3]
            lr_entries = [int((mn * ranks[:, q]).sum()) for q in range(4)]
            out.update(rank_max=int(ranks.max()), rank_mean=float(ranks.mean()), lowrank_entries=lr_entries)
        else:
            lr_entries = [0] * 4
        out["compression"] = [(dmn + e) / float(n * n) for e in lr_entries]
        # bytes one 8-column product of all four matrices streams: dense complex128 + factors complex64
        out["bytes_per_apply"] = int(sum(16 * dmn + 8 * e for e in lr_entries))
        return out

    @property
    def desc_ptr(self):
        return C_addr(self.desc)

    def matvec(self, kind, x):
        """H_kind @ x for x (n, r) complex128 in face order (r <= 8); returns (n, r)."""
        import torch
        from . import engine
        from ._lib import call
        kind = KINDS.index(kind) if isinstance(kind, str) else int(kind)
        n, r = x.shape
        c2p = self.c2p.long()
        xc = x[c2p].t().contiguous()                  # (r, n): column c at c*n
        yc = torch.zeros_like(xc)
        call("mnp_hmat_matvec", self.desc_ptr, kind, r, xc.data_ptr(), n, yc.data_ptr(), n, engine.stream_handle())
        y = torch.empty_like(x)
        y[c2p] = yc.t()
        return y

    def full(self, kind):
        """Dense (n, n) numpy matrix of kind in face order (for checks on small meshes)."""
        kind = KINDS.index(kind) if isinstance(kind, str) else int(kind)
        n = self.geo.n
        A = np.zeros((n, n), complex)
        D = self.D[kind * self.dstride:(kind + 1) * self.dstride].cpu().numpy()
        off = 0
        for r0, c0, m, k in self._dense:
            A[r0:r0 + m, c0:c0 + k] = D[off:off + m * k].reshape(k, m).T
            off += m * k
        if self.nlr:
            U = self.U.cpu().numpy().astype(complex)
            V = self.V.cpu().numpy().astype(complex)
            uo = self.uoff[kind].cpu().numpy()
            vo = self.voff[kind].cpu().numpy()
            for b, (r0, c0, m, k) in enumerate(self._low):
                rk = self.ranks[b, kind]
                if rk == 0:
                    continue
                Ub = U[uo[b]:uo[b] + m * rk].reshape(rk, m).T
                Vb = V[vo[b]:vo[b] + k * rk].reshape(rk, k).T
                A[r0:r0 + m, c0:c0 + k] = Ub @ Vb.T
        c2p = self.tree.ind[:, 0]
        out = np.empty_like(A)
        out[np.ix_(c2p, c2p)] = A
        return out

    def lowrank_blocks(self):
        """(r0, c0, m, n) of the admissible blocks in cluster order."""
        return self._low


def C_addr(struct):
    import ctypes
    return ctypes.addressof(struct)
