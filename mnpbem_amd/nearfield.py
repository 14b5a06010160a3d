"""Near-field refinement of the retarded Green function (set-up).

Everything here is wavelength independent and runs once per geometry (host numpy, or with
``device`` the near-pair integrals and the closed-surface column sums in HIP kernels):

* ``refine_pairs``: which (observer i, source j) face pairs are integrated
  rather than collocated: 2 = self term (centroid distance 0), 1 = near
  (d - r_j < RelCutoff * r_i or d - r_j < AbsCutoff), with r the element's
  bounding radius and the distance formed as |p|^2 + |q|^2 - 2 p.q
  (greenfun/refine_utils.py:12-158). Candidates come from a k-d tree, then the
  exact rule decides.
* ``refine_coefficients``: Taylor coefficients in (ik) of the integrated
  entries, order 5: g_n (scalar) and f_n (Cartesian 3-vector) so that a
  refined entry at wavenumber k is  G = sum_n g_n (ik)^n e^{ikd},
  F = n_i . sum_n f_n (ik)^n e^{ikd}  (greenfun/greenret_refined.py:122-516).
  Self terms use the polar rule of the element, near pairs its 28-point rule.
* ``closed_surface_correction``: the Fuchs-Liu correction that makes the
  area-weighted column sums of the quasistatic surface derivative equal
  -2 pi on each closed surface; added to the n = 0 coefficient of the self
  terms along the normal (greenfun/compgreen_ret.py:192-400,
  greenfun/compgreen_stat.py:219-464).
"""
from math import factorial

import numpy as np
from scipy.spatial import cKDTree

ORDER = 5
_TINY = np.finfo(float).eps


def refine_pairs(pos1, rad1, pos2, rad2, rel_cutoff=3.0, abs_cutoff=0.0):
    """Return (rows, cols, kind) of refined pairs, kind 2 = self, 1 = near,
    sorted by row then column."""
    reach = float(np.max(rad2, initial=0.0)) + max(rel_cutoff * float(np.max(rad1, initial=0.0)), abs_cutoff) + 1e-9
    tree = cKDTree(pos2)
    cand = tree.query_ball_point(pos1, reach * (1 + 1e-9) + 1e-12)
    counts = np.array([len(c) for c in cand], dtype=np.int64)
    rows = np.repeat(np.arange(len(pos1)), counts)
    cols = np.concatenate([np.asarray(c, dtype=np.int64) for c in cand]) if rows.size else np.zeros(0, np.int64)
    p, q = pos1[rows], pos2[cols]
    d2 = np.sum(p * p, axis=1) + np.sum(q * q, axis=1) - 2.0 * np.sum(p * q, axis=1)
    d2[d2 < 1e-10] = 0.0
    d = np.sqrt(d2)
    dr = d - rad2[cols]
    near = ((dr < abs_cutoff) | (dr / rad1[rows] < rel_cutoff)) & (d != 0)
    diag = d == 0
    keep = near | diag
    rows, cols = rows[keep], cols[keep]
    kind = np.where(diag[keep], 2, 1).astype(np.int8)
    order = np.lexsort((cols, rows))
    return rows[order], cols[order], kind[order]


def _powers(x, order):
    out = np.empty((order + 1,) + x.shape)
    out[0] = 1.0
    for n in range(1, order + 1):
        out[n] = out[n - 1] * x
    return out


def _self_terms(pos, nvec, t1, t2, pts, w, order):
    """Polar-rule self terms. pos/nvec/t1/t2 (m, 3); pts (m, q, 3); w (m, q)."""
    vec = pos[:, None, :] - pts
    r = np.maximum(np.sqrt(np.sum(vec * vec, axis=2)), _TINY)
    nr = np.einsum("md,mqd->mq", nvec, vec)
    in1 = np.einsum("md,mqd->mq", t1, vec)
    in2 = np.einsum("md,mqd->mq", t2, vec)
    rr = np.maximum(r, 1e-4 * r.max(axis=1, keepdims=True))
    inv_rr3 = 1.0 / rr ** 3
    rp = _powers(r, order)
    g = np.empty((len(pos), order + 1))
    f = np.empty((len(pos), 3, order + 1))
    for n in range(order + 1):
        fac = 1.0 / factorial(n)
        g[:, n] = np.sum(w * rp[n] / r, axis=1) * fac
        a1 = np.sum(w * (n - 1) * nr * rp[n] / r ** 3, axis=1) * fac
        a2 = np.sum(w * (n - 1) * in1 * rp[n] * inv_rr3, axis=1) * fac
        a3 = np.sum(w * (n - 1) * in2 * rp[n] * inv_rr3, axis=1) * fac
        f[:, :, n] = nvec * a1[:, None] + t1 * a2[:, None] + t2 * a3[:, None]
    return g, f


def _near_terms(pos_obs, pos_src, pts, w, order):
    """Near-pair terms: observers pos_obs (m, 3), source centroids pos_src (m, 3),
    source quadrature points pts (m, q, 3) and weights w (m, q)."""
    rv = pos_obs[:, None, :] - pts
    r = np.maximum(np.sqrt(np.sum(rv * rv, axis=2)), _TINY)
    r0 = np.sqrt(np.sum((pos_src - pos_obs) ** 2, axis=1))
    delta = r - r0[:, None]
    dp = _powers(delta, order)
    inv_r = 1.0 / r
    inv_r2 = inv_r * inv_r
    inv_r3 = inv_r2 * inv_r
    g = np.empty((len(pos_obs), order + 1))
    f = np.empty((len(pos_obs), 3, order + 1))
    for n in range(order + 1):
        g[:, n] = np.sum(w * dp[n] * inv_r, axis=1) / factorial(n)
        if n == 0:
            s = -inv_r3
        else:
            s = -dp[n] * inv_r3 / factorial(n) + dp[n - 1] * inv_r2 / factorial(n - 1)
        f[:, :, n] = np.einsum("mq,mqd->md", w * s, rv)
    return g, f


def _grouped_points(part, faces, polar):
    """Quadrature points of ``faces`` grouped by point count: list of
    (positions-in-faces, pts (m, q, 3), w (m, q))."""
    pts, w, owner = part.quadpol(faces) if polar else part.quad(faces)
    order = np.argsort(owner, kind="stable")
    pts, w, owner = pts[order], w[order], owner[order]
    counts = np.bincount(owner, minlength=len(faces))
    starts = np.concatenate([[0], np.cumsum(counts)])
    groups = []
    for q in np.unique(counts):
        sel = np.nonzero(counts == q)[0]
        idx = starts[sel][:, None] + np.arange(q)[None, :]
        groups.append((sel, pts[idx], w[idx]))
    return groups


def _near_terms_device(part, rows, cols, nidx, order, g, f, dev):
    """Near-pair coefficients on the device (``mnp_near_terms``), written into g[nidx], f[nidx]."""
    import torch
    from . import engine
    from ._lib import call
    pts, w, owner = part.quad(np.arange(part.nfaces))
    o = np.argsort(owner, kind="stable")
    qptr = np.concatenate([[0], np.cumsum(np.bincount(owner, minlength=part.nfaces))]).astype(np.int32)
    T = lambda a, dt=None: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    tp = T(part.pos, torch.float64)
    tq, tqp, tw = T(qptr), T(pts[o], torch.float64), T(w[o], torch.float64)
    tr, tc = T(rows[nidx].astype(np.int32)), T(cols[nidx].astype(np.int32))
    slot = T(np.arange(len(nidx), dtype=np.int64))
    tg = torch.empty((len(nidx), order + 1), dtype=torch.float64, device=dev)
    tf = torch.empty((len(nidx), 3, order + 1), dtype=torch.float64, device=dev)
    call("mnp_near_terms", len(nidx), tr.data_ptr(), tc.data_ptr(), slot.data_ptr(), tp.data_ptr(), tq.data_ptr(),
         tqp.data_ptr(), tw.data_ptr(), order, tg.data_ptr(), tf.data_ptr(), engine.stream_handle())
    g[nidx] = tg.cpu().numpy()
    f[nidx] = tf.cpu().numpy()


def refine_coefficients(part, rows, cols, kind, order=ORDER, chunk=200000, device=None):
    """Coefficients g (n_ref, order+1) and f (n_ref, 3, order+1) (real) for the
    pairs of ``refine_pairs`` on one surface ``part`` (pos, nvec, tvec1, tvec2,
    quad, quadpol). ``device``: a torch device to integrate the near pairs there
    (``mnp_near_terms``); self terms stay on the host."""
    n_ref = len(rows)
    g = np.zeros((n_ref, order + 1))
    f = np.zeros((n_ref, 3, order + 1))
    pos, nvec, t1, t2 = part.pos, part.nvec, part.tvec1, part.tvec2
    # self terms
    sidx = np.nonzero(kind == 2)[0]
    if len(sidx):
        faces = cols[sidx]
        for sel, pts, w in _grouped_points(part, faces, polar=True):
            obs = rows[sidx[sel]]
            gg, ff = _self_terms(pos[obs], nvec[obs], t1[obs], t2[obs], pts, w, order)
            g[sidx[sel]] = gg
            f[sidx[sel]] = ff
    # near pairs, grouped by source face
    nidx = np.nonzero(kind == 1)[0]
    if len(nidx) and device is not None:
        _near_terms_device(part, rows, cols, nidx, order, g, f, device)
    elif len(nidx):
        src = np.unique(cols[nidx])
        slot = np.full(part.nfaces, -1, dtype=np.int64)
        slot[src] = np.arange(len(src))
        for sel, pts, w in _grouped_points(part, src, polar=False):
            grp = np.zeros(len(src), dtype=bool)
            grp[sel] = True
            local = np.full(len(src), -1, dtype=np.int64)
            local[sel] = np.arange(len(sel))
            pairs = nidx[grp[slot[cols[nidx]]]]
            for s in range(0, len(pairs), chunk):
                pr = pairs[s:s + chunk]
                li = local[slot[cols[pr]]]
                gg, ff = _near_terms(pos[rows[pr]], pos[cols[pr]], pts[li], w[li], order)
                g[pr] = gg
                f[pr] = ff
    return g, f


def _colsum_far_device(part, loc, ind, dev):
    import torch
    from . import engine
    from ._lib import call
    T = lambda a, dt=None: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    tl, ti = T(loc.astype(np.int32)), T(ind.astype(np.int32))
    out = torch.empty(len(ind), dtype=torch.float64, device=dev)
    tp, tn, ta = T(part.pos, torch.float64), T(part.nvec, torch.float64), T(part.area, torch.float64)
    call("mnp_colsum_far", len(loc), tl.data_ptr(), len(ind), ti.data_ptr(), tp.data_ptr(), tn.data_ptr(),
         ta.data_ptr(), out.data_ptr(), engine.stream_handle())
    return out.cpu().numpy()


def stat_colsum(part, rows, cols, kind, f0_normal, loc, ind, chunk=4096, device=None):
    """f_j = sum_{i in loc} a_i F_ij / a_j for j in ``ind``, with F the refined
    quasistatic surface derivative: far entries -n_i.(r_i - r_j) a_j / d^3,
    refined entries ``f0_normal`` (= n_i . f_0 of the pair). ``device``: a torch
    device for the O(n^2) collocated sums (``mnp_colsum_far``)."""
    pos, nvec, area = part.pos, part.nvec, part.area
    loc = np.asarray(loc)
    ind = np.asarray(ind)
    if device is not None:
        acc = _colsum_far_device(part, loc, ind, device)
    else:
        acc = np.zeros(len(ind))
        for s in range(0, len(loc), chunk):
            li = loc[s:s + chunk]
            r = pos[li][:, None, :] - pos[ind][None, :, :]
            d = np.maximum(np.sqrt(np.sum(r * r, axis=2)), _TINY)
            nr = np.einsum("id,ijd->ij", nvec[li], r)
            acc += np.sum(-area[li][:, None] * nr / d ** 3, axis=0)
    # swap the collocated value of refined pairs for the integrated one
    col_slot = np.full(part.nfaces, -1, dtype=np.int64)
    col_slot[ind] = np.arange(len(ind))
    in_loc = np.zeros(part.nfaces, dtype=bool)
    in_loc[loc] = True
    m = (col_slot[cols] >= 0) & in_loc[rows]
    r_, c_ = rows[m], cols[m]
    rv = pos[r_] - pos[c_]
    d = np.maximum(np.sqrt(np.sum(rv * rv, axis=1)), _TINY)
    far = -np.sum(nvec[r_] * rv, axis=1) * area[c_] / d ** 3
    far[kind[m] == 2] = 0.0
    np.add.at(acc, col_slot[c_], area[r_] * (f0_normal[m] - far) / area[c_])
    return acc


class NearField:
    """Refinement data of a composite particle: pairs, coefficients, and the
    closed-surface correction already folded into the self terms.

    ``device`` (a torch device) integrates the near pairs and the O(n^2) column sums of the
    closed-surface correction in HIP kernels; None keeps every step on the host."""

    def __init__(self, cp, rel_cutoff=3.0, abs_cutoff=0.0, order=ORDER, device=None):
        part = cp.pc
        rad = part.bradius()
        self.order = order
        self.rows, self.cols, self.kind = refine_pairs(part.pos, rad, part.pos, rad, rel_cutoff, abs_cutoff)
        self.g, self.f = refine_coefficients(part, self.rows, self.cols, self.kind, order, device=device)
        nvec = part.nvec
        f0n = np.einsum("md,md->m", nvec[self.rows], self.f[:, :, 0])
        self.closed_correction = np.zeros(part.nfaces)
        diag_slot = np.full(part.nfaces, -1, dtype=np.int64)
        sd = np.nonzero(self.kind == 2)[0]
        diag_slot[self.rows[sd]] = sd
        for ind, loc, direction in cp.closed_sets():
            fj = stat_colsum(part, self.rows, self.cols, self.kind, f0n, loc, ind, device=device)
            corr = -2.0 * np.pi * direction - fj
            self.closed_correction[ind] = corr
            slots = diag_slot[ind]
            ok = slots >= 0
            self.f[slots[ok], :, 0] += corr[ok][:, None] * nvec[ind[ok]]

    @property
    def n_ref(self):
        return len(self.rows)
