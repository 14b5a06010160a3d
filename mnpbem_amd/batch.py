"""Particles of uneven face counts solved together (no counterpart in the reference, which runs
one particle at a time, one process per GPU: utils/multi_gpu.py:148-260).

A group of particles with face counts n_p <= n is one batch of (particle, wavelength) matrices of
size n. Particle p's Green matrices fill the leading n_p x n_p block of its entries (the
assembly writes them with leading dimension n); ``mnp_pad_batched`` turns the rest into a
constant diagonal block (G1, G2: 1, H1: 1, H2: -1; L1 = L2 = 1 and zero normals, positions,
areas and excitation on the padding faces). Every stage of the retarded chain then keeps the
matrices block diagonal, the padding rows of each solution are zero, and the particle block is
the particle's own solution: the batch runs the kernels of the single-particle sweep (assembly,
LU chain, plane-wave right-hand side, FP64 refinement, far field) with per-entry geometry (the
library's ``sface`` stride) and every launch covers all particles of the group.

``plan_groups`` cuts a set of particles, sorted by face count, into groups whose padding costs
at most a given fraction of the LU work and whose matrices fit a memory budget.
"""
import numpy as np
import torch

from . import engine
from ._lib import call
from .engine import C64, C128, _p, stream_handle
from .greenfun import CompGreenRet


def plan_groups(nfaces, nlam, max_waste=0.25, mem_bytes=120e9, precision="mixed", max_entries=1024):
    """Groups of particle indices (sorted by face count) for one batch each: the padded LU work
    (n_max^3 per entry) exceeds the particles' own (sum n_p^3) by at most ``max_waste``, and the
    group's matrices (7 factor-precision + 4 complex128 for mixed precision per entry) fit in
    ``mem_bytes``."""
    order = np.argsort(np.asarray(nfaces), kind="stable")
    per_entry = (7 * 8 + 4 * 16) if precision == "mixed" else (7 * (16 if precision == "double" else 8))
    groups, cur = [], []
    for i in order:
        cand = cur + [int(i)]
        nmax = max(nfaces[j] for j in cand)
        own = sum(float(nfaces[j]) ** 3 for j in cand)
        waste = len(cand) * float(nmax) ** 3 / own - 1.0
        mem = len(cand) * nlam * per_entry * float(nmax) ** 2
        if cur and (waste > max_waste or mem > mem_bytes or len(cand) * nlam > max_entries):
            groups.append(cur)
            cur = [int(i)]
        else:
            cur = cand
    if cur:
        groups.append(cur)
    return groups


class GroupGeometry:
    """Per-entry geometry of a particle group, padded to n faces: entry b = p nlam + l holds
    particle p at wavelength l (pos, nvec, area, side flags per medium; zeros on padding)."""

    def __init__(self, geos, nlam, n, dev):
        self.geos, self.nlam, self.n, self.dev = geos, nlam, n, dev
        self.sface = n
        P = len(geos)
        B = P * nlam
        self.batch = B
        pos = np.zeros((P, n, 3))
        nvec = np.zeros((P, n, 3))
        area = np.zeros((P, n))
        nact = np.zeros(P, np.int32)
        for p, g in enumerate(geos):
            cp = g.cp
            m = cp.nfaces
            pos[p, :m], nvec[p, :m], area[p, :m] = cp.pos, cp.nvec, cp.area
            nact[p] = m
        rep = lambda a: np.ascontiguousarray(np.repeat(a, nlam, axis=0))
        self.nact_np = np.repeat(nact, nlam)
        self.pos = torch.tensor(rep(pos), dtype=torch.float64, device=dev)
        self.nvec = torch.tensor(rep(nvec), dtype=torch.float64, device=dev)
        self.area = torch.tensor(rep(area), dtype=torch.float64, device=dev)
        self.nact = torch.tensor(self.nact_np, device=dev)
        self._side = {}

    def side(self, medium):
        if medium not in self._side:
            s = np.zeros((len(self.geos), self.n), np.int32)
            for p, g in enumerate(self.geos):
                io = g.cp.inout_faces
                s[p, :len(io)] = (io[:, 0] == medium) + 2 * (io[:, 1] == medium)
            self._side[medium] = torch.tensor(np.ascontiguousarray(np.repeat(s, self.nlam, axis=0)), device=self.dev)
        return self._side[medium]


class GroupSolver(engine.RetSolverBatch):
    """RetSolverBatch over a particle group: assembly per particle into its entries, padding,
    then the batched chain on per-entry geometry."""

    def __init__(self, ggeo, precision="mixed", nb=32, refine_tol=1e-12, max_refine=12, tuning=None):
        super().__init__(ggeo, ggeo.batch, precision, nb, refine_tol, max_refine, tuning)

    def _media(self, enei_list):
        g, nl, n = self.geo, len(enei_list), self.geo.n
        B = g.batch
        L1 = np.ones((B, n), complex)
        L2 = np.ones((B, n), complex)
        kvac = np.repeat(2 * np.pi / enei_list[None, :], len(g.geos), axis=0).ravel()
        kmats, epss = [], []
        for p, pg in enumerate(g.geos):
            km, l1, l2, _, kmat_np, eps_np = pg.media(enei_list)
            m = pg.n
            L1[p * nl:(p + 1) * nl, :m] = l1.cpu().numpy()
            L2[p * nl:(p + 1) * nl, :m] = l2.cpu().numpy()
            kmats.append(km)
            epss.append(eps_np)
        d = g.dev
        self.kmats = kmats
        self.kmat_np = np.concatenate([k.cpu().numpy() for k in kmats], axis=0)     # (B, nmat)
        self.eps_np = np.concatenate(epss, axis=0)
        self.L1 = torch.tensor(L1, dtype=C128, device=d)
        self.L2 = torch.tensor(L2, dtype=C128, device=d)
        self.kvac = torch.tensor(kvac, dtype=torch.float64, device=d)
        self.kmat = torch.tensor(self.kmat_np, dtype=C128, device=d)

    def _assemble(self, gb):
        g, nl = self.geo, len(self.enei)
        es = gb.G1.element_size()
        for p, pg in enumerate(g.geos):
            off = p * nl * gb.stride * es
            call("mnp_green_assemble", pg.n, pg.npart, pg.nmat, nl, pg.order, _p(pg.pos), _p(pg.nvec), _p(pg.area),
                 _p(pg.fpart), _p(pg.con_in), _p(pg.con_out), _p(self.kmats[p]), pg.nref, _p(pg.ref_row),
                 _p(pg.ref_col), _p(pg.ref_g), _p(pg.ref_f), gb.prec, _p(gb.G1) + off, _p(gb.G2) + off,
                 _p(gb.H1) + off, _p(gb.H2) + off, gb.ld, gb.stride, stream_handle())
        for t, d in ((gb.G1, 1.0), (gb.G2, 1.0), (gb.H1, 1.0), (gb.H2, -1.0)):
            call("mnp_pad_batched", gb.prec, g.n, g.batch, _p(g.nact), _p(t), gb.ld, gb.stride, d, 0.0,
                 stream_handle())

    def factor(self, enei_list):
        enei_list = np.atleast_1d(np.asarray(enei_list, float))
        if len(enei_list) * len(self.geo.geos) != self.batch:
            raise ValueError("the group holds {} entries; got {} wavelengths for {} particles".format(
                self.batch, len(enei_list), len(self.geo.geos)))
        g = self.geo
        self.enei = enei_list
        self.nlam = g.batch
        self._media(enei_list)
        if self.exact is not None:
            self._assemble(self.exact)
        self._assemble(self.fac)
        return self._factor_assembled()


def solve_particles(cps, enei_list, pol, direc, precision="mixed", greens=None, medium=1, max_waste=0.25,
                    mem_bytes=120e9, autotune=True, timings=None):
    """Extinction of a plane wave (polarisation pol, direction direc) for every particle of
    ``cps`` at every wavelength of ``enei_list``: (len(cps), nlam) array. Particles are grouped by
    ``plan_groups`` and each group is solved as one batch. ``timings`` (list) receives, per group,
    (particle indices, seconds from the group's start to its results)."""
    import time
    enei_list = np.atleast_1d(np.asarray(enei_list, float))
    nl = len(enei_list)
    greens = greens or [CompGreenRet(cp, cp) for cp in cps]
    geos = [g.geo for g in greens]
    nfaces = [cp.nfaces for cp in cps]
    ext = np.zeros((len(cps), nl))
    fallback = []
    for grp in plan_groups(nfaces, nl, max_waste, mem_bytes, precision):
        t0 = time.perf_counter()
        n = max(nfaces[i] for i in grp)
        gg = GroupGeometry([geos[i] for i in grp], nl, n, geos[grp[0]].dev)
        tuning = dict(gemm_split=3 if precision != "double" else 0)
        if autotune:
            from . import autotune as AT
            fdt = C128 if precision == "double" else C64
            tuning = AT.tune_lu(n, min(gg.batch, 32), fdt, gg.dev, gemm_split=tuning["gemm_split"])
        s = GroupSolver(gg, precision, tuning.get("nb", 32), tuning=tuning)
        s.factor(enei_list)
        kmed = s.kmat[:, medium - 1].contiguous()
        rhs = engine.planewave_rhs(gg, s.kvac, kmed, s.L1, s.L2, pol, direc, medium, gg.batch, gg.batch)
        x = s.solve(rhs)
        if precision == "mixed" and len(s.stats["fallback"]):
            fb = s.stats["fallback"]
            fallback.extend(int(b) for b in fb)
            sd = GroupSolver(gg, "double", s.nb)
            sd.factor(enei_list)
            x[torch.as_tensor(fb, device=x.device)] = sd.solve(rhs)[torch.as_tensor(fb, device=x.device)]
            del sd
        ff = engine.farfield(gg, x, s.kvac, kmed, np.asarray(direc, float)[None, :], medium, gg.batch)
        e = ff[:, 0, :3].cpu().numpy()
        km = s.kmat_np[:, medium - 1]
        val = np.real(4 * np.pi / km * np.imag(e @ np.conj(np.asarray(pol, float))))
        ext[grp] = val.reshape(len(grp), nl)
        if timings is not None:
            timings.append((list(grp), time.perf_counter() - t0))
        del s, x, rhs, gg
    return dict(ext=ext, fallback=fallback)
