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
        if any(getattr(g, "nested", False) for g in geos):
            raise NotImplementedError("particle groups hold separate particles; solve nested particles with BEMRet")
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
        # written in place once they exist: a captured graph keeps reading the same buffers
        for name, val, dt in (("L1", L1, C128), ("L2", L2, C128), ("kvac", kvac, torch.float64),
                              ("kmat", self.kmat_np, C128)):
            cur = getattr(self, name, None)
            if isinstance(cur, torch.Tensor) and tuple(cur.shape) == val.shape and cur.dtype == dt:
                cur.copy_(torch.from_numpy(np.ascontiguousarray(val)).to(dt))
            else:
                setattr(self, name, torch.tensor(np.ascontiguousarray(val), dtype=dt, device=d))

    def _assemble(self, gb, pads=True):
        g, nl = self.geo, len(self.enei)
        es = gb.G1.element_size()
        for p, pg in enumerate(g.geos):
            off = p * nl * gb.stride * es
            call("mnp_green_assemble", pg.n, pg.npart, pg.nmat, nl, pg.order, _p(pg.pos), _p(pg.nvec), _p(pg.area),
                 _p(pg.fpart), _p(pg.con_in), _p(pg.con_out), _p(self.kmats[p]), pg.nref, _p(pg.ref_row),
                 _p(pg.ref_col), _p(pg.ref_g), _p(pg.ref_f), gb.prec, _p(gb.G1) + off, _p(gb.G2) + off,
                 _p(gb.H1) + off, _p(gb.H2) + off, gb.ld, gb.stride, stream_handle())
        for t, d in ((gb.G1, 1.0), (gb.G2, 1.0), (gb.H1, 1.0), (gb.H2, -1.0)):
            if not pads:
                break
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


def _planewave_into(out, geo, kvac, kmed, L1, L2, pol, direc, medium, batch):
    n = geo.n
    pol = np.ascontiguousarray(np.asarray(pol, float).ravel())
    direc = np.ascontiguousarray(np.asarray(direc, float).ravel())
    call("mnp_planewave_rhs", n, batch, _p(geo.pos), _p(geo.nvec), _p(geo.side(medium)), geo.sface, _p(kmed),
         _p(kvac), _p(L1), _p(L2), pol.ctypes.data, direc.ctypes.data, _p(out), 8 * n, stream_handle())


class GraphedGroup:
    """The group's pipeline as two captured HIP graphs on one stream: F = padding, LU chain,
    plane-wave right-hand side, the first solve with the complex64 chain and its FP64 residual;
    R = one refinement step (residual cast, chain solve, FP64 update, new residual). The host
    replays F once and R until every entry's residual is at refine_tol or stops halving (those
    entries are redone with the complex128 chain, as the eager path does). Assembly (one launch
    set per particle) runs eagerly before F; workspaces are reserved before capture so nothing is
    allocated while the stream is captured, and no memset enters the graphs (flags, counters and
    norms are cleared by kernels)."""

    def __init__(self, solver, enei_list, pol, direc, medium):
        s, g = solver, solver.geo
        if s.precision != "mixed":
            raise ValueError("graph capture covers the mixed-precision pipeline")
        self.s, self.medium = s, medium
        B, n = g.batch, g.n
        d = g.dev
        self.stream = torch.cuda.Stream(device=d)
        self.rhs = torch.empty((B, 8, n), dtype=C128, device=d)
        self.r32 = torch.empty((B, 8, n), dtype=C64, device=d)
        self.dx = torch.empty_like(self.r32)
        self.x = torch.empty((B, 8, n), dtype=C128, device=d)
        self.r = torch.empty_like(self.x)
        self.norms = torch.zeros((B, 2), dtype=torch.float64, device=d)
        self.pol, self.direc = pol, direc
        s.enei = np.atleast_1d(np.asarray(enei_list, float))
        s.nlam = B
        s._media(s.enei)
        self.kmed = s.kmat[:, medium - 1].contiguous()
        g.side(medium)             # made before capture (a host-to-device copy)
        with torch.cuda.stream(self.stream):
            from .autotune import knobs
            with knobs(s.tuning):
                call("mnp_lu_reserve", n, B, s.nb, stream_handle())
        self.stream.synchronize()
        self.F = torch.cuda.CUDAGraph()
        self.R = torch.cuda.CUDAGraph()
        with knobs(s.tuning):
            with torch.cuda.graph(self.F, stream=self.stream):
                self._first()
            with torch.cuda.graph(self.R, stream=self.stream):
                self._step()

    def _pads(self):
        s, g = self.s, self.s.geo
        for gb in (s.exact, s.fac):
            for t, dval in ((gb.G1, 1.0), (gb.G2, 1.0), (gb.H1, 1.0), (gb.H2, -1.0)):
                call("mnp_pad_batched", gb.prec, g.n, g.batch, _p(g.nact), _p(t), gb.ld, gb.stride, dval, 0.0,
                     stream_handle())

    def _factor_call(self):
        s, g, f = self.s, self.s.geo, self.s.fac
        call("mnp_bemret_factor", s.prec, g.n, g.batch, _p(f.G1), _p(f.G2), _p(f.H1), _p(f.H2), _p(s.S1),
             _p(s.S2), _p(s.D), f.ld, f.stride, _p(s.ipiv), _p(s.perm), _p(s.info), _p(g.nvec), g.sface,
             _p(s.L1), _p(s.L2), _p(s.kvac), s.nb, stream_handle())

    def _first(self):
        s, g = self.s, self.s.geo
        n, B = g.n, g.batch
        self._pads()
        self._factor_call()
        _planewave_into(self.rhs, g, s.kvac, self.kmed, s.L1, s.L2, self.pol, self.direc, self.medium, B)
        call("mnp_cast_batched", 1, 8 * n, B, _p(self.rhs), 8 * n, _p(self.r32), 8 * n, stream_handle())
        s._chain(self.r32, self.dx)
        self.x.copy_(self.dx)      # x = first correction, widened to complex128 (an elementwise kernel)
        s.residual(self.x, self.rhs, self.r, self.norms)

    def _step(self):
        s, g = self.s, self.s.geo
        n, B = g.n, g.batch
        call("mnp_cast_batched", 1, 8 * n, B, _p(self.r), 8 * n, _p(self.r32), 8 * n, stream_handle())
        s._chain(self.r32, self.dx)
        call("mnp_accum_batched", 1, 8 * n, B, _p(self.dx), 8 * n, _p(self.x), 8 * n, None, stream_handle())
        s.residual(self.x, self.rhs, self.r, self.norms)

    def run(self, assemble=True):
        """Assembly (eager) + F + R until converged; returns the solution tensor and statistics."""
        s = self.s
        with torch.cuda.stream(self.stream):
            if assemble:
                s._assemble(s.exact, pads=False)
                s._assemble(s.fac, pads=False)
            self.F.replay()
            info = s.check_info()
            if np.any(info < 0):
                raise RuntimeError("LU panel synchronisation timed out (info {})".format(info))
            hist, prev = [], None
            stalled = np.zeros(s.geo.batch, bool)
            for it in range(s.max_refine + 1):
                nr = self.norms.cpu().numpy()
                rel = np.sqrt(nr[:, 0] / np.maximum(nr[:, 1], 1e-300))
                hist.append(rel)
                done = rel <= s.refine_tol
                if prev is not None:
                    stalled |= (~done) & (rel > 0.5 * prev)
                if np.all(done | stalled) or it == s.max_refine:
                    break
                prev = rel
                self.R.replay()
        torch.cuda.current_stream().wait_stream(self.stream)
        s.stats = dict(relres=hist[-1], history=np.array(hist), iterations=len(hist) - 1,
                       fallback=np.nonzero(~(hist[-1] <= s.refine_tol))[0])
        return self.x


def solve_particles(cps, enei_list, pol, direc, precision="mixed", greens=None, medium=1, max_waste=0.25,
                    mem_bytes=120e9, autotune=True, timings=None, graphs=False):
    """Extinction of a plane wave (polarisation pol, direction direc) for every particle of
    ``cps`` at every wavelength of ``enei_list``: (len(cps), nlam) array. Particles are grouped by
    ``plan_groups`` and each group is solved as one batch. ``timings`` (list) receives, per group,
    (particle indices, seconds from the group's start to its results). ``graphs=True`` (mixed
    precision) runs each group's pipeline as captured HIP graphs (GraphedGroup)."""
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
        tuning = dict(gemm_split=16 if precision != "double" else 0)
        if autotune:
            from . import autotune as AT
            fdt = C128 if precision == "double" else C64
            tuning = AT.tune_lu(n, min(gg.batch, 32), fdt, gg.dev, gemm_split=tuning["gemm_split"])
        s = GroupSolver(gg, precision, tuning.get("nb", 32), tuning=tuning)
        if graphs:
            gr = GraphedGroup(s, enei_list, pol, direc, medium)
            x = gr.run()
            kmed, rhs = gr.kmed, gr.rhs
        else:
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
