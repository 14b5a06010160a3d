"""Iterative retarded BEM solver (reference bem/bem_ret_iter.py:455-1877, bem/bem_iter.py:11-244).

The unknown is the packed 8N vector x = [sig1 h1 sig2 h2]; the operator is the exact one of
``BEMRetIter._afun`` (4 products G1/G2/H1/H2 with 4 columns, Eqs. 10, 11, 14, 17), applied on
the GPU (``mnp_bem_operator``). The Krylov solver is restarted GMRES with right
preconditioning, classical Gram-Schmidt with one re-orthogonalisation, Givens rotations; the
basis lives in HBM (``mnp_mdot`` / ``mnp_mupdate``), only the (restart+1) x restart
Hessenberg matrix is on the host. ``solver='bicgstab'`` and ``'cgs'`` run right-preconditioned
BiCGSTAB and CGS on the same device vectors. Preconditioner 'bemret' is the BEMRet LU chain of the
same wavelength (as the reference's default, bem_ret_iter.py:801-1086), in complex64
(``precond_precision='single'``) or complex128; 'none' runs unpreconditioned.

``hmatrix=True`` (options ``htol``, ``kmax``, ``cleaf``, ``eta``, ``fadmiss`` as the reference,
bem_ret_iter.py:488-580) replaces the dense operator by H-matrices: cluster tree and block
partition on the host, entries, ACA and products on the device (``mnpbem_amd.hmatrix``).

Differences from the reference: the solver accepts an initial guess ``x0`` (the reference
always starts from zero, bem_ret_iter.py:1695), and the residual reported by ``info()`` is the
true relative residual |b - A x| / |b| recomputed on the device after the last restart.
"""
import numpy as np
import torch

from . import engine
from ._lib import call, stage
from .bem import _npol, _pack_rhs
from .greenfun import CompGreenRet, CompStruct


class DeviceGMRES:
    """Restarted right-preconditioned GMRES on packed vectors of length L = 8 n (batch 1)."""

    def __init__(self, L, restart, dev):
        self.L, self.m, self.dev = L, restart, dev
        self.V = torch.empty((restart + 1, L), dtype=torch.complex128, device=dev)
        self.Z = torch.empty((restart, L), dtype=torch.complex128, device=dev)
        self.h = torch.empty(2 * (restart + 1), dtype=torch.float64, device=dev)
        self.s = torch.empty(2 * (restart + 1), dtype=torch.float64, device=dev)

    def _dot(self, nv, V, w):
        call("mnp_mdot", self.L, nv, V.data_ptr(), self.L, w.data_ptr(), self.h.data_ptr(), engine.stream_handle())
        hv = self.h[: 2 * nv].cpu().numpy()
        return hv[0::2] + 1j * hv[1::2]

    def _update(self, nv, V, coef, w):
        c = np.zeros(2 * nv)
        c[0::2], c[1::2] = coef.real, coef.imag
        self.s[: 2 * nv] = torch.from_numpy(c).to(self.dev)
        call("mnp_mupdate", self.L, nv, V.data_ptr(), self.L, self.s.data_ptr(), w.data_ptr(), engine.stream_handle())

    def solve(self, apply_A, apply_M, b, x0, tol, maxit):
        """Returns (x, flag, relres, iterations, history of relative residual estimates)."""
        L, m = self.L, self.m
        x = x0.clone() if x0 is not None else torch.zeros(L, dtype=torch.complex128, device=self.dev)
        bnorm = float(torch.linalg.vector_norm(b).item())
        if bnorm == 0:
            return torch.zeros_like(b), 0, 0.0, 0, [0.0]
        hist = []
        it = 0
        r = torch.empty_like(b)
        flag = 1
        while it < maxit:
            apply_A(x, r, rhs=b)                         # r = b - A x
            beta = float(torch.linalg.vector_norm(r).item())
            hist.append(beta / bnorm)
            if beta / bnorm <= tol:
                flag = 0
                break
            self.V[0].copy_(r / beta)
            H = np.zeros((m + 1, m), complex)
            cs = np.zeros(m, complex)
            sn = np.zeros(m, complex)
            g = np.zeros(m + 1, complex)
            g[0] = beta
            j_done = 0
            for j in range(m):
                with stage("gmres_iter"):
                    apply_M(self.V[j], self.Z[j])
                    w = self.V[j + 1]
                    apply_A(self.Z[j], w, rhs=None)
                    hcol = np.zeros(j + 2, complex)
                    for _ in range(2):                       # CGS2
                        hj = self._dot(j + 1, self.V, w)
                        self._update(j + 1, self.V, -hj, w)
                        hcol[: j + 1] += hj
                    hn = float(torch.linalg.vector_norm(w).item())
                    hcol[j + 1] = hn
                    if hn > 0:
                        w.div_(hn)
                    # Givens rotations G_i = [[conj(c), conj(s)], [-s, c]], c = a/rho, s = b/rho
                    for i in range(j):
                        t = np.conj(cs[i]) * hcol[i] + np.conj(sn[i]) * hcol[i + 1]
                        hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1]
                        hcol[i] = t
                    den = np.sqrt(abs(hcol[j]) ** 2 + abs(hcol[j + 1]) ** 2)
                    cs[j] = hcol[j] / den if den else 1.0
                    sn[j] = hcol[j + 1] / den if den else 0.0
                    hcol[j] = den
                    hcol[j + 1] = 0
                    g[j + 1] = -sn[j] * g[j]
                    g[j] = np.conj(cs[j]) * g[j]
                    H[: j + 2, j] = hcol
                    it += 1
                    j_done = j + 1
                    hist.append(abs(g[j + 1]) / bnorm)
                if abs(g[j + 1]) / bnorm <= tol or it >= maxit or hn == 0:
                    break
            k = j_done
            y = np.linalg.solve(np.triu(H[:k, :k]), g[:k]) if k else np.zeros(0)
            if k:
                self._update(k, self.Z, y, x)
        apply_A(x, r, rhs=b)
        relres = float(torch.linalg.vector_norm(r).item()) / bnorm
        if relres <= tol:
            flag = 0
        return x, flag, relres, it, hist


class DeviceBiCGSTAB:
    """Right-preconditioned BiCGSTAB (van der Vorst) on device vectors, the reference's
    solver='bicgstab' (bem_iter.py:64-68, scipy's bicgstab): solves A M y = b, x = M y.
    Returns (x, flag, relres, iterations, history) as DeviceGMRES; relres is recomputed."""

    def __init__(self, L, dev):
        self.L, self.dev = L, dev

    def solve(self, apply_A, apply_M, b, x0, tol, maxit):
        x = x0.clone() if x0 is not None else torch.zeros_like(b)
        bnorm = float(torch.linalg.vector_norm(b).item())
        if bnorm == 0:
            return torch.zeros_like(b), 0, 0.0, 0, [0.0]
        r = torch.empty_like(b)
        apply_A(x, r, rhs=b)
        rh = r.clone()
        p, v = torch.zeros_like(b), torch.zeros_like(b)
        ph, sh, t = torch.empty_like(b), torch.empty_like(b), torch.empty_like(b)
        rho = alpha = omega = 1.0 + 0j
        hist = [float(torch.linalg.vector_norm(r).item()) / bnorm]
        flag, it = 1, 0
        while it < maxit and hist[-1] > tol:
            rho_new = complex(torch.vdot(rh, r).item())
            if rho_new == 0:
                break
            beta = (rho_new / rho) * (alpha / omega)
            p.sub_(v, alpha=omega).mul_(beta).add_(r)
            apply_M(p, ph)
            apply_A(ph, v, rhs=None)
            alpha = rho_new / complex(torch.vdot(rh, v).item())
            s = r - alpha * v
            it += 1
            sn = float(torch.linalg.vector_norm(s).item()) / bnorm
            if sn <= tol:
                x.add_(ph, alpha=alpha)
                hist.append(sn)
                break
            apply_M(s, sh)
            apply_A(sh, t, rhs=None)
            tt = float(torch.linalg.vector_norm(t).item()) ** 2
            omega = complex(torch.vdot(t, s).item()) / tt if tt > 0 else 0j
            x.add_(ph, alpha=alpha).add_(sh, alpha=omega)
            r = s - omega * t
            hist.append(float(torch.linalg.vector_norm(r).item()) / bnorm)
            rho = rho_new
            if omega == 0:
                break
        apply_A(x, r, rhs=b)
        relres = float(torch.linalg.vector_norm(r).item()) / bnorm
        if relres <= tol:
            flag = 0
        return x, flag, relres, it, hist


class DeviceCGS:
    """Right-preconditioned conjugate gradients squared (Sonneveld), the reference's
    solver='cgs' (bem_iter.py:60-63, scipy's cgs). Returns as DeviceGMRES."""

    def __init__(self, L, dev):
        self.L, self.dev = L, dev

    def solve(self, apply_A, apply_M, b, x0, tol, maxit):
        x = x0.clone() if x0 is not None else torch.zeros_like(b)
        bnorm = float(torch.linalg.vector_norm(b).item())
        if bnorm == 0:
            return torch.zeros_like(b), 0, 0.0, 0, [0.0]
        r = torch.empty_like(b)
        apply_A(x, r, rhs=b)
        rh = r.clone()
        p, q = torch.zeros_like(b), torch.zeros_like(b)
        ph, vh, uh, au = (torch.empty_like(b) for _ in range(4))
        rho_prev = 1.0 + 0j
        hist = [float(torch.linalg.vector_norm(r).item()) / bnorm]
        flag, it = 1, 0
        while it < maxit and hist[-1] > tol:
            rho = complex(torch.vdot(rh, r).item())
            if rho == 0:
                break
            if it == 0:
                u = r.clone()
                p.copy_(u)
            else:
                beta = rho / rho_prev
                u = r + beta * q
                p.mul_(beta).add_(q).mul_(beta).add_(u)
            apply_M(p, ph)
            apply_A(ph, vh, rhs=None)
            alpha = rho / complex(torch.vdot(rh, vh).item())
            q = u - alpha * vh
            apply_M(u + q, uh)
            x.add_(uh, alpha=alpha)
            apply_A(uh, au, rhs=None)
            r.sub_(au, alpha=alpha)
            it += 1
            hist.append(float(torch.linalg.vector_norm(r).item()) / bnorm)
            rho_prev = rho
        apply_A(x, r, rhs=b)
        relres = float(torch.linalg.vector_norm(r).item()) / bnorm
        if relres <= tol:
            flag = 0
        return x, flag, relres, it, hist


def make_solver(name, L, restart, dev):
    """Device Krylov solver by the reference's name (bem_iter.py:11-17)."""
    if name == "gmres":
        return DeviceGMRES(L, restart, dev)
    if name == "bicgstab":
        return DeviceBiCGSTAB(L, dev)
    if name == "cgs":
        return DeviceCGS(L, dev)
    raise ValueError("[error] iterative solver not known: <{}>".format(name))


class BEMRetIter:
    """Iterative BEMRet. Options (reference names): ``solver='gmres'|'bicgstab'|'cgs'``, ``tol=1e-4``,
    ``maxit=200``, ``restart=None`` (= min(n, maxit)), ``precond='hmat'|'bemret'|None``,
    ``preconditioner='auto'|'none'``, ``precond_precision='single'|'double'``,
    ``precond_gemm_split=2|3|0`` (not in the reference), ``hmatrix``,
    ``htol=1e-6``, ``kmax=[4, 100]``, ``cleaf=200``, ``eta=2.5``, ``fadmiss``; not in the
    reference: ``precond_wavelength`` (factor the preconditioner at that wavelength once and
    reuse it for a sweep) and ``solve(exc, x0=...)``."""

    def __init__(self, p, enei=None, **op):
        if p is None:
            raise ValueError("BEMRetIter: 'p' must be a ComParticle, got None.")
        self.hmatrix = bool(op.get("hmatrix", False))
        top = lambda v: max(v) if hasattr(v, "__iter__") else v
        self.htol = float(top(op.get("htol", 1e-6)))
        self.kmax = int(top(op.get("kmax", [4, 100])))
        self.cleaf = int(op.get("cleaf", 200))
        self.eta = float(op.get("eta", 2.5))
        self.fadmiss = op.get("fadmiss", None)
        self.p = p
        self.solver = op.get("solver", "gmres")
        if self.solver not in ("gmres", "bicgstab", "cgs"):
            raise ValueError("[error] iterative solver not known: <{}>".format(self.solver))
        self.tol = float(op.get("tol", 1e-4))
        self.maxit = int(op.get("maxit", 200))
        self.restart = op.get("restart", None)
        pc = op.get("preconditioner", "auto")
        if pc not in ("auto", "none", "dense", "hlu_dense", "tree", "hlu_tree", "particle"):
            raise ValueError("BEMRetIter: unknown preconditioner {!r}".format(pc))
        # 'particle' (not in the reference): block diagonal over the particles, each block the
        # BEMRet chain of one particle alone in its media; factored for all particles in one
        # batch. Cheaper to build (sum of n_p^3 instead of n^3), GMRES resolves the coupling.
        self.precond_mode = "particle" if pc == "particle" else "chain"
        pr = op.get("precond", None if self.hmatrix else "hmat")
        # with hmatrix=True the reference preconditions with the dense BEMRet chain unless
        # preconditioner='none' (bem_ret_iter.py:1688-1692, :1745-1774)
        self.use_precond = pc != "none" if self.hmatrix else not (pc == "none" or pr is None)
        self.precond_precision = op.get("precond_precision", "single")
        # complex64 preconditioner GEMMs: operands split into 2 bfloat16 terms by default (about
        # 16 significant bits: ample for an approximate inverse, GMRES judges every iterate);
        # 3 = as accurate as FP32, 0 = the FP32 matrix cores
        self.precond_gemm_split = int(op.get("precond_gemm_split", 2))
        # None: the preconditioner is the BEMRet chain of each solve's wavelength (reference);
        # a wavelength: the chain is factored there once and reused for every later solve
        self.precond_wavelength = op.get("precond_wavelength", None)
        # a preconditioner reused across solves is inverted explicitly once, so that each GMRES
        # iteration applies it by matrix-vector products instead of latency-bound substitutions
        self.precond_inverse = bool(op.get("precond_inverse", self.precond_wavelength is not None))
        self.g = op.get("green") or CompGreenRet(p, p, **{k: v for k, v in op.items()
                                                           if k in ("RelCutoff", "AbsCutoff", "order")})
        p._mnp_geo = self.g.geo
        if self.precond_mode == "particle" and self.use_precond:
            self._particle_blocks()       # wavelength-independent set-up of the blocks
        self.enei = None
        self.timings = {}     # seconds of the last H-matrix build, preconditioner build and GMRES run
        self._flag = self._relres = self._iter = None
        self._hist = None
        if enei is not None:
            self._init_matrices(enei)

    def _init_matrices(self, enei):
        if self.enei is not None and np.isclose(self.enei, enei):
            return
        geo = self.g.geo
        self._kmat, self._L1, self._L2, self._kvac, _, _ = geo.media([enei])
        self._exact = self._hm = None
        if self.hmatrix:
            from .hmatrix import ClusterTree, HMatrixRet
            if getattr(self, "_tree", None) is None:
                self._tree = ClusterTree(self.p.pos, self.cleaf, self.p.face_particle)
            self._hm = None
            torch.cuda.empty_cache()
            t0 = _now(geo.dev)
            self._hm = HMatrixRet(geo, self._tree, enei, self.htol, self.kmax, self.eta, self.fadmiss)
            self.timings["hmatrix"] = _now(geo.dev) - t0
            self._opwork = torch.empty((64, geo.n), dtype=torch.complex128, device=geo.dev)
        else:
            self._exact = engine.GreenBatch(geo, 1, torch.complex128)
            self._exact.assemble(self._kmat, 1)
            self._opwork = torch.empty((1, 32, geo.n), dtype=torch.complex128, device=geo.dev)
        keep = (self.precond_wavelength is not None and getattr(self, "_pre", None) is not None
                and np.isclose(self._pre_enei, self.precond_wavelength))
        self.timings["precond"] = 0.0
        if not keep:
            self._pre = None
        if self.use_precond and not keep:
            torch.cuda.empty_cache()
            # complex64 (default) or complex128 chain; refinement is done by GMRES itself
            single = self.precond_precision == "single"
            tuning = dict(gemm_split=self.precond_gemm_split if single else 0)
            # one large matrix: wider outer blocks give the trailing GEMMs a deeper K (2-term GEMMs
            # at n = 24,544: outer block 128 -> 4.5 s, 256 -> 3.7 s, 512 -> 3.5 s for the chain)
            if geo.n >= 12000:
                tuning["nbo"] = 512
                # 64 x 128 plane tiles with direct LDS copies: 6 % faster than 64 x 64 on the
                # 24k-row trailing updates (2-term, K 512)
                if single and tuning["gemm_split"] in (2, 3):
                    tuning["plane_tile"] = 12
            elif geo.n >= 4000:
                tuning["nbo"] = 256
            self._pre_enei = enei if self.precond_wavelength is None else float(self.precond_wavelength)
            t0 = _now(geo.dev)
            if self.precond_mode == "particle":
                self._pre = self._particle_preconditioner(single, tuning)
            else:
                self._pre = engine.RetSolverBatch(geo, 1, "single" if single else "double", tuning=tuning)
            self._pre.factor([self._pre_enei])
            if self.precond_inverse:
                self._pre.invert()
            self.timings["precond"] = _now(geo.dev) - t0
            self.timings["precond_build"] = self.timings["precond"]     # kept while it is reused
        self.enei = enei

    def _apply_A(self, x, y, rhs=None):
        e, geo, n = self._exact, self.g.geo, self.g.geo.n
        if self._hm is not None:
            call("mnp_bem_operator_hmat", self._hm.desc_ptr, geo.nvec.data_ptr(), self._L1.data_ptr(),
                 self._L2.data_ptr(), self._kvac.data_ptr(), x.data_ptr(),
                 rhs.data_ptr() if rhs is not None else None, y.data_ptr(), self._opwork.data_ptr(), 64 * n, None,
                 engine.stream_handle())
            return
        call("mnp_bem_operator", n, 1, e.G1.data_ptr(), e.G2.data_ptr(), e.H1.data_ptr(), e.H2.data_ptr(), e.ld,
             e.stride, geo.nvec.data_ptr(), 0, self._L1.data_ptr(), self._L2.data_ptr(), self._kvac.data_ptr(),
             x.data_ptr(), 8 * n, rhs.data_ptr() if rhs is not None else None, 8 * n, y.data_ptr(), 8 * n,
             self._opwork.data_ptr(), 32 * n, None, engine.stream_handle())

    def _particle_blocks(self):
        """Host and device geometry of each particle of p alone in its media (near-field pairs,
        closed-surface terms), and the gather indices between the packed 8 N vector and the
        padded (P, 8, n) batch entries. Wavelength independent: built once."""
        from .comparticle import ComParticle
        cps = []
        for i, part in enumerate(self.p.p):
            cp = ComParticle(self.p.eps, [part], [self.p.inout[i]], 1)
            cp.pc.quad_rule = self.p.pc.quad_rule
            cps.append(cp)
        greens = [CompGreenRet(cp, cp) for cp in cps]
        for g in greens:
            g.geo
        off = self.p._offsets
        n = max(cp.nfaces for cp in cps)
        idx = np.full((len(greens), n), -1, np.int64)
        for i in range(len(greens)):
            idx[i, : off[i + 1] - off[i]] = np.arange(off[i], off[i + 1])
        self._pidx = torch.tensor(idx, device=self.g.geo.dev)
        self._pmask = self._pidx >= 0
        self._blocks = (greens, n)

    def _particle_preconditioner(self, single, tuning):
        """GroupSolver over the particles of p, each alone in its own media (one batch entry per
        particle, padded to the largest)."""
        from .batch import GroupGeometry, GroupSolver
        geo = self.g.geo
        greens, n = self._blocks
        tuning = dict(tuning)
        tuning.pop("plane_tile", None)
        if n < 12000:
            tuning.pop("nbo", None)
            if n >= 4000:
                tuning["nbo"] = 256
        gg = GroupGeometry([g.geo for g in greens], 1, n, geo.dev)
        return GroupSolver(gg, "single" if single else "double", tuning=tuning)

    def _apply_M(self, v, z):
        if self._pre is None:
            z.copy_(v)
            return
        n = self.g.geo.n
        if self.precond_mode == "particle":
            V = v.view(8, n)
            P, m = self._pidx.shape
            rhs = torch.zeros((P, 8, m), dtype=v.dtype, device=v.device)
            sel = self._pidx.clamp(min=0)
            rhs[:] = V[:, sel].permute(1, 0, 2)
            rhs.mul_(self._pmask[:, None, :])
            x = self._pre.solve(rhs)
            Z = z.view(8, n)
            Z[:, self._pidx[self._pmask]] = x.permute(1, 0, 2)[:, self._pmask]
            return
        z.copy_(self._pre.solve(v.view(1, 8, n)).view(-1))

    def _best_start(self, b, guesses):
        """Residual-minimising combination x0 = sum_i c_i g_i of candidate guesses: one operator
        application per guess, then the small least-squares problem min |b - sum c_i A g_i|
        on the host. A useless guess gets a coefficient near zero, so the start is never worse
        than zero (|b - A x0| <= |b|)."""
        AG = []
        for gv in guesses:
            y = torch.empty_like(b)
            self._apply_A(gv, y, rhs=None)
            AG.append(y)
        Am = torch.stack(AG, dim=1)                       # (L, k)
        M_ = (Am.conj().T @ Am).cpu().numpy()
        r = (Am.conj().T @ b).cpu().numpy()
        c = np.linalg.lstsq(M_, r, rcond=1e-12)[0]
        x0 = torch.zeros_like(b)
        for ci, gv in zip(c, guesses):
            x0.add_(gv, alpha=complex(ci))
        self.start_coefficients = c
        return x0

    def solve(self, exc, x0=None):
        """Solve for excitation ``exc``; ``x0`` (optional) = packed initial guess or a
        CompStruct with sig1, h1, sig2, h2, or a list of such guesses (then the start is their
        residual-minimising combination, see ``_best_start``). Returns (sig, self)."""
        enei = exc.enei
        self._init_matrices(enei)
        p, n = self.p, self.p.nfaces
        k0 = 2 * np.pi / enei
        eps1, eps2 = p.eps1(enei), p.eps2(enei)
        npol = _npol(exc)
        L = 8 * n
        restart = int(self.restart) if self.restart is not None else min(L, self.maxit)
        restart = max(1, min(restart, 128))
        gm = make_solver(self.solver, L, restart, self.g.geo.dev)
        out = np.zeros((npol, 8, n), complex)
        infos = []
        for ip in range(npol):
            b = torch.tensor(_pack_rhs(p, exc, k0, eps1, eps2, ip).ravel(), device=self.g.geo.dev)
            xg = None
            t0 = _now(self.g.geo.dev)
            if isinstance(x0, (list, tuple)):
                cands = [torch.tensor(np.asarray(_packed(g_, ip, n)).ravel(), dtype=torch.complex128,
                                      device=self.g.geo.dev) for g_ in x0 if g_ is not None]
                xg = self._best_start(b, cands) if cands else None
            elif x0 is not None:
                xg = torch.tensor(np.asarray(_packed(x0, ip, n)).ravel(), dtype=torch.complex128,
                                  device=self.g.geo.dev)
            x, flag, relres, it, hist = gm.solve(self._apply_A, self._apply_M, b, xg, self.tol, self.maxit)
            self.timings["gmres"] = _now(self.g.geo.dev) - t0
            out[ip] = x.view(8, n).cpu().numpy()
            infos.append((flag, relres, it, hist))
        self._flag, self._relres, self._iter, self._hist = infos[-1]
        o = out[0] if npol == 1 else None
        if npol == 1:
            sig = CompStruct(p, enei, sig1=o[0], h1=o[1:4].T.copy(), sig2=o[4], h2=o[5:8].T.copy())
        else:
            sig = CompStruct(p, enei, sig1=out[:, 0].T.copy(), h1=np.transpose(out[:, 1:4], (2, 1, 0)).copy(),
                             sig2=out[:, 4].T.copy(), h2=np.transpose(out[:, 5:8], (2, 1, 0)).copy())
        return sig, self

    STARTS = ("zero", "previous", "extrapolate", "learned", "extrapolate+learned")

    def solve_spectrum(self, exc, enei_list, start="extrapolate", history=3, warm_start=None):
        """Plane-wave sweep over ``enei_list``: one solve per wavelength, each GMRES started from

        * 'zero': the reference's start (bem_ret_iter.py:1695);
        * 'previous': the previous wavelength's solution;
        * 'extrapolate': the residual-minimising combination of the last ``history`` solutions
          (the start is never worse than zero, see ``_best_start``);
        * 'learned', 'extrapolate+learned': the guess ``warm_start(enei)`` (a packed 8 n vector
          or CompStruct, e.g. from the warm-start network) alone, or as one more candidate of
          the combination.

        ``start`` may be a tuple of these: every wavelength's matrices (H-matrices and the
        operator) are then built once and solved from each start in turn, each start keeping
        its own previous solutions. Returns, per start (a dict of them for a tuple), a dict of
        per-wavelength lists: 'ext' (nlam, npol), 'iterations', 'relres', 'flag',
        'initial_residual' (|b - A x0| / |b|), 'time_s' (GMRES including the start's operator
        applications, plus the guess's own construction)."""
        starts = (start,) if isinstance(start, str) else tuple(start)
        for st in starts:
            if st not in self.STARTS:
                raise ValueError("start must be one of {}, got {!r}".format(self.STARTS, st))
            if "learned" in st and warm_start is None:
                raise ValueError("start '{}' needs warm_start".format(st))
        import time
        keep = max(1, int(history))
        res = {st: {k: [] for k in ("ext", "iterations", "relres", "flag", "initial_residual", "time_s")}
               for st in starts}
        prev = {st: [] for st in starts}
        for lam in np.atleast_1d(np.asarray(enei_list, float)):
            e = exc(self.p, lam)
            guess = None
            t_guess = 0.0
            if any("learned" in st for st in starts):
                t0 = time.perf_counter()
                guess = warm_start(lam)
                t_guess = time.perf_counter() - t0
            for st in starts:
                cands = []
                if st == "previous":
                    cands = prev[st][-1:]
                elif st.startswith("extrapolate"):
                    cands = prev[st][-keep:]
                if "learned" in st:
                    cands = cands + [guess]
                x0 = None
                if cands:
                    x0 = cands[0] if st == "previous" else cands
                sig, _ = self.solve(e, x0=x0)
                flag, relres, niter = self.info()
                r = res[st]
                r["ext"].append(np.atleast_1d(exc.extinction(sig)))
                r["iterations"].append(int(niter))
                r["relres"].append(float(relres))
                r["flag"].append(int(flag))
                r["initial_residual"].append(float(self.history[0]))
                r["time_s"].append(self.timings["gmres"] + (t_guess if "learned" in st else 0.0))
                prev[st].append(sig)
                del prev[st][:-keep]
        for r in res.values():
            r["ext"] = np.array(r["ext"])
        return res[start] if isinstance(start, str) else res

    def __truediv__(self, exc):
        return self.solve(exc)[0]

    def potential(self, sig, inout=2):
        """Potentials and surface derivatives of solution ``sig`` outside (inout=2, default) or
        inside (1) of the particle surface (bem_ret.py:1833-1859, via CompGreenRet.potential)."""
        return self.g.potential(sig, inout)

    def field(self, sig, inout=2):
        """Electric and magnetic fields of solution ``sig`` outside (inout=2, default) or
        inside (1) of the particle surface (bem_ret.py:1861-1888, via CompGreenRet.field)."""
        return self.g.field(sig, inout)

    def info(self):
        """(flag, relres, iter) of the last solve (bem_iter.py:191-194)."""
        return self._flag, self._relres, self._iter

    @property
    def history(self):
        return self._hist


def _now(dev):
    import time
    torch.cuda.synchronize(dev)
    return time.perf_counter()


def _packed(x0, ipol, n):
    if isinstance(x0, CompStruct):
        def col(v, c=None):
            v = np.asarray(v)
            if c is None:
                return v[:, ipol] if v.ndim == 2 else v
            return v[:, c, ipol] if v.ndim == 3 else v[:, c]
        return np.concatenate([col(x0.sig1)] + [col(x0.h1, c) for c in range(3)] + [col(x0.sig2)] +
                              [col(x0.h2, c) for c in range(3)])
    x = np.asarray(x0)
    return x.reshape(-1, 8 * n)[ipol] if x.size > 8 * n else x.ravel()
