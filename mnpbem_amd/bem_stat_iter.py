"""Iterative quasistatic BEM solver (reference bem/bem_stat_iter.py:342-1156) and the
H-matrix quasistatic Green function (greenfun/aca_compgreen_stat.py:11-327).

The system is that of BEMStat, -(F + diag Lambda) sigma = phi', Lambda = 2 pi (eps1 + eps2) /
(eps1 - eps2) per face (bem_stat_iter.py:855-868), solved by the device Krylov
solvers of ``bem_iter`` (restarted GMRES with CGS2, BiCGSTAB or CGS, right-preconditioned). F is wavelength independent:
dense on the device (``CompGreenStat``), or, with ``hmatrix=True``, an H-matrix whose entries,
ACA and products run on the device (``hmatrix.HMatrixRet`` with the k = 0 kernels). The
preconditioner is the LU of the dense -(F + diag Lambda) of the solve's wavelength
(precond 'hmat' / 'full', bem_stat_iter.py:528-600; also built for ``hmatrix=True`` unless
``preconditioner='none'``, :973-999), in complex128 (default) or complex64.

Differences from the reference: ``solve(exc, x0=...)`` accepts an initial guess;
``info()`` reports the true relative residual recomputed after the last restart.
"""
import numpy as np
import torch

from . import engine
from ._lib import call
from .bem_iter import _now, make_solver
from .greenfun import CompGreenStat, CompStruct


class _HMatView:
    """One kernel of an H-matrix as the reference's HMatrix object: ``@`` (host numpy, (n,) or
    (n, r)), ``full()``, ``compression()``. ``shift`` adds a diagonal (per face)."""

    def __init__(self, hm, kind, shift=None):
        self.hm, self.kind, self.shift = hm, kind, shift

    @property
    def shape(self):
        return (self.hm.geo.n, self.hm.geo.n)

    def matvec_dev(self, x):
        """(n, r) complex128 device tensor -> H x (device), r <= 8."""
        y = self.hm.matvec(self.kind, x)
        if self.shift is not None:
            y = y + self.shift[:, None] * x
        return y

    def __matmul__(self, x):
        x = np.asarray(x, complex)
        xm = x.reshape(x.shape[0], -1)
        out = np.empty_like(xm)
        for c0 in range(0, xm.shape[1], 8):
            xd = torch.tensor(np.ascontiguousarray(xm[:, c0:c0 + 8]), device=self.hm.geo.dev)
            out[:, c0:c0 + 8] = self.matvec_dev(xd).cpu().numpy()
        return out.reshape(x.shape)

    def full(self):
        A = self.hm.full(self.kind)
        if self.shift is not None:
            A[np.diag_indices_from(A)] += self.shift.cpu().numpy()
        return A

    def compression(self):
        return self.hm.stats["compression"][self.kind]


class ACACompGreenStat:
    """Quasistatic G, F, H1, H2 of a ComParticle as H-matrices (aca_compgreen_stat.py:11-327):
    ``eval(*keys)`` returns objects with ``@``, ``full()`` and ``compression()``; options
    ``htol``, ``kmax``, ``cleaf``, ``eta``, ``fadmiss`` and the CompGreenStat options."""

    def __init__(self, p, htol=1e-6, kmax=100, cleaf=32, fadmiss=None, eta=2.5, **options):
        from .hmatrix import ClusterTree
        self.p = p
        self.g = CompGreenStat(p, p, **options)
        self.tree = ClusterTree(p.pos, cleaf, p.face_particle)
        self._htol, self._kmax, self._eta, self._fadmiss = float(htol), int(kmax), float(eta), fadmiss
        self._hm = {}
        self._mat_cache = None
        self._enei_cache = None

    def _hmat(self, kind):
        from .hmatrix import HMatrixRet
        if kind not in self._hm:
            self._hm[kind] = HMatrixRet(self.g.green.geo, self.tree, 1.0, self._htol, self._kmax, self._eta,
                                        self._fadmiss, static=True, kinds=(kind,))
        return self._hm[kind]

    def eval(self, *keys):
        out = []
        dev = self.g.green.geo.dev
        shift = torch.tensor(self.g.diag_shift, dtype=torch.complex128, device=dev)
        for key in keys:
            if key == "Gp":
                raise ValueError("[error] Gp not implemented for ACACompGreenStat")
            if key == "G":
                out.append(_HMatView(self._hmat(0), 0))
            elif key == "F":
                out.append(_HMatView(self._hmat(2), 2, shift - 2 * np.pi))
            elif key in ("H1", "H2"):
                out.append(_HMatView(self._hmat(2), 2, shift + (0.0 if key == "H1" else -4 * np.pi)))
            else:
                raise ValueError("[error] Unknown key <{}>".format(key))
        return out[0] if len(out) == 1 else tuple(out)

    def __getattr__(self, name):
        if name in ("G", "F", "H1", "H2"):
            return self.eval(name)
        raise AttributeError("'{}' object has no attribute '{}'".format(type(self).__name__, name))

    def full(self, key="G"):
        return self.eval(key).full()

    def compression(self, key="G"):
        return self.eval(key).compression()

    def potential(self, sig, inout=1):
        """phi = G sig, phi' = H sig inside (H1) or outside (H2) (aca_compgreen_stat.py)."""
        G, H = self.eval("G", "H1" if inout == 1 else "H2")
        s = np.asarray(sig.sig)
        t = "1" if inout == 1 else "2"
        return CompStruct(self.p, sig.enei, **{"phi" + t: G @ s, "phi" + t + "p": H @ s})

    def solve(self, exc):
        """sigma = -(diag Lambda + F)^{-1} phi' with the dense inverse (as the reference)."""
        if self._enei_cache is None or self._enei_cache != exc.enei:
            e1, e2 = self.p.eps1(exc.enei), self.p.eps2(exc.enei)
            lam = 2 * np.pi * (e1 + e2) / (e1 - e2)
            A = np.diag(lam) + self.eval("F").full()
            self._mat_cache = -np.linalg.inv(A)
            self._enei_cache = exc.enei
        return CompStruct(self.p, exc.enei, sig=self._mat_cache @ np.asarray(exc.phip))

    def __repr__(self):
        return "ACACompGreenStat(n={})".format(self.tree.n)


class BEMStatIter:
    """Iterative BEMStat. Options (reference names): ``solver='gmres'|'bicgstab'|'cgs'``, ``tol=1e-4``,
    ``maxit=200``, ``restart=None``, ``precond='hmat'|'full'|None``,
    ``preconditioner='auto'|'none'``, ``hmatrix``, ``htol=1e-6``, ``kmax=[4, 100]``,
    ``cleaf=200``, ``eta=2.5``, ``fadmiss``; not in the reference: ``precond_precision``
    ('double' default, or 'single')."""

    def __init__(self, p, enei=None, **op):
        if p is None:
            raise ValueError("BEMStatIter: 'p' must be a ComParticle, got None.")
        self.p = p
        self.hmatrix = bool(op.get("hmatrix", False))
        top = lambda v: max(v) if hasattr(v, "__iter__") else v
        self.htol = float(top(op.get("htol", 1e-6)))
        self.kmax = int(top(op.get("kmax", [4, 100])))
        self.cleaf = int(op.get("cleaf", 200))
        self.eta = float(op.get("eta", 2.5))
        self.fadmiss = op.get("fadmiss", None)
        self.solver = op.get("solver", "gmres")
        if self.solver not in ("gmres", "bicgstab", "cgs"):
            raise ValueError("[error] iterative solver not known: <{}>".format(self.solver))
        self.tol = float(op.get("tol", 1e-4))
        self.maxit = int(op.get("maxit", 200))
        self.restart = op.get("restart", None)
        pc = op.get("preconditioner", "auto")
        self.precond = op.get("precond", None if self.hmatrix else "hmat")
        if self.precond not in (None, "hmat", "full"):
            raise ValueError("[error] preconditioner not known: <{}>".format(self.precond))
        self.use_precond = pc != "none" if self.hmatrix else not (pc == "none" or self.precond is None)
        self.precond_precision = op.get("precond_precision", "double")
        gopt = {k: v for k, v in op.items() if k in ("RelCutoff", "AbsCutoff", "order", "green")}
        if self.hmatrix:
            self.g = ACACompGreenStat(p, self.htol, self.kmax, self.cleaf, self.fadmiss, self.eta, **gopt)
            self._gs = self.g.g
        else:
            self.g = CompGreenStat(p, p, **gopt)
            self._gs = self.g
        self.enei = None
        self.timings = {}
        self._flag, self._relres, self._iter, self._hist = [], [], [], None
        self._pre = None
        if enei is not None:
            self._init_matrices(enei)

    @property
    def F(self):
        return self.g.eval("F") if self.hmatrix else self.g.F

    def _init_matrices(self, enei):
        if self.enei is not None and self.enei == enei:
            return self
        geo = self._gs.green.geo
        n = geo.n
        e1, e2 = self.p.eps1(enei), self.p.eps2(enei)
        lam = 2 * np.pi * (e1 + e2) / (e1 - e2)
        # -(F + Lambda) x = -(F + 2 pi) x + d x,  d = 2 pi - Lambda - shift (F = F_dev + shift)
        d = 2 * np.pi - lam - self._gs.diag_shift
        self._d = torch.tensor(np.ascontiguousarray(d), dtype=torch.complex128, device=geo.dev)
        self._hm = None
        if self.hmatrix:
            t0 = _now(geo.dev)
            self._hm = self.g._hmat(2)
            self.timings["hmatrix"] = _now(geo.dev) - t0
        else:
            self._H1 = self._gs.device_F()[1]
        self._pre = None
        self.timings["precond"] = 0.0
        if self.use_precond:
            t0 = _now(geo.dev)
            H1 = self._gs.device_F()[1]
            prec = 1 if self.precond_precision == "single" else 0
            sh = engine.stream_handle()
            src = H1
            if prec:
                src = torch.empty((1, n, n), dtype=torch.complex64, device=geo.dev)
                call("mnp_cast_batched", 1, n * n, 1, H1.data_ptr(), n * n, src.data_ptr(), n * n, sh)
            M = torch.empty((1, n, n), dtype=src.dtype, device=geo.dev)
            dd = self._d.view(1, n).contiguous()
            call("mnp_scale_diag_batched", prec, n, 1, -1.0, src.data_ptr(), n, dd.data_ptr(), M.data_ptr(), n, n * n,
                 sh)
            del src
            ipiv = torch.empty((1, n), dtype=torch.int32, device=geo.dev)
            info = torch.empty(1, dtype=torch.int32, device=geo.dev)
            call("mnp_getrf_batched", prec, n, 1, M.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), 32, sh)
            pf = torch.empty((1, n), dtype=torch.int32, device=geo.dev)
            q = torch.empty_like(pf)
            call("mnp_perm_batched", n, 1, ipiv.data_ptr(), n, pf.data_ptr(), q.data_ptr(), n, sh)
            if int(info.item()) < 0:
                raise RuntimeError("BEMStatIter: preconditioner factorisation failed (info={})".format(int(info.item())))
            self._pre = (prec, M, pf)
            self.timings["precond"] = _now(geo.dev) - t0
        self.enei = enei
        return self

    def _apply_A(self, x, y, rhs=None):
        """y = -(F + Lambda) x, or y = rhs - (that) when rhs is given (device vectors of n)."""
        n = x.numel()
        if self._hm is not None:
            t = self._hm.matvec(2, x.view(n, 1)).view(n)
        else:
            t = torch.empty_like(x)
            call("mnp_gemv_batched", 0, n, 1, 1, self._H1.data_ptr(), n, n * n, x.data_ptr(), n, n, t.data_ptr(), n,
                 n, engine.stream_handle())
        v = self._d * x - t
        if rhs is not None:
            torch.sub(rhs, v, out=y)
        else:
            y.copy_(v)

    def _apply_M(self, v, z):
        if self._pre is None:
            z.copy_(v)
            return
        prec, M, pf = self._pre
        n = v.numel()
        src = v if prec == 0 else v.to(torch.complex64)
        out = torch.empty_like(src)
        call("mnp_getrs_batched", prec, n, 1, 1, M.data_ptr(), n, n * n, pf.data_ptr(), n, src.data_ptr(), n, n,
             out.data_ptr(), n, n, engine.stream_handle())
        z.copy_(out if prec == 0 else out.to(torch.complex128))

    def solve(self, exc, x0=None):
        """Surface charge for excitation ``exc`` (field 'phip'): returns (sig, self)."""
        enei = exc.enei
        self._init_matrices(enei)
        geo = self._gs.green.geo
        n = geo.n
        phip = np.asarray(exc.phip, complex)
        cols = phip.reshape(n, -1)
        restart = int(self.restart) if self.restart is not None else min(n, self.maxit)
        restart = max(1, min(restart, 128))
        gm = make_solver(self.solver, n, restart, geo.dev) if self.maxit > 0 else None
        out = np.zeros_like(cols)
        x0c = None if x0 is None else np.asarray(x0, complex).reshape(n, -1)
        t0 = _now(geo.dev)
        flags, res, its = [], [], []
        for ip in range(cols.shape[1]):
            b = torch.tensor(np.ascontiguousarray(cols[:, ip]), device=geo.dev)
            if self.maxit == 0:
                # preconditioner as the solver (bem_iter.py:51-55)
                if self._pre is None:
                    raise AssertionError("[error] preconditioner must be hmat when maxit=0")
                x = torch.empty_like(b)
                self._apply_M(b, x)
                flag, relres, it, hist = 0, 0.0, 0, []
            else:
                xg = None if x0c is None else torch.tensor(np.ascontiguousarray(x0c[:, ip]), device=geo.dev)
                x, flag, relres, it, hist = gm.solve(self._apply_A, self._apply_M, b, xg, self.tol, self.maxit)
            out[:, ip] = x.cpu().numpy()
            flags.append(flag)
            res.append(relres)
            its.append(it)
            self._hist = hist
        self.timings["gmres"] = _now(geo.dev) - t0
        # one record per solve, as the reference's _set_iter (bem_iter.py:103-114)
        self._flag.append(int(max(flags)))
        self._relres.append(float(max(res)))
        self._iter.append(np.array([int(max(its)), 0]))
        return CompStruct(self.p, enei, sig=out.reshape(phip.shape)), self

    def __truediv__(self, exc):
        return self.solve(exc)[0]

    def __call__(self, enei):
        return self._init_matrices(enei)

    def info(self):
        """(flag, relres, iter) lists with one entry per solve so far (bem_iter.py:191-194);
        relres is the true relative residual |b - A x| / |b|."""
        return self._flag, self._relres, self._iter

    def potential(self, sig, inout=2):
        """Quasistatic surface potentials phi = G sig, phi' = H sig (bem_stat_iter.py:1073-1085)."""
        if self.hmatrix:
            return self.g.potential(sig, inout)
        s = np.asarray(sig.sig)
        H = self.g.F + (2 * np.pi if inout == 1 else -2 * np.pi) * np.eye(len(s))
        t = "1" if inout == 1 else "2"
        return CompStruct(self.p, sig.enei, **{"phi" + t: self.g.G @ s, "phi" + t + "p": H @ s})

    def field(self, sig, inout=2):
        """e = -Hp sig on the surface (bem_stat_iter.py:1019-1072, via CompGreenStat.field)."""
        return self._gs.field(sig, inout)

    def clear(self):
        self._pre = None
        self.enei = None
        return self
