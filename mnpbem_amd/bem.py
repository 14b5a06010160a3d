"""Retarded direct BEM solver (reference bem/bem_ret.py:138-1986).

``BEMRet(p).solve(exc)`` keeps the reference's call: ``exc`` is the CompStruct of
``PlaneWaveRet.potential`` (or any CompStruct with phi1/phi2/a1/a2 and derivatives), the
result a CompStruct with sig1, sig2, h1, h2 (host numpy). Factorisation and solves run on
the GPU (libmnpbem_hip.so); ``solve_spectrum`` keeps everything on the device for a whole
wavelength sweep, batching wavelengths.

Options: ``precision`` = 'double' (default, complex128 LU chain as the reference),
'mixed' (complex64 LU chain + FP64 iterative refinement on the exact operator) or
'single'; ``batch`` = wavelengths factored together; ``nb`` = LU panel width;
``refine_tol``, ``max_refine`` for the refinement; ``autotune`` = pick the LU shape (panel
width, outer block width, lookahead) by timing (default: on when batch >= 8); ``lu_tuning`` =
a fixed LU shape instead (dict with keys of ``autotune.knobs``).
"""
import numpy as np
import torch

from . import engine
from .greenfun import CompGreenRet, CompStruct


def _field(exc, name):
    v = exc.get(name, 0) if hasattr(exc, "get") else getattr(exc, name, 0)
    return v


def _pack_rhs(p, exc, k0, eps1, eps2, ipol):
    """[phi a De alpha] (8, n) of Eqs. (10,11), (15), (18) (bem_ret.py:1360-1429)."""
    n = p.nfaces
    nv = p.nvec

    def get(name, vec):
        v = _field(exc, name)
        if isinstance(v, np.ndarray):
            if vec:
                return v[:, :, ipol] if v.ndim == 3 else v
            return v[:, ipol] if v.ndim == 2 else v
        return np.zeros((n, 3) if vec else n, complex)

    phi1, phi2 = get("phi1", False), get("phi2", False)
    phi1p, phi2p = get("phi1p", False), get("phi2p", False)
    a1, a2 = get("a1", True), get("a2", True)
    a1p, a2p = get("a1p", True), get("a2p", True)
    phi = phi2 - phi1
    a = a2 - a1
    alpha = a2p - a1p - 1j * k0 * (nv * (eps2 * phi2)[:, None] - nv * (eps1 * phi1)[:, None])
    De = eps2 * phi2p - eps1 * phi1p - 1j * k0 * (eps2 * np.sum(nv * a2, axis=1) - eps1 * np.sum(nv * a1, axis=1))
    return np.vstack([phi[None], a.T, De[None], alpha.T])


def _npol(exc):
    for name, nd in (("a1", 3), ("a2", 3), ("a1p", 3), ("a2p", 3), ("phi1", 2), ("phi2", 2)):
        v = _field(exc, name)
        if isinstance(v, np.ndarray) and v.ndim == nd:
            return v.shape[-1]
    return 1


class BEMRet:
    def __init__(self, p, enei=None, **op):
        if p is None:
            raise ValueError("BEMRet: 'p' must be a ComParticle, got None.")
        if not hasattr(p, "eps") or not hasattr(p, "inout"):
            raise TypeError("BEMRet: 'p' must be a ComParticle.")
        self.p = p
        self.op = op
        self.precision = op.get("precision", "double")
        self.batch = int(op.get("batch", 1))
        self.nb = int(op.get("nb", 32))
        self.refine_tol = float(op.get("refine_tol", 1e-12))
        self.max_refine = int(op.get("max_refine", 12))
        # LU shape autotuning (autotune.py): default on for batched sweeps (batch >= 8)
        self.autotune = op.get("autotune", None)
        self.lu_tuning = op.get("lu_tuning", None)
        # complex-float GEMMs of the mixed chain: 16 (default) operands as scaled fp16 pairs on the
        # FP16 matrix cores (3 products per real product; measured as accurate as the FP32 matrix
        # cores), 3 bfloat16 terms (6 products, as accurate as FP32), 2 bfloat16 terms (about 16
        # bits; the FP64 refinement takes the accuracy back), or 0 (the FP32 matrix cores)
        self.gemm_split = int(op.get("gemm_split", 16))
        if self.gemm_split not in (0, 2, 3, 16):
            raise ValueError("BEMRet: gemm_split must be 0, 2, 3 or 16")
        self.tuning = None
        self.g = op.get("green") or CompGreenRet(p, p, **{k: v for k, v in op.items()
                                                           if k in ("RelCutoff", "AbsCutoff", "order")})
        p._mnp_geo = self.g.geo
        self.enei = None
        self._solver = None
        self.nvec = p.nvec
        if enei is not None:
            self.init(enei)

    # ------------------------------------------------------------------ per wavelength API
    def _get_solver(self, batch, precision):
        s = self._solver
        if s is None or s.batch != batch or s.precision != precision:
            # the previous solver's buffers go back to torch's caching allocator and are reused
            # (no empty_cache: re-mapping tens of GB per solver costs more than the solves)
            self._solver = None
            nb = self.nb
            split = self.gemm_split if precision != "double" else 0
            tuning = dict(gemm_split=split)
            if self.lu_tuning is not None:
                # an explicit LU shape (keys of autotune.knobs: nb, nbo, lookahead, lu_recursive,
                # plane_tile) instead of the autotuner's timing
                tuning = dict(self.lu_tuning, gemm_split=split)
                if split != 16 and split not in (2, 3):
                    tuning.pop("plane_tile", None)
                nb = int(tuning.get("nb", nb))
            elif (self.autotune if self.autotune is not None else batch >= 8):
                from . import autotune
                fdt = torch.complex128 if precision == "double" else torch.complex64
                tuning = autotune.tune_lu(self.g.geo.n, batch, fdt, self.g.geo.dev, gemm_split=split)
                nb = tuning["nb"]
            self.tuning = tuning
            s = engine.RetSolverBatch(self.g.geo, batch, precision, nb, self.refine_tol, self.max_refine, tuning)
            s.tune_assembly = bool(self.autotune if self.autotune is not None else batch >= 8)
            self._solver = s
        return s

    def init(self, enei):
        """Factor the BEM system at one wavelength (skipped if already factored there)."""
        if self.enei is not None and np.isclose(self.enei, enei):
            return self
        self._get_solver(1, self.precision).factor([enei])
        self.enei = enei
        self.k = 2 * np.pi / enei
        return self

    def solve(self, exc):
        """Surface charges and currents for excitation ``exc``: returns (sig, self)."""
        enei = exc.enei if hasattr(exc, "enei") else exc["enei"]
        self.init(enei)
        s = self._solver
        p = self.p
        k0 = 2 * np.pi / enei
        eps1, eps2 = p.eps1(enei), p.eps2(enei)
        npol = _npol(exc)
        n = p.nfaces
        out = np.zeros((npol, 8, n), complex)
        for ip in range(npol):
            # explicit (8 n, n, 1) strides: a numpy view with a new leading axis would carry stride 0
            rhs = torch.tensor(_pack_rhs(p, exc, k0, eps1, eps2, ip), device=s.geo.dev).reshape(1, 8, n).contiguous()
            out[ip] = s.solve(rhs)[0].cpu().numpy()
        if npol == 1:
            o = out[0]
            sig = CompStruct(p, enei, sig1=o[0], h1=o[1:4].T.copy(), sig2=o[4], h2=o[5:8].T.copy())
        else:
            sig = CompStruct(p, enei, sig1=out[:, 0].T.copy(), h1=np.transpose(out[:, 1:4], (2, 1, 0)).copy(),
                             sig2=out[:, 4].T.copy(), h2=np.transpose(out[:, 5:8], (2, 1, 0)).copy())
        return sig, self

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

    def __call__(self, enei):
        return self.init(enei)

    def clear(self):
        """Drops the factors and returns their memory to the device."""
        self._solver = None
        self.enei = None
        torch.cuda.empty_cache()
        return self

    # ------------------------------------------------------------------ device sweep
    def solve_spectrum(self, exc, enei_list, scattering=False, batch=None, precision=None, timings=None,
                       solutions=False):
        """Extinction (and scattering) of a plane wave over ``enei_list`` with every stage
        on the device, ``batch`` wavelengths per factorisation. Returns a dict with 'ext'
        (nlam, npol), optionally 'sca', refinement statistics and, with ``solutions``, the
        surface charges and currents as 'solutions' (nlam, npol, 8, n) complex128 on the host,
        packed [sig1, h1 (x, y, z), sig2, h2 (x, y, z)] per face."""
        enei_list = np.atleast_1d(np.asarray(enei_list, float))
        batch = int(batch or self.batch)
        precision = precision or self.precision
        # the sweep refactors the solver object that init() may have factored: the factors no
        # longer belong to self.enei, so the next solve() factors again
        self.enei = None
        s = self._get_solver(batch, precision)
        geo = s.geo
        npol = exc.pol.shape[0]
        ext = np.zeros((len(enei_list), npol))
        sca = np.zeros((len(enei_list), npol)) if scattering else None
        refine_iters, relres, fallback, history = [], [], [], []
        sols = np.zeros((len(enei_list), npol, 8, geo.n), complex) if solutions else None
        sp = exc.spec
        for b0 in range(0, len(enei_list), batch):
            el = enei_list[b0:b0 + batch]
            nl = len(el)
            s.factor(el)
            kmed = s.kmat[:, exc.medium - 1].contiguous()
            kmed_np = s.kmat_np[:, exc.medium - 1]
            eps_med = s.eps_np[:, exc.medium - 1]
            for ip in range(npol):
                rhs = engine.planewave_rhs(geo, s.kvac, kmed, s.L1, s.L2, exc.pol[ip], exc.dir[ip], exc.medium, nl,
                                           batch)
                x = s.solve(rhs)
                if precision == "mixed":
                    refine_iters.append(s.stats["iterations"])
                    relres.append(s.stats["relres"])
                    history.append(s.stats["history"])
                    if len(s.stats["fallback"]):
                        fb = s.stats["fallback"]
                        fallback.extend(list(el[fb]))
                        x = self._redo_double(el, fb, rhs, x)
                if solutions:
                    sols[b0:b0 + nl, ip] = x[:nl].cpu().numpy()
                ff = engine.farfield(geo, x, s.kvac, kmed, exc.dir[ip:ip + 1], exc.medium, nl)
                e = ff[:, 0, :3].cpu().numpy()
                ext[b0:b0 + nl, ip] = np.real(4 * np.pi / kmed_np * np.imag(e @ np.conj(exc.pol[ip])))
                if scattering:
                    ffs = engine.farfield(geo, x, s.kvac, kmed, sp.nvec, exc.medium, nl).cpu().numpy()
                    e, h = ffs[:, :, :3], ffs[:, :, 3:]
                    dsca = 0.5 * np.real(np.sum(sp.nvec[None] * np.cross(e, np.conj(h)), axis=2))
                    nb = np.real(np.sqrt(eps_med))
                    sca[b0:b0 + nl, ip] = (dsca @ sp.area) / (0.5 * nb)
        out = dict(ext=ext, sca=sca)
        if solutions:
            out["solutions"] = sols
        if precision == "mixed":
            # relres before each refinement step and after the last: the largest over the wavelengths
            steps = max(len(h) for h in history)
            hmax = [float(max(np.max(h[i]) for h in history if len(h) > i)) for i in range(steps)]
            out.update(refine_iterations=refine_iters, relres=relres, fallback=fallback, refine_history=hmax)
        return out

    def _redo_double(self, el, idx, rhs, x):
        """Re-solve wavelengths el[idx] whose refinement stalled with the complex128 chain."""
        geo = self.g.geo
        sd = engine.RetSolverBatch(geo, len(idx), "double", self.nb)
        sd.factor(el[idx])
        sub = rhs[torch.as_tensor(idx, device=rhs.device)].contiguous()
        xd = sd.solve(sub)
        x = x.clone()
        x[torch.as_tensor(idx, device=x.device)] = xd[: len(idx)]
        return x
