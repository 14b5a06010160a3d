"""Quasistatic direct BEM solver (reference bem/bem_stat.py:101-1165) and plane-wave
excitation (simulation/planewave_stat.py:21-383).

sigma = M^{-1} phi',  M = -(diag Lambda + F),  Lambda = 2 pi (eps1 + eps2) / (eps1 - eps2).
F is wavelength independent and assembled once on the GPU; each wavelength adds its
diagonal, factors and solves there. Cross sections use the induced dipole
sum_j a_j r_j sigma_j (host, as the reference).
"""
import numpy as np
import torch

from . import engine
from ._lib import call
from .greenfun import CompGreenStat, CompStruct


class BEMStat:
    def __init__(self, p, enei=None, **op):
        if p is None:
            raise ValueError("BEMStat: 'p' must be a ComParticle (or compatible particle object), got None.")
        if not (hasattr(p, "pos") and hasattr(p, "nvec") and hasattr(p, "eps")):
            raise TypeError("BEMStat: 'p' must expose ComParticle-like attributes (pos, nvec, eps); got {!r}."
                            .format(type(p).__name__))
        self.p = p
        self.g = CompGreenStat(p, p, **op)
        self.nb = int(op.get("nb", 32))
        self.enei = None
        self._fac = None
        if enei is not None:
            self(enei)

    @property
    def F(self):
        return self.g.F

    def _factor(self, enei_list):
        enei_list = np.atleast_1d(np.asarray(enei_list, float))
        geo = self.g.green.geo
        n, B = geo.n, len(enei_list)
        _, H1 = self.g.device_F()
        e1 = np.stack([self.p.eps1(l) for l in enei_list])
        e2 = np.stack([self.p.eps2(l) for l in enei_list])
        lam = 2 * np.pi * (e1 + e2) / (e1 - e2)
        # M = -(F + diag Lambda) = -(F + 2 pi I) + diag(2 pi - Lambda - shift)
        d = torch.tensor(np.ascontiguousarray(2 * np.pi - lam - self.g.diag_shift[None, :]), dtype=torch.complex128,
                         device=geo.dev)
        M = torch.empty((B, n, n), dtype=torch.complex128, device=geo.dev)
        sh = engine.stream_handle()
        call("mnp_scale_diag_batched", 0, n, B, -1.0, H1.data_ptr(), n, d.data_ptr(), M.data_ptr(), n, n * n, sh)
        ipiv = torch.empty((B, n), dtype=torch.int32, device=geo.dev)
        info = torch.empty(B, dtype=torch.int32, device=geo.dev)
        call("mnp_getrf_batched", 0, n, B, M.data_ptr(), n, n * n, ipiv.data_ptr(), n, info.data_ptr(), self.nb, sh)
        inf = info.cpu().numpy()
        if np.any(inf < 0):
            raise RuntimeError("BEMStat: LU panel synchronisation timed out (info {}); factors are invalid".format(inf))
        if np.any(inf > 0):
            engine.warn_singular("BEMStat", inf)
        pf = torch.empty((B, n), dtype=torch.int32, device=geo.dev)
        q = torch.empty_like(pf)
        call("mnp_perm_batched", n, B, ipiv.data_ptr(), n, pf.data_ptr(), q.data_ptr(), n, sh)
        self._fac = (enei_list, M, pf, info)
        return self._fac

    def __call__(self, enei):
        if self.enei is None or self.enei != enei:
            self._factor([enei])
            self.enei = enei
        return self

    def _solve_dev(self, phip, b=0):
        """phip (n, r) numpy -> sigma (n, r) with the factor of wavelength index b."""
        enei_list, M, pf, _ = self._fac
        geo = self.g.green.geo
        n = geo.n
        phip = np.asarray(phip, complex).reshape(n, -1)
        r = phip.shape[1]
        out = np.zeros((n, r), complex)
        for c0 in range(0, r, 8):
            cols = phip[:, c0:c0 + 8]
            rr = cols.shape[1]
            rhs = torch.tensor(np.ascontiguousarray(cols.T), device=geo.dev)
            x = torch.empty_like(rhs)
            call("mnp_getrs_batched", 0, n, rr, 1, M[b].data_ptr(), n, n * n, pf[b].data_ptr(), n, rhs.data_ptr(), n,
                 n * rr, x.data_ptr(), n, n * rr, engine.stream_handle())
            out[:, c0:c0 + rr] = x.cpu().numpy().T
        return out

    def solve(self, exc):
        """Surface charge for excitation ``exc`` (field 'phip'): returns (sig, self)."""
        self(exc.enei)
        phip = np.asarray(exc.phip)
        sig = self._solve_dev(phip)
        sig = sig[:, 0] if phip.ndim == 1 else sig
        return CompStruct(self.p, exc.enei, sig=sig), self

    __truediv__ = solve

    def potential(self, sig, inout=2):
        """phi = G sig, phip = H sig outside (inout=2, default) or inside (1) of the surface
        (bem_stat.py:936-960, via CompGreenStat.potential)."""
        return self.g.potential(sig, inout)

    def field(self, sig, inout=2):
        """e = -Hp sig outside (inout=2, default) or inside (1) of the surface (bem_stat.py:
        863-935, deriv='cart', via CompGreenStat.field)."""
        return self.g.field(sig, inout)

    def clear(self):
        self._fac = None
        self.enei = None
        return self


class PlaneWaveStat:
    """Quasistatic plane wave: phi' = -n . pol; dipole cross sections."""

    def __init__(self, pol, medium=1, **options):
        if pol is None:
            raise ValueError("PlaneWaveStat: 'pol' must be a polarization vector, got None.")
        pol = np.asarray(pol, float)
        self.pol = pol.reshape(1, 3) if pol.ndim == 1 else pol
        if self.pol.shape[1] != 3:
            raise ValueError("PlaneWaveStat: 'pol' must have shape (3,) or (npol, 3).")
        self.medium = options.get("medium", medium)

    def potential(self, p, enei):
        phip = -p.nvec @ self.pol.T
        if self.pol.shape[0] == 1:
            phip = phip[:, 0]
        return CompStruct(p, enei, phip=phip)

    __call__ = potential

    def _dip(self, sig):
        s = np.asarray(sig.sig)
        s = s[:, None] if s.ndim == 1 else s
        return (sig.p.area[:, None] * sig.p.pos).T @ s          # (3, npol)

    def absorption(self, sig):
        dip = self._dip(sig)
        _, k = sig.p.eps[self.medium - 1](sig.enei)
        a = np.real(4 * np.pi * k * np.imag(np.sum(self.pol[: dip.shape[1]] * dip.T, axis=1)))
        return float(a[0]) if a.size == 1 else a

    def scattering(self, sig):
        dip = self._dip(sig)
        _, k = sig.p.eps[self.medium - 1](sig.enei)
        s = np.real(8 * np.pi / 3 * np.real(k) ** 4 * np.sum(np.abs(dip) ** 2, axis=0))
        return float(s[0]) if s.size == 1 else s

    def extinction(self, sig):
        return self.scattering(sig) + self.absorption(sig)
