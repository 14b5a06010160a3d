"""Plane-wave excitation and far-field observables (retarded).

Mirrors simulation/planewave_ret.py:81-504 and spectrum/spectrum_ret.py:136-459. The
potential is formed on the host (it is the API's input container); far fields are reduced
on the GPU by libmnpbem_hip.so.
"""
import os

import numpy as np
import torch

from . import engine
from .greenfun import CompStruct

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


class SpectrumRet:
    """Unit sphere of directions for far-field integration; default: the toolbox's
    508-direction trisphere(256, 2) set."""

    def __init__(self, pinfty=None, medium=1):
        self.medium = medium
        if pinfty is None:
            d = np.load(os.path.join(_DATA, "pinfty256.npz"))
            self.nvec, self.area = d["nvec"], d["area"]
        elif isinstance(pinfty, np.ndarray):
            self.nvec = np.atleast_2d(pinfty)
            self.area = np.full(len(self.nvec), 4 * np.pi / len(self.nvec))
        else:
            self.nvec, self.area = np.asarray(pinfty.nvec), np.asarray(pinfty.area)
        self.ndir = len(self.nvec)


def _as_rows(v, name):
    a = np.asarray(v, dtype=float)
    if a.dtype == object or a.size == 0:
        raise ValueError("PlaneWaveRet: '{}' must be a numeric array of shape (3,) or (npol, 3).".format(name))
    if a.ndim == 1:
        if a.size != 3:
            raise ValueError("PlaneWaveRet: 1D '{}' must have length 3, got length {}.".format(name, a.size))
        a = a.reshape(1, 3)
    elif a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("PlaneWaveRet: '{}' must have shape (npol, 3), got {}.".format(name, a.shape))
    return a


def _geo_for(p):
    g = getattr(p, "_mnp_geo", None)
    if g is None:
        g = engine.DeviceGeometry(p, None, engine.device())
        p._mnp_geo = g
    return g


def _pack_solution(sig, ipol, dev):
    n = sig.p.nfaces

    def col(v, c=None):
        v = np.asarray(v)
        if c is None:
            return v[:, ipol] if v.ndim == 2 else v
        return v[:, c, ipol] if v.ndim == 3 else v[:, c]

    cols = [col(sig.sig1)] + [col(sig.h1, c) for c in range(3)] + [col(sig.sig2)] + [col(sig.h2, c) for c in range(3)]
    x = np.stack(cols, axis=0).astype(complex).reshape(1, 8, n)
    return torch.tensor(x, device=dev)


class PlaneWaveRet:
    def __init__(self, pol, dir, medium=1, **options):
        if pol is None:
            raise ValueError("PlaneWaveRet: 'pol' must be a polarization vector, got None.")
        if dir is None:
            raise ValueError("PlaneWaveRet: 'dir' must be a propagation direction vector, got None.")
        self.pol = _as_rows(pol, "pol")
        self.dir = _as_rows(dir, "dir")
        self.medium = options.get("medium", medium)
        self.spec = SpectrumRet(options.get("pinfty"), medium=self.medium)

    def potential(self, p, enei):
        """CompStruct with a1, a1p, a2, a2p (planewave_ret.py:257-361); (n, 3) for one
        polarisation, (n, 3, npol) otherwise."""
        if not np.allclose(np.sum(self.pol * self.dir, axis=1), 0):
            raise ValueError("Polarization and propagation direction must be orthogonal")
        eps, _ = p.eps[self.medium - 1](enei)
        k0 = 2 * np.pi / enei
        k = k0 * np.sqrt(complex(np.asarray(eps).ravel()[0]))
        npol = self.pol.shape[0]
        io = p.inout_faces
        exc = CompStruct(p, enei)
        for side, (na, nap) in enumerate((("a1", "a1p"), ("a2", "a2p"))):
            a = np.zeros((p.nfaces, 3, npol), complex)
            ap = np.zeros((p.nfaces, 3, npol), complex)
            ind = np.nonzero(io[:, side] == self.medium)[0]
            for i in range(npol):
                ph = np.exp(1j * k * (p.pos[ind] @ self.dir[i])) / (1j * k0)
                a[ind, :, i] = ph[:, None] * self.pol[i]
                ap[ind, :, i] = (1j * k * (p.nvec[ind] @ self.dir[i]))[:, None] * ph[:, None] * self.pol[i]
            if npol == 1:
                a, ap = a[:, :, 0], ap[:, :, 0]
            exc.set(**{na: a, nap: ap})
        return exc

    __call__ = potential

    def field(self, p, enei, inout=1):
        """Incident fields e = pol e^{ik dir.r}, h = n_b dir x e (planewave_ret.py:157-240) on the
        faces of a ComParticle whose side ``inout`` (1 inside, 2 outside) is the excitation
        medium, or at the points of a ComPoint that lie in it; zero elsewhere. (n, 3) for one
        polarisation, (n, 3, npol) otherwise."""
        from .compoint import ComPoint
        if not np.allclose(np.sum(self.pol * self.dir, axis=1), 0):
            raise ValueError("Polarization and propagation direction must be orthogonal")
        eps, _ = p.eps[self.medium - 1](enei)
        nb = np.sqrt(complex(np.asarray(eps).ravel()[0]))
        k = 2 * np.pi / enei * nb
        if isinstance(p, ComPoint):
            n, ind = p.n, np.nonzero(p.medium == self.medium)[0]
        else:
            n, ind = p.nfaces, np.nonzero(p.inout_faces[:, inout - 1] == self.medium)[0]
        npol = self.pol.shape[0]
        e = np.zeros((n, 3, npol), complex)
        h = np.zeros((n, 3, npol), complex)
        for i in range(npol):
            e[ind, :, i] = np.exp(1j * k * (p.pos[ind] @ self.dir[i]))[:, None] * self.pol[i]
            h[:, :, i] = nb * np.cross(self.dir[i][None, :], e[:, :, i])
        if npol == 1:
            e, h = e[:, :, 0], h[:, :, 0]
        return CompStruct(p, enei, e=e, h=h)

    def _farfield(self, sig, dirs, ipol):
        p = sig.p
        geo = _geo_for(p)
        kmat, _, _, kvac, kmat_np, _ = geo.media([sig.enei])
        kmed = kmat[:, self.medium - 1].contiguous()
        x = _pack_solution(sig, ipol, geo.dev)
        ff = engine.farfield(geo, x, kvac, kmed, dirs, self.medium, 1)
        return ff[0].cpu().numpy(), kmat_np[0, self.medium - 1]

    def _npol_sig(self, sig):
        s1 = np.asarray(sig.sig1)
        return 1 if s1.ndim == 1 else s1.shape[1]

    def extinction(self, sig):
        """ext = 4 pi / k Im(conj(pol) . e(dir)) (planewave_ret.py:437-504)."""
        npol = self.pol.shape[0]
        ext = np.zeros(npol)
        for i in range(npol):
            ip = min(i, self._npol_sig(sig) - 1)
            ff, k = self._farfield(sig, self.dir[i:i + 1], ip)
            e = ff[0, :3]
            ext[i] = np.real(4 * np.pi / k * np.imag(np.sum(np.conj(self.pol[i]) * e)))
        return float(ext[0]) if npol == 1 else ext

    def scattering(self, sig):
        """(sca, dsca): Poynting flux through the direction sphere over 0.5 n_b
        (planewave_ret.py:391-435, spectrum_ret.py:395-459)."""
        nsig = self._npol_sig(sig)
        eps, _ = sig.p.eps[self.medium - 1](sig.enei)
        nb = np.real(np.sqrt(complex(np.asarray(eps).ravel()[0])))
        sca = np.zeros(nsig)
        dsca = np.zeros((self.spec.ndir, nsig))
        for ip in range(nsig):
            ff, _ = self._farfield(sig, self.spec.nvec, ip)
            e, h = ff[:, :3], ff[:, 3:]
            dsca[:, ip] = 0.5 * np.real(np.sum(self.spec.nvec * np.cross(e, np.conj(h)), axis=1))
            sca[ip] = np.dot(self.spec.area, dsca[:, ip])
        sca = sca / (0.5 * nb)
        dsca = dsca / (0.5 * nb)
        if nsig == 1:
            return float(sca[0]), CompStruct(None, sig.enei, dsca=dsca[:, 0])
        return sca, CompStruct(None, sig.enei, dsca=dsca)

    def absorption(self, sig):
        sca, _ = self.scattering(sig)
        return self.extinction(sig) - sca

    def __repr__(self):
        return "PlaneWaveRet(pol={}, dir={}, medium={})".format(self.pol.tolist(), self.dir.tolist(), self.medium)
