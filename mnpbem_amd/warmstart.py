"""Learned GMRES warm start for BEMRetIter (no counterpart in the reference, whose iterative
solver starts from zero, bem_ret_iter.py:1695).

A per-face MLP 12 -> 128 -> 128 -> 128 -> 4 (GELU) maps local geometry, eps(lambda) and the
excitation phase to the surface charge sigma and the inside current along the polarisation,
h1.pol, from which the packed initial guess [sigma1, h1, sigma2, h2] = [s, (h1.pol) pol, s, 0]
is built (sigma1 and sigma2 agree to a few per cent on metal particles, h1 lies along the
polarisation, and h1 is what carries the residual of the 8N system). Forward and backward run
on the BF16 matrix cores with FP32 accumulation (``mnp_mlp_gemm``), training uses Adam
(``mnp_mlp_adam``) on FP32 master weights. GMRES judges every result: the guess only changes
where the iteration starts.

Features per face: centroid relative to its particle's centre (3), unit normal (3), the
particle's half-length along the polarisation, local mean curvature, Re/Im of the inside
eps(lambda), cos/sin of the excitation phase k (dir . r) at the centroid. In a solve the
learned guess is one candidate of BEMRetIter's residual-minimising start (with the previous
wavelength's solution when there is one), so a poor prediction costs one operator
application and never a worse start than zero.
Targets: Re/Im sigma2, Re/Im (h1 . pol).
"""
import numpy as np
import torch
from scipy.spatial import cKDTree

from . import engine
from ._lib import call

DIMS = (12, 128, 128, 128, 4)
NPARAM = sum(a * b + b for a, b in zip(DIMS[:-1], DIMS[1:]))   # 35,204


def mean_curvature(cp, k=7):
    """Per-face mean curvature estimate from the k nearest centroids of the same particle:
    H_i = mean_j (n_i - n_j).(x_i - x_j) / |x_i - x_j|^2 (1/R on a sphere of radius R)."""
    pos, nvec = cp.pos, cp.nvec
    fp = cp.face_particle
    H = np.zeros(len(pos))
    for ip in np.unique(fp):
        idx = np.nonzero(fp == ip)[0]
        kk = min(k, len(idx))
        _, nb = cKDTree(pos[idx]).query(pos[idx], kk)
        nb = idx[nb[:, 1:]]
        dx = pos[idx, None, :] - pos[nb]
        dn = nvec[idx, None, :] - nvec[nb]
        H[idx] = np.mean(np.sum(dn * dx, axis=2) / np.maximum(np.sum(dx * dx, axis=2), 1e-30), axis=1)
    return H


def features(cp, enei, pol, direc, medium=1, curvature=None):
    """(n, 12) float64 feature matrix of ComParticle ``cp`` at wavelength ``enei``: position
    relative to the face's particle centre (3), unit normal (3), the particle's half-length
    along the polarisation, local mean curvature, Re/Im inside eps, cos/sin of the excitation
    phase. Mesh-density independent (no face area) and independent of where the particle sits,
    except through the phase."""
    pos, nvec = cp.pos, cp.nvec
    H = mean_curvature(cp) if curvature is None else curvature
    p = np.asarray(pol, float)
    p = p / np.linalg.norm(p)
    fp = cp.face_particle
    rel = np.empty_like(pos)
    half = np.empty(len(pos))
    for ip in np.unique(fp):
        idx = fp == ip
        c = pos[idx].mean(axis=0)
        rel[idx] = pos[idx] - c
        s = rel[idx] @ p
        half[idx] = 0.5 * (s.max() - s.min())
    eps_in = np.asarray(cp.eps1(enei))
    eps_med, k_med = cp.eps[medium - 1](enei)
    phase = np.real(k_med) * (pos @ np.asarray(direc, float))
    return np.column_stack([rel, nvec, half, H, eps_in.real, eps_in.imag, np.cos(phase), np.sin(phase)])


def targets(cp, sig, pol):
    """(n, 4): Re/Im sigma2, Re/Im (h1 . pol) of a solution CompStruct."""
    p = np.asarray(pol, float)
    p = p / np.linalg.norm(p)
    hp = np.asarray(sig.h1) @ p
    s2 = np.asarray(sig.sig2)
    return np.column_stack([s2.real, s2.imag, hp.real, hp.imag])


def guess(cp, pred, pol):
    """Packed complex initial guess [sigma1, h1, sigma2, h2] (8 n) from MLP outputs (n, 4)."""
    n = cp.nfaces
    p = np.asarray(pol, float)
    p = p / np.linalg.norm(p)
    s = pred[:, 0] + 1j * pred[:, 1]
    hp = pred[:, 2] + 1j * pred[:, 3]
    x = np.zeros((8, n), complex)
    x[0] = s
    x[1:4] = (hp[:, None] * p[None, :]).T
    x[4] = s
    return x.ravel()


class WarmStartMLP:
    """The network, its normalisation and its optimiser state on the device."""

    def __init__(self, dev=None, seed=0):
        self.dev = dev or engine.device()
        rng = np.random.default_rng(seed)
        parts = []
        for a, b in zip(DIMS[:-1], DIMS[1:]):
            parts.append(rng.standard_normal(a * b) * np.sqrt(2.0 / a))
            parts.append(np.zeros(b))
        self.params = torch.tensor(np.concatenate(parts), dtype=torch.float32, device=self.dev)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.x_mean = np.zeros(DIMS[0])
        self.x_std = np.ones(DIMS[0])
        self.y_mean = np.zeros(DIMS[-1])
        self.y_std = np.ones(DIMS[-1])
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.dev)

    def _views(self, t):
        out, o = [], 0
        for a, b in zip(DIMS[:-1], DIMS[1:]):
            W = t[o:o + a * b]
            o += a * b
            bb = t[o:o + b]
            o += b
            out.append((W, bb))
        return out

    # --------------------------------------------------------------- device passes
    def forward(self, X):
        """X: (N, 12) float32 device tensor (normalised). Returns Y (N, 4) and the cache."""
        N = X.shape[0]
        st = engine.stream_handle()
        h = X.contiguous()
        cache = []
        layers = self._views(self.params)
        for li, (W, b) in enumerate(layers):
            fo = DIMS[li + 1]
            Z = torch.empty((N, fo), dtype=torch.float32, device=self.dev)
            last = li == len(layers) - 1
            H = Z if last else torch.empty_like(Z)
            call("mnp_mlp_gemm", 0, 0, 3 if last else 2, N, fo, DIMS[li], h.data_ptr(), DIMS[li], W.data_ptr(), fo,
                 b.data_ptr(), Z.data_ptr(), H.data_ptr(), fo, st)
            cache.append((h, Z))
            h = H
        return h, cache

    def backward(self, cache, dY):
        """Gradients of the loss with output gradient dY (N, 4) into self.grad."""
        st = engine.stream_handle()
        N = dY.shape[0]
        layers = self._views(self.params)
        glayers = self._views(self.grad)
        d = dY.contiguous()
        for li in range(len(layers) - 1, -1, -1):
            h, Z = cache[li]
            W, _ = layers[li]
            gW, gb = glayers[li]
            fi, fo = DIMS[li], DIMS[li + 1]
            if li < len(layers) - 1:
                dz = torch.empty_like(Z)
                call("mnp_mlp_gelu_bwd", N * fo, d.data_ptr(), Z.data_ptr(), dz.data_ptr(), st)
            else:
                dz = d
            # dW = h^T dz (K = N), db = column sums of dz, d(h) = dz W^T
            call("mnp_mlp_gemm", 1, 0, 0, fi, fo, N, h.data_ptr(), fi, dz.data_ptr(), fo, None, gW.data_ptr(), None, fo,
                 st)
            call("mnp_mlp_colsum", N, fo, dz.data_ptr(), fo, gb.data_ptr(), st)
            if li > 0:
                dh = torch.empty((N, fi), dtype=torch.float32, device=self.dev)
                call("mnp_mlp_gemm", 0, 1, 0, N, fi, fo, dz.data_ptr(), fo, W.data_ptr(), fo, None, dh.data_ptr(),
                     None, fi, st)
                d = dh

    def loss_grad(self, X, T):
        """MSE loss and gradient (device), X (N, 12), T (N, 4) float32."""
        Y, cache = self.forward(X)
        dY = torch.empty_like(Y)
        call("mnp_mlp_mse", Y.numel(), Y.data_ptr(), T.data_ptr(), dY.data_ptr(), self._loss.data_ptr(),
             engine.stream_handle())
        self.backward(cache, dY)
        return self._loss

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        call("mnp_mlp_adam", NPARAM, self.params.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
             self.v.data_ptr(), lr, b1, b2, eps, self.step, engine.stream_handle())

    # --------------------------------------------------------------- training / use
    def fit(self, X, T, steps=2000, batch=8192, lr=2e-3, seed=0):
        """Normalises features and targets with their statistics, then runs Adam on random
        minibatches. X (S, 12), T (S, 4) numpy. Returns the loss history (every 50 steps)."""
        self.x_mean, self.x_std = X.mean(axis=0), X.std(axis=0) + 1e-12
        self.y_mean, self.y_std = T.mean(axis=0), T.std(axis=0) + 1e-30
        Xd = torch.tensor((X - self.x_mean) / self.x_std, dtype=torch.float32, device=self.dev)
        Td = torch.tensor((T - self.y_mean) / self.y_std, dtype=torch.float32, device=self.dev)
        g = torch.Generator(device=self.dev).manual_seed(seed)
        hist = []
        S = X.shape[0]
        for it in range(steps):
            idx = torch.randint(0, S, (min(batch, S),), device=self.dev, generator=g)
            loss = self.loss_grad(Xd[idx].contiguous(), Td[idx].contiguous())
            self.adam(lr * (0.5 * (1 + np.cos(np.pi * it / steps))) + 1e-5)
            if it % 50 == 0 or it == steps - 1:
                hist.append(float(loss.item()))
        return hist

    def predict(self, X):
        """X (n, 12) numpy features -> (n, 4) numpy outputs in physical units."""
        Xd = torch.tensor((X - self.x_mean) / self.x_std, dtype=torch.float32, device=self.dev)
        Y, _ = self.forward(Xd)
        return Y.double().cpu().numpy() * self.y_std + self.y_mean


def training_set(cps, lams, pol, direc, precision="mixed", batch=50):
    """Features and targets from direct solves of every ComParticle in ``cps`` over its
    wavelengths for one plane wave: ``lams`` is one array for all particles or one array per
    particle. Each particle's wavelengths run as device sweeps (BEMRet.solve_spectrum, ``batch``
    wavelengths per factorisation, solutions copied back). Returns (X (S, 12), T (S, 4))."""
    from .bem import BEMRet
    from .planewave import PlaneWaveRet
    exc = PlaneWaveRet(np.asarray(pol, float), np.asarray(direc, float))
    p = np.asarray(pol, float)
    p = p / np.linalg.norm(p)
    per = isinstance(lams, (list, tuple)) or np.ndim(lams) == 2
    Xs, Ts = [], []
    for i, cp in enumerate(cps):
        el = np.atleast_1d(np.asarray(lams[i] if per else lams, float))
        H = mean_curvature(cp)
        bem = BEMRet(cp, precision=precision, batch=min(batch, len(el)), autotune=False)
        sol = bem.solve_spectrum(exc, el, solutions=True)["solutions"]
        for j, lam in enumerate(el):
            x = sol[j, 0]
            hp = x[1:4].T @ p
            Xs.append(features(cp, lam, pol, direc, curvature=H))
            Ts.append(np.column_stack([x[4].real, x[4].imag, hp.real, hp.imag]))
        bem.clear()
        del bem
    return np.vstack(Xs), np.vstack(Ts)
