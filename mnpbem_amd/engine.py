"""Device-resident retarded BEM pipeline for batches of wavelengths.

Buffers are torch tensors on the GPU (torch provides memory and the stream); all arithmetic
is done by libmnpbem_hip.so. A matrix of the batch is stored column-major: a tensor of shape
(B, n, ld) holds matrix b with element (i, j) at [b, j, i]. Packed vectors of eight columns
(phi a De alpha, or sig1 h1 sig2 h2) are tensors of shape (B, 8, n).

Per wavelength batch:
  assemble G1 G2 H1 H2 -> factor (LU chain) -> plane-wave right-hand side -> solve
  (+ FP64 iterative refinement in mixed precision) -> far field -> cross sections.
"""
import warnings

import numpy as np
import torch

from . import _lib
from ._lib import call, stage

C128 = torch.complex128
C64 = torch.complex64


def _p(t):
    return t.data_ptr() if t is not None else None


def stream_handle():
    return torch.cuda.current_stream().cuda_stream


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("mnpbem_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def connectivity(inout):
    """Region connectivity of the composite particle (greenfun/compgreen_ret.py:435-505):
    con[r1][r2][a, b] = material joining region r1 of particle a and region r2 of particle b."""
    io = np.asarray(inout)
    con = [[None, None], [None, None]]
    for r1 in range(2):
        for r2 in range(2):
            a = io[:, r1][:, None]
            b = io[:, r2][None, :]
            con[r1][r2] = np.where(a == b, a, 0).astype(np.int32)
    return con


class SingularMatrixWarning(RuntimeWarning):
    """An LU factor met an exact zero pivot (LAPACK info > 0): the solution holds inf/NaN."""


def warn_singular(who, info):
    """Warn about zero pivots (LAPACK zgetrf convention: info j > 0 = column j-1), as scipy's
    LinAlgWarning does; the reference does not check (utils/gpu.py:189-191)."""
    bad = np.asarray(info).ravel()
    bad = bad[bad > 0]
    warnings.warn("{}: exactly singular factor (zero pivot in column {}); results hold inf/NaN"
                  .format(who, int(bad[0]) - 1), SingularMatrixWarning, stacklevel=3)


class DeviceGeometry:
    """Per-face geometry, refinement data and media of a ComParticle, uploaded once."""

    def __init__(self, cp, nearfield, dev=None):
        self.dev = dev or device()
        self.cp = cp
        self.n = cp.nfaces
        self.sface = 0          # one geometry for every matrix of a batch
        d = self.dev
        f64 = torch.float64
        self.pos = torch.tensor(np.ascontiguousarray(cp.pos), dtype=f64, device=d)
        self.nvec = torch.tensor(np.ascontiguousarray(cp.nvec), dtype=f64, device=d)
        self.area = torch.tensor(np.ascontiguousarray(cp.area), dtype=f64, device=d)
        self.fpart = torch.tensor(cp.face_particle.astype(np.int32), device=d)
        con = connectivity(cp.inout)
        # nested particles (a core inside a shell: one particle's outside medium is another's
        # inside) add the cross-region blocks G1 = G(in, in) - G(out, in), G2 = G(out, out) -
        # G(in, out) (bem_ret.py:407-433). For a pair of particles at most one of the two terms of
        # each side exists (both would need a particle whose inside and outside media agree), so
        # the tables hold one signed material per pair: +m for G(in, in) / G(out, out), -m for the
        # subtracted cross term, 0 for none.
        self.nested = bool(np.any(con[1][0]) or np.any(con[0][1]))
        if np.any((con[0][0] != 0) & (con[1][0] != 0)) or np.any((con[1][1] != 0) & (con[0][1] != 0)):
            raise ValueError("a particle's inside and outside media must differ")
        self.npart = cp.inout.shape[0]
        self.con_in = torch.tensor((con[0][0] - con[1][0]).ravel().astype(np.int32), device=d)
        self.con_out = torch.tensor((con[1][1] - con[0][1]).ravel().astype(np.int32), device=d)
        self.nmat = len(cp.eps)
        nf = nearfield
        if nf is not None:   # None: far-field use only (no matrix assembly)
            self.order = nf.order
            self.nref = nf.n_ref
            self.ref_row = torch.tensor(nf.rows.astype(np.int32), device=d)
            self.ref_col = torch.tensor(nf.cols.astype(np.int32), device=d)
            self.ref_g = torch.tensor(np.ascontiguousarray(nf.g), dtype=f64, device=d)
            self.ref_f = torch.tensor(np.ascontiguousarray(nf.f), dtype=f64, device=d)
        self._nf = nf
        self._csr = None
        self._side = {}

    def refined_csr(self):
        """Refined pairs as a CSR with ascending columns per row: (rowptr, col, g, f) on the
        device, for entry-wise evaluation (H-matrices)."""
        if self._csr is None:
            nf, d = self._nf, self.dev
            order = np.lexsort((nf.cols, nf.rows))
            rows = nf.rows[order]
            rowptr = np.searchsorted(rows, np.arange(self.n + 1)).astype(np.int32)
            self._csr = (torch.tensor(rowptr, device=d),
                         torch.tensor(nf.cols[order].astype(np.int32), device=d),
                         torch.tensor(np.ascontiguousarray(nf.g[order]), dtype=torch.float64, device=d),
                         torch.tensor(np.ascontiguousarray(nf.f[order]), dtype=torch.float64, device=d))
        return self._csr

    def side(self, medium):
        """Per face: bit 0 = inside medium is ``medium``, bit 1 = outside medium is."""
        if medium not in self._side:
            io = self.cp.inout_faces
            s = (io[:, 0] == medium).astype(np.int32) + 2 * (io[:, 1] == medium).astype(np.int32)
            self._side[medium] = torch.tensor(s, device=self.dev)
        return self._side[medium]

    def media(self, enei_list):
        """k per (wavelength, material), eps inside/outside per (wavelength, face), k vacuum."""
        enei_list = np.atleast_1d(np.asarray(enei_list, float))
        ek = [f(enei_list) for f in self.cp.eps]
        kmat = np.stack([np.broadcast_to(np.asarray(e[1], complex), enei_list.shape) for e in ek], axis=1)
        epsm = np.stack([np.broadcast_to(np.asarray(e[0], complex), enei_list.shape) for e in ek], axis=1)
        io = self.cp.inout_faces
        L1 = epsm[:, io[:, 0] - 1]
        L2 = epsm[:, io[:, 1] - 1]
        d = self.dev
        return (torch.tensor(np.ascontiguousarray(kmat), dtype=C128, device=d),
                torch.tensor(np.ascontiguousarray(L1), dtype=C128, device=d),
                torch.tensor(np.ascontiguousarray(L2), dtype=C128, device=d),
                torch.tensor(2 * np.pi / enei_list, dtype=torch.float64, device=d),
                kmat, epsm)


class GreenBatch:
    """G1, G2, H1, H2 for a batch of wavelengths in one precision."""

    def __init__(self, geo, batch, dtype=C128):
        self.geo, self.batch, self.dtype = geo, batch, dtype
        n = geo.n
        self.ld = n
        self.stride = n * self.ld
        mk = lambda: torch.empty((batch, n, self.ld), dtype=dtype, device=geo.dev)
        self.G1, self.G2, self.H1, self.H2 = mk(), mk(), mk(), mk()

    @property
    def prec(self):
        return 0 if self.dtype == C128 else 1

    def assemble(self, kmat, nlam):
        g = self.geo
        call("mnp_green_assemble", g.n, g.npart, g.nmat, nlam, g.order, _p(g.pos), _p(g.nvec), _p(g.area),
             _p(g.fpart), _p(g.con_in), _p(g.con_out), _p(kmat), g.nref, _p(g.ref_row), _p(g.ref_col),
             _p(g.ref_g), _p(g.ref_f), self.prec, _p(self.G1), _p(self.G2), _p(self.H1), _p(self.H2),
             self.ld, self.stride, stream_handle())

    def assemble_pair(self, other, kmat, nlam):
        """This (complex128) batch and its complex-float copy ``other`` from one pass."""
        g = self.geo
        call("mnp_green_assemble_pair", g.n, g.npart, g.nmat, nlam, g.order, _p(g.pos), _p(g.nvec), _p(g.area),
             _p(g.fpart), _p(g.con_in), _p(g.con_out), _p(kmat), g.nref, _p(g.ref_row), _p(g.ref_col),
             _p(g.ref_g), _p(g.ref_f), _p(self.G1), _p(self.G2), _p(self.H1), _p(self.H2), self.ld, self.stride,
             _p(other.G1), _p(other.G2), _p(other.H1), _p(other.H2), other.ld, other.stride, stream_handle())

    @staticmethod
    def to_numpy(t, b):
        """Column-major device matrix b -> (n, n) numpy array."""
        return t[b].transpose(0, 1).cpu().numpy()


class RetSolverBatch:
    """Dense BEMRet factors for up to ``batch`` wavelengths.

    precision: 'double' (complex128 chain), 'single' (complex64 chain, no refinement) or
    'mixed' (complex64 chain as approximate inverse, FP64 iterative refinement against the
    exact complex128 operator; wavelengths whose refinement stalls are redone in 'double').
    """

    def __init__(self, geo, batch, precision="double", nb=32, refine_tol=1e-12, max_refine=12, tuning=None):
        if precision not in ("double", "single", "mixed"):
            raise ValueError("precision must be 'double', 'single' or 'mixed'")
        self.geo, self.batch, self.precision, self.nb = geo, batch, precision, nb
        self.tuning = tuning          # LU knobs applied around the factorisation (autotune)
        self.refine_tol, self.max_refine = refine_tol, max_refine
        n, d = geo.n, geo.dev
        fdt = C128 if precision == "double" else C64
        self.fac = GreenBatch(geo, batch, fdt)   # G1, G2 -> LU; H1, H2 scratch
        mk = lambda: torch.empty((batch, n, n), dtype=fdt, device=d)
        self.S1, self.S2, self.D = mk(), mk(), mk()
        self.exact = GreenBatch(geo, batch, C128) if precision == "mixed" else None
        self.ipiv = torch.empty((4, batch, n), dtype=torch.int32, device=d)
        self.perm = torch.empty((8, batch, n), dtype=torch.int32, device=d)
        self.info = torch.empty((4, batch), dtype=torch.int32, device=d)
        self.work = torch.empty((batch, 24, n), dtype=fdt, device=d)
        self.opwork = torch.empty((batch, 32, n), dtype=C128, device=d) if precision == "mixed" else None
        self.nlam = 0
        self.stats = {}
        self.tune_assembly = False     # time the assembly shapes on the first factorisation
        self.asm_cfg = {}              # chosen shape per precision of the assembled matrices

    @property
    def prec(self):
        return self.fac.prec

    def factor(self, enei_list):
        enei_list = np.atleast_1d(np.asarray(enei_list, float))
        nl = len(enei_list)
        if nl > self.batch:
            raise ValueError("batch holds {} wavelengths, got {}".format(self.batch, nl))
        g = self.geo
        self.kmat, self.L1, self.L2, self.kvac, self.kmat_np, self.eps_np = g.media(enei_list)
        self.enei = enei_list
        self.nlam = nl
        if self.exact is not None:
            # mixed precision: the exact complex128 operator and the complex-float factor
            # matrices from one pass (each exponential evaluated once)
            from .autotune import knobs, tune_assembly
            key = str(C128)
            if self.tune_assembly and key not in self.asm_cfg:
                self.asm_cfg[key] = tune_assembly(self.exact, self.kmat, nl)
            with knobs(self.asm_cfg.get(key)):
                self.exact.assemble_pair(self.fac, self.kmat, nl)
        else:
            self._assemble_tuned(self.fac, nl)
        return self._factor_assembled()

    def _assemble_tuned(self, gb, nl):
        from .autotune import knobs, tune_assembly
        key = str(gb.dtype)
        if self.tune_assembly and key not in self.asm_cfg:
            self.asm_cfg[key] = tune_assembly(gb, self.kmat, nl)
        with knobs(self.asm_cfg.get(key)):
            gb.assemble(self.kmat, nl)

    def _needs_general(self):
        """Whether the chain needs L1 = G1 eps1 G1^-1, L2 = G2 eps2 G2^-1 as full matrices
        (bem_ret.py:450-465): the inside of one particle is the outside of another (nested
        particles) and eps is not one value over all faces. Otherwise L1, L2 are the per-face
        eps (diagonal)."""
        if not getattr(self.geo, "nested", False):
            return False
        nl = self.nlam
        e1, e2 = self.L1[:nl], self.L2[:nl]
        uniform = bool(torch.all(e1 == e1[:, :1]).item()) and bool(torch.all(e2 == e2[:, :1]).item())
        return not uniform

    def _factor_assembled(self):
        """LU chain of the assembled G1, G2, H1, H2 of the first self.nlam matrices."""
        g, f, nl = self.geo, self.fac, self.nlam
        self.explicit = False
        self.general = self._needs_general()
        from .autotune import knobs
        with knobs(self.tuning):
            if self.general:
                self._factor_general()
            else:
                call("mnp_bemret_factor", self.prec, g.n, nl, _p(f.G1), _p(f.G2), _p(f.H1), _p(f.H2), _p(self.S1),
                     _p(self.S2), _p(self.D), f.ld, f.stride, _p(self.ipiv), _p(self.perm), _p(self.info),
                     _p(g.nvec), g.sface, _p(self.L1), _p(self.L2), _p(self.kvac), self.nb, stream_handle())
        info = self.check_info()
        if np.any(info < 0):
            raise RuntimeError("LU panel synchronisation timed out (info {}); factors are invalid".format(info))
        self.stats["zero_pivot"] = int(np.count_nonzero(info))
        # mixed precision: a zero pivot of the complex64 chain shows up as a stalled refinement,
        # and that wavelength is redone with the complex128 chain
        if self.precision != "mixed" and np.any(info > 0):
            warn_singular("BEMRet ({})".format(self.precision), info)
        return self

    # -- general chain (nested particles with non-uniform eps) ---------------------------------
    # Same storage as the diagonal chain: G1, G2 hold their LUs, S1 = Sigma1, D = Delta^-1,
    # S2 = LU(Sigma); pivots, permutations and info in the layout of mnp_bemret_factor
    # (factor q of G1, G2, Delta, Sigma at offset q nlam n, permutations (pf, qinv) at 2q, 2q + 1).
    # L1, L2 are kept as full matrices for the solve. Elementwise steps run on torch device tensors;
    # the LUs, right-side solves and products run on the library's kernels.

    def _lu(self, q, A):
        n, nl, pb = self.geo.n, self.nlam, self.nlam * self.geo.n
        ip = self.ipiv.view(-1)[q * pb:(q + 1) * pb]
        pm = self.perm.view(-1)
        call("mnp_getrf_batched", self.prec, n, nl, _p(A), n, n * n, _p(ip), n,
             self.info.data_ptr() + 4 * q * nl, self.nb, stream_handle())
        call("mnp_perm_batched", n, nl, _p(ip), n, pm.data_ptr() + 4 * 2 * q * pb,
             pm.data_ptr() + 4 * (2 * q + 1) * pb, n, stream_handle())

    def _rsolve(self, q, A, B, X):
        """X = B A^-1 (A: LU number q; B is overwritten)."""
        n, nl, pb = self.geo.n, self.nlam, self.nlam * self.geo.n
        call("mnp_rsolve_batched", self.prec, n, n, nl, _p(A), n, n * n,
             self.perm.data_ptr() + 4 * (2 * q + 1) * pb, n, _p(B), n, n * n, _p(X), n, n * n, self.nb,
             stream_handle())

    def _gemm_sub(self, C, A, B):
        """C -= A B for the first nlam matrices."""
        n = self.geo.n
        call("mnp_gemm_sub_batched", self.prec, n, n, n, _p(A), n, n * n, _p(B), n, n * n, _p(C), n, n * n,
             self.nlam, stream_handle())

    def _factor_general(self):
        """bem_ret.py:438-495 with L1 = G1 eps1 G1^-1, L2 = G2 eps2 G2^-1 (right-side solves of
        the column-scaled G1, G2) and Sigma = Sigma1 L1 - Sigma2 L2 + k^2 ((L Delta^-1) o n n^T) L."""
        g, f, nl, n = self.geo, self.fac, self.nlam, self.geo.n
        fdt = f.dtype
        if getattr(self, "_Lm", None) is None:
            mk = lambda: torch.empty((self.batch, n, n), dtype=fdt, device=g.dev)
            self._Lm = [mk(), mk(), mk()]
        L1m, L2m, T = self._Lm
        G1, G2, H1, H2 = f.G1[:nl], f.G2[:nl], f.H1[:nl], f.H2[:nl]
        S1, S2, D = self.S1[:nl], self.S2[:nl], self.D[:nl]
        e1 = self.L1[:nl].to(fdt)
        e2 = self.L2[:nl].to(fdt)
        # matrix (i, j) of entry b sits at [b, j, i]: scaling column j is a scaling of row j here
        torch.mul(G1, e1[:, :, None], out=T[:nl])
        torch.mul(G2, e2[:, :, None], out=L2m[:nl])
        self._lu(0, f.G1)
        self._lu(1, f.G2)
        self._rsolve(0, f.G1, T, L1m)                # L1 = (G1 eps1) G1^-1
        self._rsolve(1, f.G2, L2m, T)                # L2 = (G2 eps2) G2^-1, in T
        L2m, T = T, L2m
        self._Lm = [L1m, L2m, T]
        self._rsolve(0, f.G1, f.H1, self.S1)         # Sigma1 = H1 G1^-1
        self._rsolve(1, f.G2, f.H2, self.S2)         # Sigma2 = H2 G2^-1 (in S2 for now)
        torch.sub(S1, S2, out=H1)                    # Delta
        self._lu(2, f.H1)
        H2.zero_()
        H2.diagonal(dim1=1, dim2=2).fill_(1)
        self._rsolve(2, f.H1, f.H2, self.D)          # Delta^-1
        H1.zero_()
        self._gemm_sub(f.H1, self.S2, L2m)           # H1 = -Sigma2 L2
        torch.sub(L2m[:nl], L1m[:nl], out=S2)        # S2 = -L
        H2.zero_()
        self._gemm_sub(f.H2, self.S2, self.D)        # H2 = L Delta^-1
        nv = g.nvec.to(fdt)
        nn = nv @ nv.transpose(0, 1)                 # symmetric: its layout does not matter
        k2 = (self.kvac[:nl] ** 2).to(fdt)
        H2.mul_(nn[None]).mul_(k2[:, None, None])    # k^2 (L Delta^-1) o n n^T
        S1.neg_()
        self._gemm_sub(f.H1, self.S1, L1m)           # + Sigma1 L1
        S1.neg_()
        self._gemm_sub(f.H1, f.H2, self.S2)          # + k^2 ((L Delta^-1) o n n^T) L
        S2.copy_(H1)
        self._lu(3, self.S2)

    def _inv_apply(self, q, M, X, Y, r):
        """Y = M^-1 X (r columns) from LU number q of M, or M X when M holds the inverse."""
        n, nl, pb = self.geo.n, self.nlam, self.nlam * self.geo.n
        if getattr(self, "explicit", False):
            call("mnp_gemv_batched", self.prec, n, r, nl, _p(M), n, n * n, _p(X), n, r * n, _p(Y), n, r * n,
                 stream_handle())
        else:
            call("mnp_getrs_batched", self.prec, n, r, nl, _p(M), n, n * n, self.perm.data_ptr() + 4 * 2 * q * pb,
                 n, _p(X), n, r * n, _p(Y), n, r * n, stream_handle())

    def _gemv(self, M, X, r):
        n, nl = self.geo.n, self.nlam
        Y = torch.empty((nl, r, n), dtype=X.dtype, device=X.device)
        call("mnp_gemv_batched", self.prec, n, r, nl, _p(M), n, n * n, _p(X), n, r * n, _p(Y), n, r * n,
             stream_handle())
        return Y

    def _chain_general(self, rhs, out):
        """Eqs. (19)-(20) of Garcia de Abajo and Howie (bem_ret.py:1708-1752) with full L1, L2:
        rhs = [phi a De alpha], out = [sig1 h1 sig2 h2] (B, 8, n) of the factor precision."""
        g, f, nl, n = self.geo, self.fac, self.nlam, self.geo.n
        L1m, L2m = self._Lm[0], self._Lm[1]
        x = rhs[:nl]
        ik = (1j * self.kvac[:nl]).to(x.dtype)[:, None]
        nv = g.nvec.to(x.dtype).transpose(0, 1)            # (3, n)
        phi, a, De, alpha = x[:, 0], x[:, 1:4], x[:, 4], x[:, 5:8]
        La = self._gemv(L1m, x[:, 0:4].contiguous(), 4)     # [L1 phi, L1 a]
        L1phi, L1a = La[:, 0], La[:, 1:4]
        W = self._gemv(self.S1, torch.cat([a, L1phi[:, None]], 1).contiguous(), 4)
        alpha2 = alpha - W[:, 0:3] + ik[:, None] * nv[None] * L1phi[:, None]
        De2 = De - W[:, 3] + ik * torch.sum(nv[None] * L1a, 1)
        U = self._gemv(self.D, alpha2.contiguous(), 3)
        LU_ = self._gemv(L1m, U, 3) - self._gemv(L2m, U, 3)
        R = (De2 + ik * torch.sum(nv[None] * LU_, 1))[:, None].contiguous()
        sig2 = torch.empty_like(R)
        self._inv_apply(3, self.S2, R, sig2, 1)
        w = self._gemv(L1m, sig2, 1) - self._gemv(L2m, sig2, 1)
        Tt = (ik[:, None] * nv[None] * w + alpha2).contiguous()
        h2 = self._gemv(self.D, Tt, 3)
        P1 = torch.cat([sig2 + phi[:, None], h2 + a], 1).contiguous()
        P2 = torch.cat([sig2, h2], 1).contiguous()
        o1 = torch.empty_like(P1)
        o2 = torch.empty_like(P2)
        self._inv_apply(0, f.G1, P1, o1, 4)
        self._inv_apply(1, f.G2, P2, o2, 4)
        out[:nl, 0:4] = o1
        out[:nl, 4:8] = o2

    def invert(self):
        """Replaces the LU factors of G1, G2 and Sigma by their explicit inverses (right-side
        solves of the identity, 16/3 n^3 complex flops each), so that every later solve applies
        them as matrix-vector products: for a factorisation reused by many solves (a fixed GMRES
        preconditioner), whose triangular substitutions are latency-bound."""
        if getattr(self, "explicit", False):
            return self
        g, f, nl, n = self.geo, self.fac, self.nlam, self.geo.n
        pb = nl * n
        qinv = {0: 1, 1: 3, 2: 7}              # perm layout: (pf, qinv) pairs of G1, G2, Delta, Sigma
        eye, out = f.H1, f.H2                   # scratch after the factorisation
        for i, M in enumerate((f.G1, f.G2, self.S2)):
            eye[:nl].zero_()
            eye[:nl].diagonal(dim1=1, dim2=2).fill_(1)
            call("mnp_rsolve_batched", self.prec, n, n, nl, _p(M), f.ld, f.stride,
                 self.perm.data_ptr() + 4 * qinv[i] * pb, n, _p(eye), f.ld, f.stride, _p(out), f.ld, f.stride,
                 self.nb, stream_handle())
            M[:nl].copy_(out[:nl])
        self.explicit = True
        return self

    def check_info(self):
        """(4, nlam) LU info of G1, G2, Delta, Sigma: 0 = ok, j = zero pivot in column j-1.
        The factor call lays pivots and info out for the wavelengths it was given."""
        nl = self.nlam
        return self.info.view(-1)[: 4 * nl].reshape(4, nl).cpu().numpy()

    def _chain(self, rhs, out):
        """out = M rhs with the factor precision (rhs, out: (B, 8, n) of that dtype)."""
        if getattr(self, "general", False):
            self._chain_general(rhs, out)
            return
        g, f, nl, n = self.geo, self.fac, self.nlam, self.geo.n
        perm = None if getattr(self, "explicit", False) else _p(self.perm)
        call("mnp_bemret_solve", self.prec, n, nl, _p(f.G1), _p(f.G2), _p(self.S1), _p(self.D), _p(self.S2),
             f.ld, f.stride, perm, _p(g.nvec), g.sface, _p(self.L1), _p(self.L2), _p(self.kvac), _p(rhs),
             8 * n,
             _p(out), 8 * n, _p(self.work), 24 * n, stream_handle())

    def residual(self, x, rhs, out, norms=None):
        """out = rhs - A x with the exact complex128 operator; norms (B, 2) = |r|^2, |rhs|^2."""
        e, g, n, nl = self.exact, self.geo, self.geo.n, self.nlam
        call("mnp_bem_operator", n, nl, _p(e.G1), _p(e.G2), _p(e.H1), _p(e.H2), e.ld, e.stride, _p(g.nvec),
             g.sface, _p(self.L1), _p(self.L2), _p(self.kvac), _p(x), 8 * n, _p(rhs), 8 * n, _p(out), 8 * n,
             _p(self.opwork), 32 * n, _p(norms), stream_handle())

    def solve(self, rhs):
        """Solve for packed right-hand sides rhs (B, 8, n) complex128 -> (B, 8, n) complex128."""
        n, nl = self.geo.n, self.nlam
        if rhs.dim() != 3 or rhs.shape[1:] != (8, n) or rhs.stride() != (8 * n, n, 1):
            raise ValueError("rhs must be a contiguous (batch, 8, n) tensor")
        if rhs.shape[0] < nl:
            raise ValueError("rhs holds {} right-hand sides; the factors cover {} wavelengths"
                             .format(rhs.shape[0], nl))
        if rhs.dtype != C128:
            raise ValueError("rhs must be complex128")
        if self.precision == "double":
            out = torch.empty(rhs.shape, dtype=rhs.dtype, device=rhs.device)
            self._chain(rhs, out)
            return out
        r32 = torch.empty((self.batch, 8, n), dtype=C64, device=rhs.device)
        dx = torch.empty_like(r32)
        call("mnp_cast_batched", 1, 8 * n, nl, _p(rhs), 8 * n, _p(r32), 8 * n, stream_handle())
        self._chain(r32, dx)
        x = torch.zeros(rhs.shape, dtype=rhs.dtype, device=rhs.device)
        call("mnp_accum_batched", 1, 8 * n, nl, _p(dx), 8 * n, _p(x), 8 * n, None, stream_handle())
        if self.precision == "single":
            return x
        r = torch.empty(rhs.shape, dtype=rhs.dtype, device=rhs.device)
        norms = torch.zeros((self.batch, 2), dtype=torch.float64, device=rhs.device)
        active = torch.ones(self.batch, dtype=torch.int32, device=rhs.device)
        hist = []
        prev = None
        stalled = np.zeros(nl, dtype=bool)
        for it in range(self.max_refine + 1):
            self.residual(x, rhs, r, norms)
            nr = norms[:nl].cpu().numpy()
            rel = np.sqrt(nr[:, 0] / np.maximum(nr[:, 1], 1e-300))
            hist.append(rel)
            done = rel <= self.refine_tol
            if prev is not None:
                stalled |= (~done) & (rel > 0.5 * prev)
            if np.all(done | stalled) or it == self.max_refine:
                break
            prev = rel
            act = (~done) & (~stalled)
            active[:nl] = torch.tensor(act.astype(np.int32), device=rhs.device)
            with stage("refine"):
                call("mnp_cast_batched", 1, 8 * n, nl, _p(r), 8 * n, _p(r32), 8 * n, stream_handle())
                self._chain(r32, dx)
                call("mnp_accum_batched", 1, 8 * n, nl, _p(dx), 8 * n, _p(x), 8 * n, _p(active), stream_handle())
        self.stats = dict(relres=hist[-1], history=np.array(hist), iterations=len(hist) - 1,
                          fallback=np.nonzero(~(hist[-1] <= self.refine_tol))[0])
        return x


def planewave_rhs(geo, kvac, kmed, L1, L2, pol, direc, medium, nlam, batch):
    n = geo.n
    out = torch.empty((batch, 8, n), dtype=C128, device=geo.dev)
    pol = np.ascontiguousarray(np.asarray(pol, float).ravel())
    direc = np.ascontiguousarray(np.asarray(direc, float).ravel())
    call("mnp_planewave_rhs", n, nlam, _p(geo.pos), _p(geo.nvec), _p(geo.side(medium)), geo.sface, _p(kmed),
         _p(kvac),
         _p(L1), _p(L2), pol.ctypes.data, direc.ctypes.data, _p(out), 8 * n, stream_handle())
    return out


def farfield(geo, x, kvac, kmed, dirs, medium, nlam):
    """(nlam, ndir, 6) complex: e (3), h (3) per direction."""
    n = geo.n
    dirs_t = torch.tensor(np.ascontiguousarray(np.atleast_2d(dirs), dtype=float), device=geo.dev)
    ndir = dirs_t.shape[0]
    out = torch.empty((nlam, ndir, 6), dtype=C128, device=geo.dev)
    call("mnp_farfield", n, nlam, _p(geo.pos), _p(geo.area), _p(geo.side(medium)), geo.sface, _p(kmed), _p(kvac),
         ndir,
         _p(dirs_t), _p(x), 8 * n, _p(out), stream_handle())
    return out
