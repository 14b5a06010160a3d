"""What a GMRES start can gain on the wavelength sweep of the learned-start workload, measured on
the oracle (CPU, small rod dimer).

Setting of the learned warm start: BEMRetIter on a gold-Drude rod dimer (two 20 x 80 nm capped
rods, 2 nm gap, axial plane wave), GMRES(50) to 1e-6 with the BEMRet chain factored once at
750 nm as right preconditioner (bem/bem_ret_iter.py:1636-1744 with the preconditioner of
:801-1086 held fixed), wavelengths 600-900 nm.

The starts compared (each through the residual-minimising combination of its candidate vectors
that BEMRetIter.solve uses for a list of guesses):

* the span of the exact solution's four per-face channels sigma1, sigma2, h1.pol, h2.pol (the
  best any per-face predictor of those quantities could give, whatever their amplitude);
* the exact solution with 5 % per-face random error (a predictor far better than a per-face
  network reaches);
* the previous wavelengths' solutions (1 and 3 of them).

Measured: the first two leave the iteration count where the zero start has it, while the
previous solutions cut it. GMRES spends its iterations on the spread of the preconditioned
operator's spectrum between the preconditioner's wavelength and the solve's, and only an error
that is both small and smooth in wavelength removes them. This pins the warm-start design: the
learned start is one candidate beside the previous solutions, and the previous-solution
extrapolation is the start that pays.
"""
import numpy as np
import pytest

from oracle import bem_oracle as O
import mnpbem_amd as M

DRUDE = (10.0, 9.065, 0.0708)


def _dimer(nrod, gap=2.0):
    p1 = M.trirod(20, 80, nrod, triangles=True).shift([0, 0, -(40 + gap / 2)])
    p2 = M.trirod(20, 80, nrod, triangles=True).shift([0, 0, 40 + gap / 2])
    return M.ComParticle([M.EpsConst(1.0), M.EpsDrude(*DRUDE)], [p1, p2], [[2, 1], [2, 1]], 1, 2, interp="curv")


def _gmres(A, Minv, b, x0, tol=1e-6, m=50, maxit=300):
    """Restarted right-preconditioned GMRES (CGS2), iterations to |b - A x| <= tol |b|."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    bn = np.linalg.norm(b)
    it = 0
    while it < maxit:
        r = b - A(x)
        beta = np.linalg.norm(r)
        if beta / bn <= tol:
            return x, it
        V, Z = [r / beta], []
        H = np.zeros((m + 1, m), complex)
        for j in range(m):
            z = Minv(V[j])
            Z.append(z)
            w = A(z)
            for _ in range(2):
                h = np.array([np.vdot(v, w) for v in V])
                H[: j + 1, j] += h
                w = w - np.array(V).T @ h
            H[j + 1, j] = np.linalg.norm(w)
            V.append(w / H[j + 1, j])
            it += 1
            e = np.zeros(j + 2, complex)
            e[0] = beta
            y = np.linalg.lstsq(H[: j + 2, : j + 1], e, rcond=None)[0]
            if np.linalg.norm(H[: j + 2, : j + 1] @ y - e) / bn <= tol or it >= maxit:
                break
        x = x + np.array(Z).T @ y
    return x, it


def _best(A, b, cands):
    AG = np.array([A(g) for g in cands]).T
    c = np.linalg.lstsq(AG, b, rcond=1e-12)[0]
    return np.array(cands).T @ c


def test_start_gain_needs_wavelength_smooth_error():
    cp = _dimer([10, 4, 8])
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    fp = cp.face_particle
    n = s.n
    green = O.Green(s, fp, np.asarray(cp.inout), [(np.nonzero(fp == a)[0],) * 2 + (1.0,) for a in range(2)])
    pol, direc = np.array([0, 0, 1.0]), np.array([1.0, 0, 0])

    def system(lam):
        ek = [O.eps_const(lam, 1.0), O.eps_drude(lam, *DRUDE)]
        G1, G2, H1, H2 = O.matrices(green, [e[1] for e in ek])
        e1, e2 = np.full(n, ek[1][0]), np.full(n, ek[0][0])
        k0 = 2 * np.pi / lam
        a2, a2p = O.planewave(s.pos, s.nvec, pol, direc, lam, 1.0, np.arange(n))
        z = np.zeros_like(a2)
        phi, a, alpha, De = O.excitation(s.nvec, k0, e1, e2, z, z, a2, a2p)
        return (G1, G2, H1, H2, e1, e2, k0), O.pack(phi, a, De, alpha)

    mats750, _ = system(750.0)
    bem = O.BEMRetOracle(*mats750[:4], s.nvec, *mats750[4:])

    def Minv(v):
        V = v.reshape(8, n)
        s1, h1, s2, h2 = bem.solve(V[0], V[1:4].T, V[5:8].T, V[4])
        return np.concatenate([s1, h1.T.ravel(), s2, h2.T.ravel()])

    rng = np.random.default_rng(3)
    total = dict(zero=0, channels=0, noisy=0, previous=0, previous3=0)
    prev = []
    for lam in np.linspace(600, 700, 11):
        mats, b = system(lam)
        A = lambda x, mats=mats: O.afun(*mats[:4], s.nvec, *mats[4:], x)
        xe = _gmres(A, Minv, b, None, tol=1e-12)[0]
        X = xe.reshape(8, n)

        def vec(slot, val):
            v = np.zeros((8, n), complex)
            v[slot] = val
            return v.ravel()
        chans = [vec(0, X[0]), vec(4, X[4]), vec(3, X[1:4].T @ pol), vec(7, X[5:8].T @ pol)]
        iters = dict(zero=_gmres(A, Minv, b, None)[1],
                     channels=_gmres(A, Minv, b, _best(A, b, chans))[1],
                     noisy=_gmres(A, Minv, b, xe * (1 + 0.05 * rng.standard_normal(xe.shape)))[1])
        if len(prev) >= 3:
            iters["previous"] = _gmres(A, Minv, b, prev[-1])[1]
            iters["previous3"] = _gmres(A, Minv, b, _best(A, b, prev[-3:]))[1]
            for k in ("previous", "previous3"):
                total[k] += iters[k]
            total["zero_late"] = total.get("zero_late", 0) + iters["zero"]
        for k in ("zero", "channels", "noisy"):
            total[k] += iters[k]
        prev.append(xe)
    print("GMRES iterations over the sweep:", total)
    # exact per-face channels with free complex amplitudes, or the exact solution with a 5 %
    # per-face error: no gain worth an operator application
    assert total["channels"] >= 0.9 * total["zero"], total
    assert total["noisy"] >= 0.85 * total["zero"], total
    # the previous wavelengths' solutions do cut the iterations
    assert total["previous"] < total["zero_late"], total
    assert total["previous3"] <= 0.6 * total["zero_late"], total
