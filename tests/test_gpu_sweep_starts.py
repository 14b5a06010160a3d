"""BEMRetIter over a wavelength sweep with a fixed preconditioner: explicit inverse factors and
the GMRES starts of BEMRetIter.solve_spectrum (the reference solves each wavelength from zero with
the preconditioner of that wavelength, bem_ret_iter.py:1636-1744)."""
import numpy as np
import pytest

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import mnpbem_amd as M                                   # noqa: E402
from mnpbem_amd.bem_iter import BEMRetIter              # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)


def _rod_dimer(nrod, gap=2.0):
    p1 = M.trirod(20, 80, nrod, triangles=True).shift([0, 0, -(40 + gap / 2)])
    p2 = M.trirod(20, 80, nrod, triangles=True).shift([0, 0, 40 + gap / 2])
    return M.ComParticle([M.EpsConst(1.0), M.EpsDrude(*DRUDE)], [p1, p2], [[2, 1], [2, 1]], 1, 2, interp="curv")


@pytest.fixture(scope="module")
def dimer():
    cp = _rod_dimer([12, 5, 10])
    exc = M.PlaneWaveRet([0, 0, 1.0], [1.0, 0, 0])
    return cp, exc


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b)))


def test_fixed_preconditioner_inverse_equals_substitution(dimer):
    """A preconditioner factored once at 750 nm and applied through explicit inverses of G1, G2
    and Sigma gives the same GMRES iterates as the substitutions (to rounding) and the direct
    solution at another wavelength."""
    cp, exc = dimer
    lam = 640.0
    banner("fixed preconditioner inverse vs substitution n={}".format(cp.nfaces))
    direct, _ = M.BEMRet(cp).solve(exc(cp, lam))
    res = {}
    for inv in (False, True):
        it = BEMRetIter(cp, tol=1e-10, maxit=200, restart=50, precond_wavelength=750.0, precond_inverse=inv)
        sig, _ = it.solve(exc(cp, lam))
        flag, relres, niter = it.info()
        assert flag == 0 and relres <= 1e-10
        res[inv] = (sig, niter)
        for f in ("sig1", "h1", "sig2", "h2"):
            assert _rel(getattr(sig, f), getattr(direct, f)) < 1e-7, (inv, f)
    assert abs(res[True][1] - res[False][1]) <= 1
    assert _rel(res[True][0].sig2, res[False][0].sig2) < 1e-8


def test_sweep_starts_agree_and_extrapolation_cuts_iterations(dimer):
    """Every start reaches the same cross sections (GMRES to 1e-8); the residual-minimising
    combination of the previous three solutions needs far fewer iterations than the zero start;
    a useless learned guess (zeros) costs nothing but its operator application."""
    cp, exc = dimer
    lams = np.linspace(620, 700, 9)
    banner("sweep starts n={} {} wavelengths".format(cp.nfaces, len(lams)))
    it = BEMRetIter(cp, tol=1e-8, maxit=300, restart=50, precond_wavelength=750.0)
    zero_guess = lambda lam: np.zeros(8 * cp.nfaces, complex)
    out = it.solve_spectrum(exc, lams, start=("zero", "previous", "extrapolate", "extrapolate+learned"), history=3,
                            warm_start=zero_guess)
    for start in out:
        assert all(f == 0 for f in out[start]["flag"]), start
    single = it.solve_spectrum(exc, lams[:2], start="previous")
    assert single["iterations"][0] == out["previous"]["iterations"][0]
    ref = out["zero"]["ext"]
    for start in out:
        assert _rel(out[start]["ext"], ref) < 1e-6, start
    tot = {k: sum(v["iterations"][3:]) for k, v in out.items()}
    print("iterations after the first three wavelengths:", tot)
    assert tot["previous"] < tot["zero"]
    assert tot["extrapolate"] <= 0.7 * tot["zero"]
    assert tot["extrapolate+learned"] <= tot["extrapolate"] + 2
    assert out["extrapolate"]["initial_residual"][0] == 1.0
    assert max(out["extrapolate"]["initial_residual"][3:]) < 0.1


@pytest.mark.parametrize("hmatrix", [False, True])
def test_particle_block_preconditioner(dimer, hmatrix):
    """preconditioner='particle': the BEMRet chain of each rod alone (one batch of two entries)
    as a block-diagonal preconditioner; GMRES reaches the direct solution, in more iterations
    than with the chain of the whole dimer but far fewer than without a preconditioner."""
    cp, exc = dimer
    lam = 720.0
    banner("particle block preconditioner n={} hmatrix={}".format(cp.nfaces, hmatrix))
    direct, _ = M.BEMRet(cp).solve(exc(cp, lam))
    opts = dict(tol=1e-9, maxit=400, restart=50)
    if hmatrix:
        opts.update(hmatrix=True, htol=1e-10, kmax=200, cleaf=32, eta=2.5)
    res = {}
    for pc in ("particle", "auto"):
        it = BEMRetIter(cp, preconditioner=pc, **opts)
        sig, _ = it.solve(exc(cp, lam))
        flag, relres, niter = it.info()
        assert flag == 0 and relres <= 1e-9, pc
        res[pc] = niter
        assert _rel(sig.sig2, direct.sig2) < (1e-5 if hmatrix else 1e-7), pc
    print("iterations:", res)
    assert res["auto"] < res["particle"] <= 40
