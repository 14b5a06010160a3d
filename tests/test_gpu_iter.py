"""BEMRetIter on the GPU: device GMRES over the exact 8N operator (bem_ret_iter.py:1362-1744)."""
import numpy as np
import pytest

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import mnpbem_amd as M                                   # noqa: E402
from mnpbem_amd.bem_iter import BEMRetIter              # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)


@pytest.fixture(scope="module")
def sphere():
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsDrude(*DRUDE)], [M.trisphere(144, 20)], [[2, 1]], 1, interp="curv")
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    return cp, exc


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_iterative_equals_direct_with_bemret_preconditioner(sphere):
    """tests/test_bem_iter.py:88-135 of the reference: iterative vs direct < 1e-8."""
    cp, exc = sphere
    lam = 600.0
    banner("bemretiter precond=bemret lam={}".format(lam))
    direct, _ = M.BEMRet(cp).solve(exc(cp, lam))
    it = BEMRetIter(cp, tol=1e-11, maxit=50, restart=20)
    sig, _ = it.solve(exc(cp, lam))
    flag, relres, niter = it.info()
    assert flag == 0 and relres <= 1e-11 and niter <= 10
    for f in ("sig1", "h1", "sig2", "h2"):
        assert _rel(getattr(sig, f), getattr(direct, f)) < 1e-8


def test_unpreconditioned_gmres_converges(sphere):
    cp, exc = sphere
    lam = 450.0
    banner("bemretiter precond=None")
    direct, _ = M.BEMRet(cp).solve(exc(cp, lam))
    it = BEMRetIter(cp, precond=None, tol=1e-8, maxit=600, restart=50)
    sig, _ = it.solve(exc(cp, lam))
    flag, relres, niter = it.info()
    assert flag == 0 and relres <= 1e-8
    assert niter > 5
    assert _rel(sig.sig1, direct.sig1) < 1e-5
    ext_d, ext_i = exc.extinction(direct), exc.extinction(sig)
    assert abs(ext_i - ext_d) / abs(ext_d) < 1e-6


def test_stall_is_reported_not_raised(sphere):
    cp, exc = sphere
    banner("bemretiter maxit stall")
    it = BEMRetIter(cp, precond=None, tol=1e-12, maxit=3, restart=3)
    sig, _ = it.solve(exc(cp, 700.0))
    flag, relres, niter = it.info()
    assert flag == 1 and relres > 1e-12 and niter == 3
    assert np.all(np.isfinite(sig.sig1))


def test_warm_start_from_solution(sphere):
    cp, exc = sphere
    lam = 650.0
    banner("bemretiter warm start")
    direct, _ = M.BEMRet(cp).solve(exc(cp, lam))
    it = BEMRetIter(cp, precond=None, tol=1e-9, maxit=100, restart=30)
    sig, _ = it.solve(exc(cp, lam), x0=direct)
    flag, relres, niter = it.info()
    assert flag == 0 and niter == 0


@pytest.mark.parametrize("solver", ["bicgstab", "cgs"])
def test_bicgstab_and_cgs_equal_direct(sphere, solver):
    """solver='bicgstab' / 'cgs' (bem_iter.py:60-68): with the BEMRet preconditioner they reach
    the direct solution (reference test_bem_iter.py:75-85 asks 1e-6); unpreconditioned BiCGSTAB
    converges on the same system."""
    cp, exc = sphere
    lam = 600.0
    banner("bemretiter solver={}".format(solver))
    direct, _ = M.BEMRet(cp).solve(exc(cp, lam))
    it = BEMRetIter(cp, solver=solver, tol=1e-10, maxit=100)
    sig, _ = it.solve(exc(cp, lam))
    flag, relres, niter = it.info()
    assert flag == 0 and relres <= 1e-10 and niter <= 20
    for f in ("sig1", "h1", "sig2", "h2"):
        assert _rel(getattr(sig, f), getattr(direct, f)) < 1e-7
    if solver == "bicgstab":
        un = BEMRetIter(cp, solver=solver, precond=None, tol=1e-8, maxit=2000)
        s2, _ = un.solve(exc(cp, lam))
        assert un.info()[0] == 0
        assert _rel(s2.sig2, direct.sig2) < 1e-5
