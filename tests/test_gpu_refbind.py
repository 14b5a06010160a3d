"""The reference's LU dispatch (utils/gpu.py:145-301) bound to the library (mnpbem_amd.refbind),
called as the reference's BEMRet.init calls it (bem/bem_ret.py:438-479), against scipy."""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from mnpbem_amd import refbind as R                      # noqa: E402
from oracle import bem_oracle as O                        # noqa: E402
import mnpbem_amd as M                                    # noqa: E402


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def bem_matrices():
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsDrude(10.0, 9.065, 0.0708)], [M.trisphere(144, 20)], [[2, 1]], 1,
                       interp="curv")
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    g = O.Green(s, np.zeros(s.n, int), [[2, 1]], [(np.arange(s.n), np.arange(s.n), 1.0)])
    lam = 600.0
    ks = [e(lam)[1] for e in cp.eps]
    G1, G2, H1, H2 = O.matrices(g, ks)
    return s, G1, G2, H1, H2, cp.eps1(lam), cp.eps2(lam), 2 * np.pi / lam


def test_bemret_init_sequence_through_binding(bem_matrices):
    """lu_factor_dispatch, eye_like_lu, lu_solve_native on the identity, to_host: the inverses and
    the Sigma/Delta chain of bem_ret.py:438-479 equal scipy's to 1e-10."""
    s, G1, G2, H1, H2, e1, e2, k = bem_matrices
    n = s.n
    banner("refbind BEMRet.init sequence n={}".format(n))
    G1_lu = R.lu_factor_dispatch(G1)
    G2_lu = R.lu_factor_dispatch(G2)
    assert R.lu_backend(G1_lu) == "mi355x" and R.info(G1_lu) == 0
    G1i = R.to_host(R.lu_solve_native(G1_lu, R.eye_like_lu(G1_lu, n)))
    G2i = R.to_host(R.lu_solve_native(G2_lu, R.eye_like_lu(G2_lu, n)))
    assert isinstance(G1i, np.ndarray) and G1i.shape == (n, n)
    assert _rel(G1i, np.linalg.inv(G1)) < 1e-10
    assert _rel(G2i, np.linalg.inv(G2)) < 1e-10
    Sigma1 = H1 @ G1i
    Sigma2 = H2 @ G2i
    Delta_lu = R.lu_factor_dispatch(Sigma1 - Sigma2)
    Deltai = R.to_host(R.lu_solve_native(Delta_lu, R.eye_like_lu(Delta_lu, n)))
    assert _rel(Deltai, np.linalg.inv(Sigma1 - Sigma2)) < 1e-10
    # the rest of the reference's solve: the final factor and one solve against the oracle chain
    L1, L2 = e1[0], e2[0]
    Sigma = Sigma1 * L1 - Sigma2 * L2 + k ** 2 * (L1 - L2) * (Deltai * (s.nvec @ s.nvec.T)) * (L1 - L2)
    Sigma_lu = R.lu_factor_dispatch(Sigma)
    rhs = np.random.default_rng(0).standard_normal(n) + 0j
    x = R.lu_solve_dispatch(Sigma_lu, rhs)
    ref = sla.lu_solve(sla.lu_factor(Sigma), rhs)
    assert x.shape == (n,) and _rel(x, ref) < 1e-10


@pytest.mark.parametrize("trans", [0, 1, 2, "T"])
@pytest.mark.parametrize("ncol", [None, 3, 8, 21])
def test_lu_solve_shapes_and_trans(bem_matrices, trans, ncol):
    s, G1, G2, H1, H2, e1, e2, k = bem_matrices
    n = s.n
    banner("refbind solve trans={} ncol={}".format(trans, ncol))
    rng = np.random.default_rng(7)
    b = rng.standard_normal((n,) if ncol is None else (n, ncol)) + 1j * rng.standard_normal(
        (n,) if ncol is None else (n, ncol))
    pkg = R.lu_factor_dispatch(H2)
    lu = sla.lu_factor(H2)
    t = {"T": 1}.get(trans, trans)
    want = sla.lu_solve(lu, b, trans=t)
    got = R.lu_solve_dispatch(pkg, b, trans=trans)
    assert got.shape == b.shape and _rel(got, want) < 1e-10
    dev = R.lu_solve_native(pkg, torch.tensor(b, device="cuda"), trans=trans)
    assert isinstance(dev, torch.Tensor) and dev.is_cuda
    assert _rel(dev.cpu().numpy(), want) < 1e-10
    assert isinstance(R.lu_solve_native(pkg, b, trans=trans), np.ndarray)


def test_zero_pivot_is_reported_not_raised():
    A = np.eye(40, dtype=complex)
    A[:, 7] = 0
    banner("refbind zero pivot")
    pkg = R.lu_factor_dispatch(A)
    assert R.info(pkg) == 8
