"""H-matrices on the GPU: dense blocks, device ACA, products, and BEMRetIter(hmatrix=True)
(greenfun/hmatrix.py:159-541, aca_compgreen_ret.py:20-208, bem_ret_iter.py:488-580)."""
import numpy as np
import pytest

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import mnpbem_amd as M                                   # noqa: E402
from mnpbem_amd import engine                            # noqa: E402
from mnpbem_amd.bem_iter import BEMRetIter              # noqa: E402
from mnpbem_amd.hmatrix import ClusterTree, HMatrixRet  # noqa: E402
from oracle.aca_oracle import aca                        # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)
LAM = 600.0
HTOL = 1e-6


@pytest.fixture(scope="module")
def built():
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsDrude(*DRUDE)], [M.trisphere(784, 20)], [[2, 1]], 1)
    g = M.CompGreenRet(cp, cp)
    geo = g.geo
    banner("hmatrix build sphere n={} lam={} htol={}".format(cp.nfaces, LAM, HTOL))
    tree = ClusterTree(cp.pos, 32, cp.face_particle)
    hm = HMatrixRet(geo, tree, LAM, htol=HTOL, kmax=100)
    dense = engine.GreenBatch(geo, 1, torch.complex128)
    kmat = geo.media([LAM])[0]
    dense.assemble(kmat, 1)
    torch.cuda.synchronize()
    A = [engine.GreenBatch.to_numpy(t, 0) for t in (dense.G1, dense.G2, dense.H1, dense.H2)]
    return cp, tree, hm, A


def test_dense_blocks_and_aca(built):
    cp, tree, hm, A = built
    assert hm.nlr > 20 and hm.nd > 20
    c2p = tree.ind[:, 0]
    for q in range(4):
        H = hm.full(q)
        Ac = A[q][np.ix_(c2p, c2p)]
        Hc = H[np.ix_(c2p, c2p)]
        scale = np.abs(Ac).max()
        for r0, c0, m, n in hm._dense:
            blk = (slice(r0, r0 + m), slice(c0, c0 + n))
            assert np.abs(Hc[blk] - Ac[blk]).max() <= 1e-13 * scale
        errs, same, diff = [], 0, []
        for b, (r0, c0, m, n) in enumerate(hm.lowrank_blocks()):
            blk = Ac[r0:r0 + m, c0:c0 + n]
            errs.append(np.linalg.norm(Hc[r0:r0 + m, c0:c0 + n] - blk) / np.linalg.norm(blk))
            _, _, k = aca(blk, HTOL, 100, block=b)
            same += int(k == hm.ranks[b, q])
            diff.append(abs(k - hm.ranks[b, q]))
        errs = np.array(errs)
        print("kind {} median {:.2e} max {:.2e} ranks equal {}/{}".format(q, np.median(errs), errs.max(), same,
                                                                          len(errs)))
        assert np.median(errs) < 3 * HTOL
        assert errs.max() < 100 * HTOL
        assert same >= 0.9 * len(errs) and max(diff) <= 3


def test_matvec_matches_dense(built):
    cp, tree, hm, A = built
    n = cp.nfaces
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, 8)) + 1j * rng.standard_normal((n, 8))
    xt = torch.tensor(x, device=hm.geo.dev)
    for q in range(4):
        y = hm.matvec(q, xt).cpu().numpy()
        ref = A[q] @ x
        assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 2e-5


def test_operator_matches_dense_operator(built):
    cp, tree, hm, A = built
    n = cp.nfaces
    banner("hmatrix 8N operator vs dense")
    itd = BEMRetIter(cp, enei=LAM)
    ith = BEMRetIter(cp, enei=LAM, hmatrix=True, htol=HTOL, cleaf=32, preconditioner="none")
    rng = np.random.default_rng(5)
    x = torch.tensor(rng.standard_normal(8 * n) + 1j * rng.standard_normal(8 * n), device=hm.geo.dev)
    yd, yh = torch.empty_like(x), torch.empty_like(x)
    itd._apply_A(x, yd)
    ith._apply_A(x, yh)
    yd, yh = yd.cpu().numpy(), yh.cpu().numpy()
    assert np.linalg.norm(yh - yd) / np.linalg.norm(yd) < 2e-5
    assert ith._hm.stats["compression"][0] < 0.8


def _rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_iterative_hmatrix_equals_direct_small_sphere():
    """mnpbem/tests/test_hmatrix_iter.py:39-63 of the reference (there: < 1e-3)."""
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsConst(-10 + 0.5j)], [M.trisphere(144, 10)], [[2, 1]], 1)
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    banner("bemretiter hmatrix sphere144 tol 1e-8")
    direct, _ = M.BEMRet(cp).solve(exc(cp, 600.0))
    it = BEMRetIter(cp, tol=1e-8, maxit=200, hmatrix=True, htol=1e-8, kmax=[4, 100], cleaf=32)
    sig, _ = it.solve(exc(cp, 600.0))
    flag, relres, niter = it.info()
    assert flag == 0 and relres <= 1e-8
    for f in ("sig1", "sig2", "h1", "h2"):
        assert _rel2(getattr(sig, f), getattr(direct, f)) < 1e-5


def test_iterative_hmatrix_dimer_vs_direct():
    p1 = M.trisphere(400, 20).shift([0, 0, -11.0])
    p2 = M.trisphere(400, 20).shift([0, 0, 11.0])
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsDrude(*DRUDE)], [p1, p2], [[2, 1], [2, 1]], 1, 2)
    exc = M.PlaneWaveRet([0, 0, 1.0], [1.0, 0, 0])
    banner("bemretiter hmatrix dimer n={}".format(cp.nfaces))
    direct, _ = M.BEMRet(cp).solve(exc(cp, 550.0))
    for pc in ("auto", "none"):
        it = BEMRetIter(cp, tol=1e-8, maxit=400, restart=60, hmatrix=True, htol=1e-6, cleaf=32, preconditioner=pc)
        sig, _ = it.solve(exc(cp, 550.0))
        flag, relres, niter = it.info()
        print(pc, flag, relres, niter, it._hm.stats)
        assert flag == 0 and relres <= 1e-8
        assert _rel2(sig.sig2, direct.sig2) < 1e-4
        ext_d, ext_i = exc.extinction(direct), exc.extinction(sig)
        assert abs(ext_i - ext_d) / abs(ext_d) < 1e-5
