"""BEM path parity on the GPU against the CPU oracle and the reference's MATLAB tables."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import bem_oracle as O                                       # noqa: E402
import mnpbem_amd as M                                                   # noqa: E402
from mnpbem_amd.engine import GreenBatch, RetSolverBatch                 # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)
GOLD = np.loadtxt(os.path.join(ROOT, "mnpbem_amd", "data", "gold.dat"), comments="%")
PINF = np.load(os.path.join(ROOT, "mnpbem_amd", "data", "pinfty256.npz"))


def _setup(parts, inout, eps):
    cp = M.ComParticle(eps, parts, inout, *range(1, len(parts) + 1), interp="curv")
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    fp = cp.face_particle
    closed = [(np.nonzero(fp == a)[0], np.nonzero(fp == a)[0], 1.0) for a in range(len(parts))]
    return cp, s, fp, closed


@pytest.fixture(scope="module")
def sphere():
    return _setup([M.trisphere(144, 20)], [[2, 1]], [M.EpsConst(1.0), M.EpsDrude(*DRUDE)])


@pytest.fixture(scope="module")
def sphere_oracle(sphere):
    cp, s, fp, closed = sphere
    return O.Green(s, fp, [[2, 1]], closed)


def _oracle_eps(cp):
    funcs = []
    for e in cp.eps:
        funcs.append(lambda lam, e=e: e(lam))
    return funcs


def test_assembly_matches_oracle(sphere, sphere_oracle):
    cp, s, fp, closed = sphere
    bem = M.BEMRet(cp)
    geo = bem.g.geo
    lams = np.array([450.0, 612.5])
    banner("assembly sphere n={} nlam=2".format(s.n))
    gb = GreenBatch(geo, 2, torch.complex128)
    gb.assemble(geo.media(lams)[0], 2)
    for b, lam in enumerate(lams):
        ks = [e(lam)[1] for e in cp.eps]
        ref = O.matrices(sphere_oracle, ks)
        for t, R in zip((gb.G1, gb.G2, gb.H1, gb.H2), ref):
            got = GreenBatch.to_numpy(t, b)
            assert np.max(np.abs(got - R)) / np.max(np.abs(R)) < 1e-9
            # far entries are bit-for-bit formulas: compare tighter away from refined pairs
            far = sphere_oracle.kind == 0
            assert np.max(np.abs(got[far] - R[far])) / np.max(np.abs(R[far])) < 1e-13


def test_assembly_dimer_blocks():
    a = M.trisphere(60, 10)
    b = M.trisphere(60, 10).shift([0, 0, 11.0])
    cp, s, fp, closed = _setup([a, b], [[2, 1], [2, 1]], [M.EpsConst(1.33), M.EpsDrude(*DRUDE)])
    banner("assembly dimer n={}".format(s.n))
    g = O.Green(s, fp, [[2, 1], [2, 1]], closed)
    bem = M.BEMRet(cp)
    gb = GreenBatch(bem.g.geo, 1, torch.complex128)
    gb.assemble(bem.g.geo.media([700.0])[0], 1)
    ks = [e(700.0)[1] for e in cp.eps]
    for t, R in zip((gb.G1, gb.G2, gb.H1, gb.H2), O.matrices(g, ks)):
        got = GreenBatch.to_numpy(t, 0)
        assert np.max(np.abs(got - R)) / np.max(np.abs(R)) < 1e-9


def test_bemret_solution_matches_oracle(sphere, sphere_oracle):
    """sig1, h1, sig2, h2 of the complex128 chain vs the oracle's LAPACK chain (1e-10)."""
    cp, s, fp, closed = sphere
    lam = 520.0
    banner("bemret fp64 solution sphere lam={}".format(lam))
    bem = M.BEMRet(cp)
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    sig, _ = bem.solve(exc.potential(cp, lam))
    ks = [e(lam)[1] for e in cp.eps]
    G1, G2, H1, H2 = O.matrices(sphere_oracle, ks)
    e1, e2 = cp.eps1(lam), cp.eps2(lam)
    ob = O.BEMRetOracle(G1, G2, H1, H2, s.nvec, e1, e2, 2 * np.pi / lam)
    pot = exc.potential(cp, lam)
    phi, a, alpha, De = O.excitation(s.nvec, 2 * np.pi / lam, e1, e2, pot.a1, pot.a1p, pot.a2, pot.a2p)
    ref = ob.solve(phi, a, alpha, De)
    for got, want in zip((sig.sig1, sig.h1, sig.sig2, sig.h2), ref):
        assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-10
    # cross sections through the device far field
    pinf_ok = exc.spec.ndir == 508
    assert pinf_ok
    ext = exc.extinction(sig)
    k = ks[0]
    ext_ref = O.extinction(np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), lam, k, 1.0, s.pos, s.area, ref,
                           np.zeros(0, int), np.arange(s.n))
    assert abs(ext - ext_ref) / abs(ext_ref) < 1e-10
    sca, _ = exc.scattering(sig)
    sca_ref = O.scattering(lam, k, 1.0, s.pos, s.area, ref, np.zeros(0, int), np.arange(s.n), PINF["nvec"],
                           PINF["area"])
    assert abs(sca - sca_ref) / abs(sca_ref) < 1e-10


def test_gold_sphere_matches_matlab_table():
    """validation/03_bemret/sphere (trisphere(144, 20), gold): GPU sweep vs MATLAB (1e-9)."""
    ml = np.loadtxt(os.path.join(GOLDEN, "bemret_sphere_matlab.csv"), delimiter=",", skiprows=1)
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsTable("gold.dat")], [M.trisphere(144, 20)], [[2, 1]], 1,
                       interp="curv")
    banner("gold sphere sweep 41 wavelengths batch 8")
    bem = M.BEMRet(cp, batch=8)
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    out = bem.solve_spectrum(exc, ml[:, 0], scattering=True)
    ext, sca = out["ext"][:, 0], out["sca"][:, 0]
    assert np.max(np.abs(ext - ml[:, 1]) / ml[:, 1]) < 1e-9
    assert np.max(np.abs(sca - ml[:, 2]) / ml[:, 2]) < 1e-9
    assert np.max(np.abs(ext - sca - ml[:, 3]) / ml[:, 3]) < 1e-9


def test_drude_sphere_sweep_vs_oracle(sphere):
    """Config 1 (Drude sphere, 41 wavelengths): device complex128 vs oracle (1e-10)."""
    cp, s, fp, closed = sphere
    lams = np.linspace(400, 800, 41)
    sel = [0, 13, 27, 40]
    banner("drude sphere sweep")
    bem = M.BEMRet(cp, batch=16)
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    ext = bem.solve_spectrum(exc, lams)["ext"][:, 0]
    ref = O.solve_spectrum(s, fp, [[2, 1]], _oracle_eps(cp), lams[sel], np.array([1.0, 0, 0]),
                           np.array([0, 0, 1.0]), closed_sets=closed)["ext"]
    assert np.max(np.abs(ext[sel] - ref) / np.abs(ref)) < 1e-10
    assert np.all(np.isfinite(ext))


@pytest.mark.parametrize("split", [0, 2, 3])
def test_mixed_precision_refinement(sphere, split):
    """complex64 factorisation (GEMMs on the FP32 matrix cores, or on the BF16 ones with 2 or 3
    operand terms) + FP64 refinement reaches the complex128 answer (1e-8 on ext)."""
    cp, s, fp, closed = sphere
    lams = np.linspace(400, 800, 9)
    banner("mixed precision sphere gemm_split={}".format(split))
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    ref = M.BEMRet(cp, batch=9).solve_spectrum(exc, lams)["ext"][:, 0]
    out = M.BEMRet(cp, batch=9, precision="mixed", gemm_split=split, autotune=False).solve_spectrum(exc, lams)
    ext = out["ext"][:, 0]
    assert np.max(np.abs(ext - ref) / np.abs(ref)) < 1e-8
    assert np.max(out["relres"][0]) <= 1e-12


def test_rod_quads_matches_matlab_table():
    """validation/03_bemret/rod (curved quads), both polarisations at 3 wavelengths."""
    ml = np.loadtxt(os.path.join(GOLDEN, "bemret_rod_matlab.csv"), delimiter=",", skiprows=1)
    sel = [0, 20, 40]
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsTable("gold.dat")], [M.trirod(10.0, 40.0, [15, 15, 15])], [[2, 1]],
                       1, interp="curv")
    banner("gold rod n={}".format(cp.nfaces))
    bem = M.BEMRet(cp, batch=3)
    exc = M.PlaneWaveRet([[1.0, 0, 0], [0, 0, 1.0]], [[0, 0, 1.0], [1.0, 0, 0]])
    out = bem.solve_spectrum(exc, ml[sel, 0], scattering=True)
    for ip, (ce, cs) in enumerate(((1, 2), (3, 4))):
        assert np.max(np.abs(out["ext"][:, ip] - ml[sel, ce]) / ml[sel, ce]) < 1e-9
        assert np.max(np.abs(out["sca"][:, ip] - ml[sel, cs]) / ml[sel, cs]) < 1e-9


def test_bemstat_matches_matlab_table():
    """validation/02_bemstat/sphere (quasistatic, gold): device LU vs MATLAB (1e-9), and the
    k = 0 device assembly of F vs the oracle."""
    ml = np.loadtxt(os.path.join(GOLDEN, "bemstat_sphere_matlab.csv"), delimiter=",", skiprows=1)
    cp = M.ComParticle([M.EpsConst(1.0), M.EpsTable("gold.dat")], [M.trisphere(144, 20)], [[2, 1]], 1,
                       interp="curv")
    banner("bemstat gold sphere")
    bem = M.BEMStat(cp)
    exc = M.PlaneWaveStat([1.0, 0, 0])
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    kind = O.refinement_kind(s)
    corr, F = O.closed_correction(s, kind, [(np.arange(s.n), np.arange(s.n), 1.0)])
    F = F + np.diag(corr)
    Fd = bem.F
    assert np.max(np.abs(Fd - F)) / np.max(np.abs(F)) < 1e-9
    Gd = bem.g.G
    Go = O.stat_G(s, kind)
    assert np.max(np.abs(Gd - Go)) / np.max(np.abs(Go)) < 1e-9
    for row in ml[::5]:
        sig, _ = bem.solve(exc(cp, row[0]))
        assert abs(exc.extinction(sig) - row[1]) / row[1] < 1e-9
        assert abs(exc.scattering(sig) - row[2]) / row[2] < 1e-9
