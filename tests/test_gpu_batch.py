"""Particles of uneven face counts solved as one padded batch (mnpbem_amd.batch) against the
per-particle sweep and the oracle."""
import numpy as np
import pytest

from conftest import banner

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import bem_oracle as O                                       # noqa: E402
import mnpbem_amd as M                                                   # noqa: E402
from mnpbem_amd import batch as MB                                       # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)


def _particles():
    import bench
    eps = [M.EpsConst(1.33), M.EpsDrude(*DRUDE)]
    s1 = bench.sphere_mesh(320, 20.0)
    e1 = bench.sphere_mesh(456, 2.0).scale([8, 8, 20])
    d1 = bench.sphere_mesh(200, 20.0).shift([0, 0, -11.5])
    d2 = bench.sphere_mesh(200, 20.0).shift([0, 0, 11.5])
    return [M.ComParticle(eps, [s1], [[2, 1]], 1, interp="curv"),
            M.ComParticle(eps, [e1], [[2, 1]], 1, interp="curv"),
            M.ComParticle(eps, [d1, d2], [[2, 1], [2, 1]], 1, 2, interp="curv")]


def test_plan_groups_bounds_waste_and_memory():
    nf = [300, 310, 2900, 1000, 1050, 2950, 640]
    groups = MB.plan_groups(nf, 32, max_waste=0.25, mem_bytes=1e12)
    assert sorted(i for g in groups for i in g) == list(range(len(nf)))
    for g in groups:
        nmax = max(nf[i] for i in g)
        assert len(g) * nmax ** 3 <= 1.25 * sum(nf[i] ** 3 for i in g)
    small = MB.plan_groups(nf, 32, mem_bytes=32 * 120 * 3000 ** 2)
    assert all(len(g) == 1 for g in small if max(nf[i] for i in g) > 2000)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_group_equals_per_particle_sweeps(precision):
    cps = _particles()
    lams = np.linspace(450, 900, 6)
    exc = M.PlaneWaveRet([1.0, 0, 0], [0, 0, 1.0])
    banner("particle group {} faces {}".format(precision, [cp.nfaces for cp in cps]))
    timings = []
    # one group of all three (waste bound off), padded to the largest particle
    out = MB.solve_particles(cps, lams, [1.0, 0, 0], [0, 0, 1.0], precision=precision, max_waste=100.0,
                             autotune=False, timings=timings)
    assert len(timings) == 1 and sorted(timings[0][0]) == [0, 1, 2]
    assert not out["fallback"]
    for i, cp in enumerate(cps):
        ref = M.BEMRet(cp, batch=6, precision=precision, autotune=False).solve_spectrum(exc, lams)["ext"][:, 0]
        # mixed: both sides are refined to |r| / |b| <= 1e-12 from different complex64 factors
        # (the padded group's GEMM tiles differ), which leaves ~1e-10 in the smallest cross sections
        tol = 1e-12 if precision == "double" else 1e-9
        assert np.max(np.abs(out["ext"][i] - ref) / np.abs(ref)) < tol


def test_group_matches_oracle():
    cps = _particles()
    lams = np.array([520.0, 700.0])
    banner("particle group vs oracle")
    out = MB.solve_particles(cps, lams, [1.0, 0, 0], [0, 0, 1.0], precision="mixed", max_waste=100.0,
                             autotune=False)
    for i in (0, 2):
        cp = cps[i]
        p = cp.pc
        s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
        funcs = [lambda lam, e=e: e(lam) for e in cp.eps]
        ref = O.solve_spectrum(s, cp.face_particle, cp.inout, funcs, lams, np.array([1.0, 0, 0]),
                               np.array([0, 0, 1.0]))["ext"]
        assert np.max(np.abs(out["ext"][i] - ref) / np.abs(ref)) < 1e-8


def test_graphed_group_equals_eager_and_replays_consistently():
    """The captured-graph pipeline of a group against the eager one, and two runs of the same
    graphs with eager work in between give the same result (every replay checked)."""
    cps = _particles()
    lams = np.linspace(450, 900, 4)
    pol, direc = [1.0, 0, 0], [0, 0, 1.0]
    banner("graphed particle group")
    eager = MB.solve_particles(cps, lams, pol, direc, precision="mixed", max_waste=100.0, autotune=False)
    graphed = MB.solve_particles(cps, lams, pol, direc, precision="mixed", max_waste=100.0, autotune=False,
                                 graphs=True)
    assert not graphed["fallback"]
    # the graphed pipeline replays refinement steps until every entry is done, the eager one
    # stops each entry on its own: both end at relres <= 1e-12, a step apart at most, which
    # the conditioning of these spectra turns into up to ~1e-10 in the extinction
    assert np.max(np.abs(graphed["ext"] - eager["ext"]) / np.abs(eager["ext"])) < 1e-9
    greens = [M.CompGreenRet(cp, cp) for cp in cps]
    geos = [g.geo for g in greens]
    n = max(cp.nfaces for cp in cps)
    gg = MB.GroupGeometry(geos, len(lams), n, geos[0].dev)
    s = MB.GroupSolver(gg, "mixed", 32, tuning=dict(gemm_split=3))
    gr = MB.GraphedGroup(s, lams, pol, direc, 1)
    x1 = gr.run().clone()
    it1 = s.stats["iterations"]
    # eager work on the same buffers between the replays
    s.factor(lams)
    _ = s.solve(gr.rhs)
    torch.cuda.synchronize()
    x2 = gr.run().clone()
    assert s.stats["iterations"] == it1
    assert torch.equal(x1, x2)
