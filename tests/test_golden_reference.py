"""Golden vectors computed by the reference implementation (tests/golden/make_reference_fixtures.py):
the config-2 rod mesh and per-face surface charges/currents of the reference's BEMRet for the
Drude sphere and rod. CPU: the product's mesh and geometry and the oracle's chain against them."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import bem_oracle as O
from mnpbem_amd.mesh import trirod, trisphere
from mnpbem_amd.comparticle import ComParticle
from mnpbem_amd.materials import EpsConst, EpsDrude

DRUDE = (10.0, 9.065, 0.0708)
MESH = np.load(os.path.join(GOLDEN, "reference_rod2496_mesh.npz"))
SOL = np.load(os.path.join(GOLDEN, "reference_drude_solutions.npz"))


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b)))


def test_rod_mesh_topology_and_geometry_equal_reference():
    cp = ComParticle([EpsConst(1.77), EpsDrude(*DRUDE)], [trirod(10.0, 40.0, [25, 11, 34], triangles=True)],
                     [[2, 1]], 1, interp="curv")
    pc = cp.pc
    assert np.array_equal(np.isnan(pc.faces), np.isnan(MESH["faces"]))
    assert np.array_equal(np.nan_to_num(pc.faces, nan=-1), np.nan_to_num(MESH["faces"], nan=-1))
    assert np.array_equal(np.nan_to_num(pc.faces2, nan=-1), np.nan_to_num(MESH["faces2"], nan=-1))
    assert np.max(np.abs(pc.verts - MESH["verts"])) < 1e-12
    assert np.max(np.abs(pc.verts2 - MESH["verts2"])) < 1e-12
    assert np.max(np.abs(cp.pos - MESH["pos"])) < 1e-12
    assert np.max(np.abs(cp.nvec - MESH["nvec"])) < 1e-12
    assert np.max(np.abs(cp.area / MESH["area"] - 1)) < 1e-12
    # the oracle's own geometry from the reference's mesh
    s = O.Surface(MESH["verts"], MESH["faces"], MESH["verts2"], MESH["faces2"])
    assert np.max(np.abs(s.pos - MESH["pos"])) < 1e-12
    assert np.max(np.abs(s.nvec - MESH["nvec"])) < 1e-12
    assert np.max(np.abs(s.area / MESH["area"] - 1)) < 1e-12


def test_oracle_sphere_solution_equals_reference():
    """The oracle's chain (assembly, refinement, closed-surface correction, LU chain, solve)
    reproduces the reference's per-face charges and currents to 3e-10 (the polar self-term
    integrals sample points 1e-7 nm from the centroid and are conditioned at about 1e-10; h2 is
    the smallest field and differs by 1.4e-10 of its maximum)."""
    cp = ComParticle([EpsConst(1.0), EpsDrude(*DRUDE)], [trisphere(144, 20.0)], [[2, 1]], 1, interp="curv")
    p = cp.pc
    s = O.Surface(p.verts, p.faces, p.verts2, p.faces2)
    funcs = [lambda lam, e=e: e(lam) for e in cp.eps]
    lams = SOL["sphere_lams"]
    out = O.solve_spectrum(s, cp.face_particle, cp.inout, funcs, lams, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]),
                           return_fields=True)
    for q in range(len(lams)):
        sol = out["fields"][q]["sol"]
        for k, got in zip(("sig1", "h1", "sig2", "h2"), sol):
            assert _rel(got, SOL["sphere_" + k][q]) < 3e-10, (k, q)
    assert _rel(out["ext"], SOL["sphere_ext"]) < 1e-10
