"""Writes golden vectors from the reference implementation (run in the build container only,
where the reference's sources are at /root/reference; nothing of the reference is stored, only
the arrays it computes):

  reference_rod2496_mesh.npz   trirod(10, 40, [25, 11, 34], triangles=True) with curved
                               interpolation: verts, faces, verts2, faces2 (mesh topology),
                               pos, nvec, area per face (mesh_generators.py:592-660,
                               particle.py:551-613)
  reference_drude_solutions.npz  sig1, h1, sig2, h2 and extinction of the reference's BEMRet
                               (bem_ret.py:293-525, :1569-1831) for
                               - the gold-Drude sphere trisphere(144, 20) in vacuum, x-polarised
                                 plane wave along z, at 450 and 612.5 nm;
                               - the gold-Drude rod above in eps 1.77, plane wave polarised along
                                 the axis travelling along x, at 600 and 800 nm.
  reference_coreshell.npz      sig1, h1, sig2, h2 and extinction of the reference's BEMRet for
                               a nested particle: a gold-Drude core trisphere(144, 12) inside a
                               shell trisphere(144, 20) of eps 2.1 in vacuum (inout [[2, 1],
                               [3, 2]], each surface closed), x-polarised plane wave along z, at
                               480 and 530 nm; and the same with the shell's eps 1 (a shell that
                               light cannot see); of the eps 2.1 case at 480 nm also the
                               diagonals of the reference's G1, G2, H1, H2 (bem_ret.py:407-433)
                               and their rows 0 (a shell face) and 400 (a core face).
Drude parameters (10, 9.065 eV, 0.0708 eV).

Usage: python tests/golden/make_reference_fixtures.py [/root/reference] [--only coreshell]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ARGS = [a for a in sys.argv[1:] if not a.startswith("--")]
ONLY = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
ARGS = [a for a in ARGS if a != ONLY]
REF = ARGS[0] if ARGS else "/root/reference"
sys.path.insert(0, REF)

import mnpbem as R                                            # noqa: E402
from mnpbem.geometry.mesh_generators import trirod, trisphere  # noqa: E402
from mnpbem.greenfun import CompGreenRet  # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)


def _solve(cp, pol, direc, lams):
    out = {k: [] for k in ("sig1", "h1", "sig2", "h2", "ext")}
    exc = R.PlaneWaveRet(np.array([pol], float), np.array([direc], float))
    bem = R.BEMRet(cp)
    for lam in lams:
        sig, _ = bem.solve(exc.potential(cp, lam))
        for k in ("sig1", "h1", "sig2", "h2"):
            out[k].append(np.asarray(getattr(sig, k)).reshape(cp.nfaces, -1).squeeze())
        out["ext"].append(np.atleast_1d(exc.extinction(sig))[0])
    return {k: np.array(v) for k, v in out.items()}


def coreshell():
    out = {}
    for tag, eps_shell in (("", 2.1), ("clear_", 1.0)):
        eps = [R.EpsConst(1.0), R.EpsConst(eps_shell), R.EpsDrude(*DRUDE)]
        cp = R.ComParticle(eps, [trisphere(144, 20.0), trisphere(144, 12.0)], [[2, 1], [3, 2]], 1, 2,
                           interp="curv")
        sol = _solve(cp, [1.0, 0, 0], [0, 0, 1.0], [480.0, 530.0])
        print("core-shell eps_shell", eps_shell, "ext", sol["ext"], flush=True)
        out.update({tag + k: v for k, v in sol.items()})
        if not tag:
            g = CompGreenRet(cp, cp)
            ev = lambda i, j, key: np.asarray(g.eval(i, j, key, 480.0))
            sub = lambda a, b: a if np.isscalar(b) or np.ndim(b) == 0 else a - b
            mats = {"G1": sub(ev(0, 0, "G"), ev(1, 0, "G")), "G2": sub(ev(1, 1, "G"), ev(0, 1, "G")),
                    "H1": sub(ev(0, 0, "H1"), ev(1, 0, "H1")), "H2": sub(ev(1, 1, "H2"), ev(0, 1, "H2"))}
            for k, v in mats.items():
                out["mat_diag_" + k] = np.diag(v).copy()
                out["mat_rows_" + k] = v[[0, 400]].copy()
    np.savez_compressed(os.path.join(HERE, "reference_coreshell.npz"), lams=np.array([480.0, 530.0]), **out)


def main():
    if ONLY == "coreshell":
        coreshell()
        return
    rod = trirod(10.0, 40.0, [25, 11, 34], triangles=True)
    cp_rod = R.ComParticle([R.EpsConst(1.77), R.EpsDrude(*DRUDE)], [rod], [[2, 1]], 1, interp="curv")
    pc = cp_rod.pc
    np.savez_compressed(os.path.join(HERE, "reference_rod2496_mesh.npz"), verts=pc.verts, faces=pc.faces,
                        verts2=pc.verts2, faces2=pc.faces2, pos=pc.pos, nvec=pc.nvec, area=pc.area)
    print("mesh", pc.nfaces, flush=True)
    sph = trisphere(144, 20.0)
    cp_sph = R.ComParticle([R.EpsConst(1.0), R.EpsDrude(*DRUDE)], [sph], [[2, 1]], 1, interp="curv")
    s1 = _solve(cp_sph, [1.0, 0, 0], [0, 0, 1.0], [450.0, 612.5])
    print("sphere ext", s1["ext"], flush=True)
    s2 = _solve(cp_rod, [0, 0, 1.0], [1.0, 0, 0], [600.0, 800.0])
    print("rod ext", s2["ext"], flush=True)
    np.savez_compressed(os.path.join(HERE, "reference_drude_solutions.npz"),
                        sphere_lams=np.array([450.0, 612.5]), rod_lams=np.array([600.0, 800.0]),
                        **{"sphere_" + k: v for k, v in s1.items()}, **{"rod_" + k: v for k, v in s2.items()})


if __name__ == "__main__":
    main()
