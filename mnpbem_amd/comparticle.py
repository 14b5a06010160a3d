"""Composite particle: particles plus the dielectric media on either side.

Reference: geometry/comparticle.py (constructor :55-127, ``eps1/eps2``
:596-652, ``inout_faces`` :682-698, ``index_func`` :752-779,
``closedparticle`` :267-346). ``inout`` rows are 1-based material indices
``[inside, outside]`` per particle.
"""
import numpy as np

from .particle import Particle


class ComParticle:
    def __init__(self, eps, particles, inout, *closed, **kwargs):
        if eps is None:
            raise ValueError("ComParticle: 'eps' must be a list of dielectric functions, got None.")
        if particles is None:
            raise ValueError("ComParticle: 'particles' must be a list of Particle, got None.")
        if inout is None:
            raise ValueError("ComParticle: 'inout' must be an inside/outside index table, got None.")
        if not hasattr(eps, "__len__"):
            raise TypeError("ComParticle: 'eps' must be a sequence (list) of dielectric functions.")
        if len(eps) == 0:
            raise ValueError("ComParticle: 'eps' must contain at least one dielectric function.")
        self.eps = list(eps)
        if isinstance(particles, Particle):
            particles = [particles]
        flat = []
        for item in particles:
            flat.extend(item if isinstance(item, (list, tuple)) else [item])
        self.p = flat
        interp = kwargs.get("interp")
        if interp is not None:
            for part in self.p:
                part.curved() if interp == "curv" else part.flat()
        io = np.atleast_2d(np.asarray(inout))
        if io.ndim != 2 or io.shape[1] != 2:
            raise ValueError("ComParticle: 'inout' must be a list of [in, out] pairs with shape (n, 2); "
                             "got shape {}.".format(io.shape))
        if io.shape[0] != len(self.p):
            raise ValueError("ComParticle: 'inout' must have one row per particle; got {} particles vs {} "
                             "inout rows.".format(len(self.p), io.shape[0]))
        if int(io.max()) > len(self.eps):
            raise ValueError("ComParticle: 'inout' references eps[{}] but only {} dielectric functions "
                             "provided.".format(int(io.max()), len(self.eps)))
        if int(io.min()) < 1:
            raise ValueError("ComParticle: 'inout' indices are 1-based; got minimum {}.".format(int(io.min())))
        self.inout = io.astype(np.int64)
        # closed surfaces: list of 1-based particle indices per particle (None = open)
        self.closed = [None] * len(self.p)
        for arg in closed:
            idx = [arg] if np.isscalar(arg) else list(arg)
            for i in idx:
                if self.closed[abs(i) - 1] is None:
                    self.closed[abs(i) - 1] = list(idx)
        self.pc = self.p[0] if len(self.p) == 1 else self.p[0].vertcat(*self.p[1:])
        # integration options (bemoptions 'rule', 'npol', 'refine'): set on the particles after
        # their geometry, so that they change the integration points of the refined entries and
        # not the element areas, as the reference does (geometry/comparticle.py:144-161, :176-186)
        if any(k in kwargs for k in ("rule", "npol", "refine")):
            from .quadrature import QuadFace
            q = QuadFace(rule=kwargs.get("rule", 18), npol=kwargs.get("npol", (7, 5)),
                         refine=kwargs.get("refine", None))
            for part in self.p + [self.pc]:
                part.quad_rule = q
        self._offsets = np.concatenate([[0], np.cumsum([q.nfaces for q in self.p])]).astype(np.int64)

    # ---------------- geometry delegated to the concatenated particle ----------------
    @property
    def pos(self):
        return self.pc.pos

    @property
    def nvec(self):
        return self.pc.nvec

    @property
    def tvec1(self):
        return self.pc.tvec1

    @property
    def tvec2(self):
        return self.pc.tvec2

    @property
    def area(self):
        return self.pc.area

    @property
    def nfaces(self):
        return int(self._offsets[-1])

    n = nfaces

    @property
    def np(self):
        return len(self.p)

    def quad(self, ind=None):
        return self.pc.quad(ind)

    quad_integration = quad

    def quadpol(self, ind=None):
        return self.pc.quadpol(ind)

    def bradius(self):
        return self.pc.bradius()

    # ---------------- media ----------------
    def index_func(self, particle_indices):
        """Face indices of the (1-based) particles."""
        ids = [particle_indices] if np.isscalar(particle_indices) else list(particle_indices)
        out = [np.arange(self._offsets[i - 1], self._offsets[i]) for i in ids]
        return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)

    @property
    def face_particle(self):
        """0-based particle index of every face."""
        return np.repeat(np.arange(len(self.p)), np.diff(self._offsets))

    @property
    def inout_faces(self):
        return self.inout[self.face_particle]

    def _eps_side(self, enei, col):
        vals = np.array([complex(np.asarray(self.eps[m - 1](enei)[0]).ravel()[0]) for m in self.inout[:, col]])
        return vals[self.face_particle]

    def eps1(self, enei):
        """Inside dielectric function of every face."""
        return self._eps_side(enei, 0)

    def eps2(self, enei):
        """Outside dielectric function of every face."""
        return self._eps_side(enei, 1)

    def closed_sets(self):
        """Per closed particle: (face indices of the particle, face indices of its
        closed surface within ``pc``, direction). When no closed argument was
        given at all, every particle is its own closed surface
        (greenfun/compgreen_ret.py:192-272; comparticle.py:267-346)."""
        closed = self.closed
        if all(c is None for c in closed):
            closed = [[i + 1] for i in range(len(self.p))]
        out = []
        for i, cl in enumerate(closed):
            if cl is None:
                continue
            if any(c <= 0 for c in cl):
                raise NotImplementedError("closed surfaces with flipped sub-particles are not provided")
            loc = np.concatenate([self.index_func(c) for c in cl])
            out.append((self.index_func(i + 1), loc, 1.0))
        return out

    def __repr__(self):
        return "ComParticle(nparticles={}, nfaces={})".format(len(self.p), self.nfaces)
