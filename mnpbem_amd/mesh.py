"""Mesh generators: spheres, capped rods, ellipsoids and rod/sphere dimers.

Reference: geometry/mesh_generators.py (``trisphere`` :13-130,
``trispheresegment`` :551-584, ``fvgrid`` :519-547, ``trirod`` :592-636).
Sphere point sets are the toolbox's tabulated, energy-minimised sets
(``data/sphere{n}.npz``).
"""
import os

import numpy as np

from .particle import Particle, add_flat_midpoints

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
_SPHERE_SIZES = np.array([32, 60, 144, 256, 400, 484, 576, 676, 784, 900, 1024, 1225, 1444])


def trisphere(n, diameter=1.0, **kwargs):
    """Curved triangulated sphere from the tabulated point set closest to ``n``
    vertices; edge mid-points are projected onto the sphere."""
    if n is None or int(n) <= 0:
        raise ValueError("trisphere: 'n' must be a positive integer, got {!r}.".format(n))
    if not np.isfinite(diameter) or diameter <= 0:
        raise ValueError("trisphere: 'diameter' must be a positive finite float, got {!r}.".format(diameter))
    nn = int(_SPHERE_SIZES[np.argmin(np.abs(_SPHERE_SIZES - int(n)))])
    dat = np.load(os.path.join(_DATA, "sphere{}.npz".format(nn)))
    base = Particle(0.5 * diameter * dat["verts"], dat["faces"].astype(float), norm=False)
    add_flat_midpoints(base)
    v2 = base.verts2
    v2 = 0.5 * diameter * v2 / np.linalg.norm(v2, axis=1, keepdims=True)
    kwargs.setdefault("interp", "curv")
    return Particle(v2, base.faces2, **kwargs)


def _surf2patch(x, y, z, triangles=False):
    """Faces/vertices of a parametric grid (m, n), vertices in column-major order."""
    m, n = x.shape
    verts = np.column_stack([x.ravel(order="F"), y.ravel(order="F"), z.ravel(order="F")])
    jj, ii = np.meshgrid(np.arange(n - 1), np.arange(m - 1), indexing="ij")
    v00 = (jj * m + ii).ravel()
    v10 = v00 + 1
    v01 = v00 + m
    v11 = v01 + 1
    if triangles:
        faces = np.vstack([np.stack([v00, v01, v11], 1), np.stack([v00, v11, v10], 1)])
    else:
        faces = np.stack([v00, v01, v11, v10], 1)
    return faces, verts


def fvgrid(x, y, triangles=False):
    """Planar grid with flat mid-points: returns (verts2, faces2)."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    if x.ndim == 1 and y.ndim == 1:
        x, y = np.meshgrid(x, y)
    faces, verts = _surf2patch(x, y, np.zeros_like(x), triangles)
    p = Particle(verts, faces[:, ::-1].astype(float), norm=False)
    add_flat_midpoints(p)
    return p.verts2, p.faces2


def trispheresegment(phi, theta, diameter=1.0, triangles=False, **kwargs):
    """Curved sphere segment over the (phi, theta) grid."""
    P, T = np.meshgrid(np.asarray(phi, float), np.asarray(theta, float))
    r = 0.5 * diameter
    faces, verts = _surf2patch(r * np.sin(T) * np.cos(P), r * np.sin(T) * np.sin(P), r * np.cos(T), triangles)
    p = Particle(verts, faces.astype(float)).clean()
    add_flat_midpoints(p)
    v2 = p.verts2
    v2 = r * v2 / np.maximum(np.linalg.norm(v2, axis=1, keepdims=True), 1e-30)
    return Particle(v2, p.faces2, **kwargs)


def trirod(diameter, height, n=None, triangles=False, **kwargs):
    """Cylinder of ``diameter`` with hemispherical caps, total length ``height``
    along z. ``n = [nphi, ntheta, nz]`` grid points (default [15, 20, 20])."""
    if n is None:
        n = [15, 20, 20]
    if len(n) != 3:
        raise ValueError("trirod: n must have 3 elements [nphi, ntheta, nz]")
    nphi, ntheta, nz = (int(v) for v in n)
    phi = np.linspace(0, 2 * np.pi, nphi)
    theta = np.linspace(0, 0.5 * np.pi, ntheta)
    cap1 = trispheresegment(phi, theta, diameter, triangles=triangles, **kwargs)
    cap1.shift([0, 0, 0.5 * (height - diameter)])
    cap2 = cap1.flip(2)
    zc = 0.5 * np.linspace(-1, 1, nz) * (height - diameter)
    v2, f2 = fvgrid(phi, zc, triangles=triangles)
    cyl_verts = np.column_stack([0.5 * diameter * np.cos(v2[:, 0]), 0.5 * diameter * np.sin(v2[:, 0]), v2[:, 1]])
    cyl = Particle(cyl_verts, f2, **kwargs)
    return (cap1 + cap2 + cyl).clean()


def triellipsoid(n, axes, **kwargs):
    """Curved ellipsoid with semi-axes ``axes``: a tabulated sphere of unit
    radius scaled along x, y, z (mid-points scaled with it)."""
    p = trisphere(n, 2.0, **kwargs)
    return p.scale(np.asarray(axes, float))


def rod_for_faces(diameter, length, target_faces, triangles=True):
    """Capped rod whose face count is close to ``target_faces``: grid points
    nphi = nz and ntheta grow together; returns (particle, n)."""
    best = None
    for nphi in range(8, 200):
        ntheta = max(4, int(round(nphi * 0.5 * np.pi * diameter / (np.pi * diameter) * 0.9)) + 1)
        nz = max(3, int(round((length - diameter) / (np.pi * diameter / (nphi - 1)))) + 1)
        nf = 2 * (nphi - 1) * ((ntheta - 1) * 2 - 1) + 2 * (nphi - 1) * (nz - 1) if triangles else 0
        if best is None or abs(nf - target_faces) < abs(best[0] - target_faces):
            best = (nf, [nphi, ntheta, nz])
        if nf > 1.5 * target_faces:
            break
    p = trirod(diameter, length, best[1], triangles=triangles)
    return p, best[1]
