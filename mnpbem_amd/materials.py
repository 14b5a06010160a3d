"""Dielectric functions eps(enei) -> (eps, k) with k = 2*pi/enei*sqrt(eps).

Wavelengths ``enei`` are vacuum wavelengths in nm. Host-side scalar work (one
complex number per wavelength and material); nothing here touches the GPU.

Reference interfaces kept: ``EpsConst`` (materials/eps_const.py:31-72),
``EpsDrude`` (materials/eps_drude.py:53-107) and ``EpsTable``
(materials/eps_table.py:51-152).
"""
import os

import numpy as np
from scipy.interpolate import CubicSpline

# eV <-> nm conversion of the MNPBEM toolbox (misc/units.py:8).
EV2NM = 1.0 / 8.0655477e-4

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def _wavenumber(enei, eps):
    return 2.0 * np.pi / enei * np.sqrt(eps)


class EpsConst:
    """Constant dielectric function (materials/eps_const.py:48-72)."""

    def __init__(self, eps):
        if eps is None:
            raise ValueError("EpsConst: 'eps' must be a numeric value, got None.")
        if not isinstance(eps, (int, float, complex, np.number)):
            raise TypeError("EpsConst: 'eps' must be numeric, got {!r}.".format(type(eps).__name__))
        self.eps = eps

    def __call__(self, enei):
        enei = np.asarray(enei, dtype=float)
        eps = np.full(enei.shape, self.eps, dtype=complex)
        return eps, _wavenumber(enei, eps)

    def wavenumber(self, enei):
        return self(enei)[1]

    def __repr__(self):
        return "EpsConst({})".format(self.eps)


class EpsDrude:
    """Drude metal eps = eps0 - wp^2 / (w (w + i gammad)), w = EV2NM / enei
    (materials/eps_drude.py:73-107). ``wp`` and ``gammad`` are in eV."""

    def __init__(self, eps0, wp, gammad, name=None):
        self.eps0 = eps0
        self.wp = wp
        self.gammad = gammad
        self.name = name

    def __call__(self, enei):
        enei = np.asarray(enei, dtype=float)
        w = EV2NM / enei
        eps = self.eps0 - self.wp ** 2 / (w * (w + 1j * self.gammad))
        return eps, _wavenumber(enei, eps)

    def wavenumber(self, enei):
        return self(enei)[1]

    def __repr__(self):
        return "EpsDrude({}, {}, {})".format(self.eps0, self.wp, self.gammad)


class EpsTable:
    """Tabulated refractive index n + ik from a text file of (eV, n, k) rows,
    interpolated in wavelength with a not-a-knot cubic spline
    (materials/eps_table.py:51-152). Out-of-range wavelengths raise ValueError."""

    def __init__(self, filename):
        path = filename if os.path.exists(filename) else os.path.join(_DATA, filename)
        if not os.path.exists(path):
            raise FileNotFoundError("Material data file not found: {}".format(filename))
        rows = []
        with open(path) as fh:
            for line in fh:
                line = line.strip()
                if not line or line[0] in "%#":
                    continue
                try:
                    vals = [float(t) for t in line.split()]
                except ValueError:
                    continue
                if len(vals) >= 3:
                    rows.append(vals[:3])
        if not rows:
            raise ValueError("No valid data found in {}".format(path))
        tab = np.asarray(rows)
        lam = EV2NM / tab[:, 0]
        order = np.argsort(lam)
        self.enei = lam[order]
        self.ni = CubicSpline(self.enei, tab[order, 1])
        self.ki = CubicSpline(self.enei, tab[order, 2])
        self.filename = os.path.basename(path)

    def __call__(self, enei):
        enei = np.asarray(enei, dtype=float)
        lo, hi = self.enei.min(), self.enei.max()
        if np.any(enei < lo) or np.any(enei > hi):
            raise ValueError("Wavelength out of range. Valid range: {:.1f} - {:.1f} nm, requested: "
                             "{:.1f} - {:.1f} nm".format(lo, hi, enei.min(), enei.max()))
        eps = (self.ni(enei) + 1j * self.ki(enei)) ** 2
        return eps, _wavenumber(enei, eps)

    def wavenumber(self, enei):
        return self(enei)[1]

    def __repr__(self):
        return "EpsTable('{}')".format(self.filename)
