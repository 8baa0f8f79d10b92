"""Analytic rectangular-waveguide port mode (the mode function a Robin port hands to the
surface assembly).

Restates ``RectangularWaveguide`` (src/emerge/_emerge/physics/microwave/bcs/
port_bcs.py:862-1078) and the two ``CoordinateSystem`` transforms it uses (cs.py:526-565):
TE(m, n) field in the port's local frame, amplitude for the requested power, propagation
constant, ``gamma = j beta`` for the Robin blocks and ``U_inc = -2 j beta E`` for the
excitation. Host NumPy, like the reference: it is evaluated once per frequency at the
surface rule's points.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# const.py:20-24
C0 = 299792458.0
Z0 = 376.73031366857
EPS0 = 8.854187818814e-12
MU0 = 1 / (C0 * C0 * EPS0)


@dataclass
class LocalFrame:
    """Origin and orthonormal axes of a port face (cs.py CoordinateSystem)."""
    origin: np.ndarray
    xax: np.ndarray
    yax: np.ndarray
    zax: np.ndarray
    _basis_inv: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        self.origin = np.asarray(self.origin, dtype=np.float64)
        self.xax = np.asarray(self.xax, dtype=np.float64)
        self.yax = np.asarray(self.yax, dtype=np.float64)
        self.zax = np.asarray(self.zax, dtype=np.float64)
        self._basis_inv = np.linalg.pinv(np.array([self.xax, self.yax, self.zax]).T)

    def in_local_cs(self, x, y, z):
        """cs.py:526-546."""
        B = self._basis_inv
        xg, yg, zg = x - self.origin[0], y - self.origin[1], z - self.origin[2]
        return (B[0, 0] * xg + B[0, 1] * yg + B[0, 2] * zg,
                B[1, 0] * xg + B[1, 1] * yg + B[1, 2] * zg,
                B[2, 0] * xg + B[2, 1] * yg + B[2, 2] * zg)

    def in_global_basis(self, x, y, z):
        """cs.py:548-565."""
        return (self.xax[0] * x + self.yax[0] * y + self.zax[0] * z,
                self.xax[1] * x + self.yax[1] * y + self.zax[1] * z,
                self.xax[2] * x + self.yax[2] * y + self.zax[2] * z)


@dataclass
class RectangularWaveguideMode:
    """TE(m, n) mode of a width x height rectangular port (port_bcs.py:862-1077)."""
    frame: LocalFrame
    width: float
    height: float
    mode: tuple = (1, 0)
    er: float = 1.0
    power: float = 1.0
    polarization: float = 1.0

    def get_beta(self, k0: float):
        """port_bcs.py:963-972."""
        return np.sqrt(self.er * k0 ** 2 - (np.pi * self.mode[0] / self.width) ** 2
                       - (np.pi * self.mode[1] / self.height) ** 2)

    def get_gamma(self, k0: float) -> complex:
        """The Robin constant j beta (port_bcs.py:974-983)."""
        return 1j * self.get_beta(k0)

    def portZ0(self, k0: float):
        return k0 * C0 * MU0 / self.get_beta(k0)

    def get_amplitude(self, k0: float):
        """port_bcs.py:952-961."""
        Zte = k0 * C0 * MU0 / self.get_beta(k0)
        m, n = self.mode
        multiplier = 8.0 / ((2.0 if m == 0 else 1.0) * (2.0 if n == 0 else 1.0))
        return np.sqrt(self.power * multiplier * Zte / (self.width * self.height))

    def port_mode_3d(self, xl, yl, k0: float):
        """E in the local frame (port_bcs.py:1002-1030, which = "E")."""
        m, n = self.mode
        amp = self.polarization * self.get_amplitude(k0)
        Ev = amp * np.cos(np.pi * m * xl / self.width) * np.cos(np.pi * n * yl / self.height)
        Eh = amp * np.sin(np.pi * m * xl / self.width) * np.sin(np.pi * n * yl / self.height)
        return np.array([Eh, Ev, 0 * Eh])

    def port_mode_3d_global(self, x, y, z, k0: float, mode_nr=None, which: str = "E"):
        """port_bcs.py:1064-1077 (which "E" or "Exy"; the TE field has no local z part)."""
        if which not in ("E", "Exy"):
            raise ValueError(f"which must be 'E' or 'Exy', not {which}")
        xl, yl, _ = self.frame.in_local_cs(x, y, z)
        Ex, Ey, Ez = self.port_mode_3d(xl, yl, k0)
        return np.array(self.frame.in_global_basis(Ex, Ey, 0 * Ez))

    def get_Uinc(self, x, y, z, k0: float):
        """-2 j beta E (port_bcs.py:985-1000)."""
        return -2 * 1j * self.get_beta(k0) * self.port_mode_3d_global(x, y, z, k0)
