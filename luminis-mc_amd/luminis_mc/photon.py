"""Host-side value types of the reference's Python surface that the GPU path
does not store per photon: ``Photon`` (include/luminis/core/photon.hpp:65-243,
bound at python/python_module.cpp:173-204), the laser's spatial sampling
functions (src/laser.cpp:11-27, bound at python_module.cpp:705-710) and the
logger controls (python_module.cpp:676-687).

On the GPU a photon's state lives in registers of the transport kernel and is
never materialised as an object; ``Photon`` here is the reference's plain
record for scripts that build or inspect one on the host.
"""
from __future__ import annotations

import enum
import math

from .linalg import CVec2, Matrix, Rng, Vec3


class Photon:
    """Photon(position, direction, m, n, wavelength) (src/photon.cpp:24-55):
    position and previous position at `position`, frame rows m, n, s =
    direction, k = 2 pi / wavelength."""

    def __init__(self, position: Vec3 | None = None, direction: Vec3 | None = None, m: Vec3 | None = None,
                 n: Vec3 | None = None, wavelength: float = 0.0):
        self.prev_pos = Vec3()
        self.pos = Vec3()
        self.P_local = Matrix(3, 3)
        self.events = 0
        self.penetration_depth = 0.0
        self.alive = True
        self.wavelength_nm = 0.0
        self.k = 0.0
        self.opticalpath = 0.0
        self.launch_time = 0.0
        self.velocity = 1.0
        self.weight = 1.0
        self.current_layer = 0
        self.polarized = True
        self.polarization = CVec2(1.0, 0.0)
        z3 = lambda: Matrix(3, 3)
        self._P0, self._P1, self._Pn2, self._Pn1, self._Pn = z3(), z3(), z3(), z3(), z3()
        self._r_1, self._r_n = Vec3(), Vec3()
        self._T = Matrix(2, 2)
        self._T_buf = Matrix(2, 2)
        if position is None:
            return
        self.pos = Vec3(position.x, position.y, position.z)
        self.prev_pos = Vec3(position.x, position.y, position.z)
        self.wavelength_nm = float(wavelength)
        self.k = 2.0 * math.pi / self.wavelength_nm
        for row, v in enumerate((m, n, direction)):
            self.P_local.data[row] = (v.x, v.y, v.z)

    # read-only CBS history of the reference's object (never filled on the host)
    P0 = property(lambda self: self._P0)
    P1 = property(lambda self: self._P1)
    Pn2 = property(lambda self: self._Pn2)
    Pn1 = property(lambda self: self._Pn1)
    Pn = property(lambda self: self._Pn)
    r_1 = property(lambda self: self._r_1)
    r_n = property(lambda self: self._r_n)
    matrix_T = property(lambda self: self._T)
    matrix_T_buffer = property(lambda self: self._T_buf)

    def set_polarization(self, polarization: CVec2) -> None:
        """Photon::set_polarization (src/photon.cpp:56-65)."""
        self.polarized = True
        self.polarization = CVec2(polarization.m, polarization.n)

    def get_stokes_parameters(self) -> list[float]:
        """Photon::get_stokes_parameters (src/photon.cpp:67-90): (I, Q, U, V)
        with V = -2 Im(Em En*); (1, 0, 0, 0) when unpolarized."""
        if not self.polarized:
            return [1.0, 0.0, 0.0, 0.0]
        em, en = self.polarization.m, self.polarization.n
        a, b = abs(em) ** 2, abs(en) ** 2
        x = em * en.conjugate()
        return [a + b, a - b, 2.0 * x.real, -2.0 * x.imag]


def uniform_distribution(rng: Rng, center: Vec3, sigma: float) -> Vec3:
    """A point uniform in the disk of radius sigma about center, z kept
    (src/laser.cpp:11-20): theta = 2 pi u1, r = sigma sqrt(u2)."""
    theta = rng.uniform() * 2.0 * math.pi
    r = math.sqrt(rng.uniform()) * sigma
    return Vec3(center.x + r * math.cos(theta), center.y + r * math.sin(theta), center.z)


def gaussian_distribution(rng: Rng, center: Vec3, sigma: float) -> Vec3:
    """Independent normals of standard deviation sigma about center in x and
    y, z kept (src/laser.cpp:22-27)."""
    x = rng.normal(center.x, sigma)
    y = rng.normal(center.y, sigma)
    return Vec3(x, y, center.z)


class LogLevel(enum.IntEnum):
    """Logger levels with the reference's values (include/luminis/log/logger.hpp:22)."""
    debug = 1
    info = 2
    warn = 3
    error = 4
    off = 6


_level = LogLevel.info


def set_log_level(level: LogLevel) -> None:
    """Set the package's log level (python_module.cpp:684-687).  The GPU path
    itself logs nothing; the level is kept for get_log_level and for scripts
    written against the reference."""
    global _level
    _level = LogLevel(int(level))


def get_log_level() -> LogLevel:
    return _level

