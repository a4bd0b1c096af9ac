"""Phase functions: theta samplers and their inverse-CDF tables.

Mirrors include/luminis/sample/phase.hpp and src/phase.cpp.  Tables are built
on the host (numpy + the library's Mie series and SamplingTable builder) and
handed to the GPU as flat arrays; ``sample_theta`` here evaluates the same
table on the host for tests and interactive use.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from .linalg import CMatrix, CVec2, Rng


def form_factor(theta, k: float, radius: float):
    """Sphere form factor 3(sin u - u cos u)/u^3, u = 2 k r sin(theta/2)
    (src/phase.cpp:10-20); exactly 1 at theta == 0."""
    th = np.asarray(theta, dtype=np.float64)
    u = 2.0 * k * radius * np.sin(th / 2.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = 3.0 * (np.sin(u) - u * np.cos(u)) / u ** 3
    f = np.where(np.abs(th) == 0.0, 1.0, f)
    return float(f) if np.ndim(theta) == 0 else f


def mie_amplitudes(x: float, m: complex, mu) -> tuple[np.ndarray, np.ndarray]:
    """S1, S2 at cosines ``mu`` (MIEV0 conventions, src/mie/dmiev.cpp:49-62)."""
    mu = np.ascontiguousarray(np.atleast_1d(mu), dtype=np.float64)
    s1 = np.zeros(2 * mu.size)
    s2 = np.zeros(2 * mu.size)
    L = _lib.lib()
    _lib.check(L.lmc_mie_amplitudes(float(x), float(m.real), float(m.imag), mu.size, _lib.dptr(mu),
                                     _lib.dptr(s1), _lib.dptr(s2)), "lmc_mie_amplitudes")
    return s1[0::2] + 1j * s1[1::2], s2[0::2] + 1j * s2[1::2]


def mie_info(x: float, m: complex) -> tuple[float, float, float]:
    """(Q_ext, Q_sca, g) (src/mie/dmiev.cpp:65-81)."""
    import ctypes
    qe, qs, g = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    L = _lib.lib()
    _lib.check(L.lmc_mie_info(float(x), float(m.real), float(m.imag), ctypes.byref(qe), ctypes.byref(qs),
                              ctypes.byref(g)), "lmc_mie_info")
    return qe.value, qs.value, g.value


class SamplingTable:
    """Inverse-CDF table on nDiv+1 uniform nodes (src/table.cpp:6-66)."""

    def __init__(self, pdf_values: np.ndarray, n_div: int, min_val: float, max_val: float):
        if n_div <= 0:
            raise ValueError("nDiv must be greater than 0")
        if min_val >= max_val:
            raise ValueError("minVal must be less than maxVal")
        pdf = np.ascontiguousarray(pdf_values, dtype=np.float64)
        if pdf.size != n_div + 1:
            raise ValueError("pdf must hold nDiv+1 values")
        self.n_div = int(n_div)
        self.min_val = float(min_val)
        self.max_val = float(max_val)
        self.step = (self.max_val - self.min_val) / self.n_div
        self.values = self.min_val + np.arange(n_div + 1) * self.step
        self.cdf = np.zeros(n_div + 1)
        _lib.check(_lib.lib().lmc_sampling_cdf(int(n_div), self.min_val, self.max_val, _lib.dptr(pdf),
                                               _lib.dptr(self.cdf)), "SamplingTable")

    def sample(self, u: float) -> float:
        idx = int(np.searchsorted(self.cdf, u, side="left"))
        if idx == 0:
            return float(self.values[0])
        if idx >= self.values.size:
            return float(self.values[-1])
        t = (u - self.cdf[idx - 1]) / (self.cdf[idx] - self.cdf[idx - 1])
        return float(self.values[idx - 1] + t * (self.values[idx] - self.values[idx - 1]))


class PhaseFunction:
    """Base class (include/luminis/sample/phase.hpp:35-67)."""

    kind = _lib.PHASE_UNIFORM
    table: SamplingTable | None = None
    g_hg = 0.0

    def sample(self, x: float) -> float:
        return self.sample_cos(x)

    def sample_cos(self, x: float) -> float:
        return math.cos(self.sample_theta(x))

    def sample_theta(self, x: float) -> float:
        raise NotImplementedError

    def sample_phi(self, x: float) -> float:
        return 2.0 * math.pi * x

    def sample_phi_conditional(self, theta: float, S: CMatrix, E: CVec2, k: float, rng: Rng) -> float:
        """Azimuth by rejection against F(phi) (src/phase.cpp:27-49)."""
        s2sq = abs(S.data[0, 0]) ** 2
        s1sq = abs(S.data[1, 1]) ** 2
        e1sq, e2sq = abs(E.m) ** 2, abs(E.n) ** 2
        e12 = (E.m * np.conj(E.n)).real
        a = s2sq * e1sq + s1sq * e2sq
        b = s1sq * e1sq + s2sq * e2sq
        c = 2.0 * (s2sq - s1sq) * e12
        fmax = 0.5 * (a + b) + 0.5 * math.sqrt((a - b) ** 2 + c * c)
        while True:
            phi = 2.0 * math.pi * rng.uniform()
            cp, sp = math.cos(phi), math.sin(phi)
            F = max(a * cp * cp + b * sp * sp + c * cp * sp, 0.0)
            if rng.uniform() * fmax <= F:
                return phi

    def scattering_efficiency(self) -> float:
        raise NotImplementedError("PhaseFunction::scattering_efficiency: Not implemented for this phase function.")

    def scattering_cross_section(self) -> float:
        raise NotImplementedError("PhaseFunction::scattering_cross_section: Not implemented for this phase function.")

    def get_anisotropy_factor(self, n_samples: int = 200000):
        rng = np.random.default_rng()
        c = np.array([self.sample_cos(u) for u in rng.random(int(n_samples))])
        g = float(c.mean())
        return [g, float(math.sqrt(max(c.var(), 0.0) / len(c)))]


class UniformPhaseFunction(PhaseFunction):
    kind = _lib.PHASE_UNIFORM

    def sample_cos(self, x: float) -> float:
        return 2.0 * x - 1.0

    def sample_theta(self, x: float) -> float:
        return math.acos(self.sample_cos(x))


class HenyeyGreensteinPhaseFunction(PhaseFunction):
    """Closed-form HG inversion (src/phase.cpp:102-115)."""

    kind = _lib.PHASE_HG

    def __init__(self, g: float):
        self.g = float(g)
        self.g_hg = self.g

    def sample_cos(self, x: float) -> float:
        g = self.g
        gg = g * g
        frac = (1.0 - gg) / (1.0 - g + 2.0 * g * x)
        return (1.0 + gg - frac * frac) / (2.0 * g)

    def sample_theta(self, x: float) -> float:
        return math.acos(self.sample_cos(x))


class _TablePhase(PhaseFunction):
    kind = _lib.PHASE_TABLE_THETA

    def _build(self, n_div: int, min_val: float, max_val: float):
        nodes = min_val + np.arange(int(n_div) + 1) * ((max_val - min_val) / int(n_div))
        self.table = SamplingTable(self.pdf(nodes), int(n_div), min_val, max_val)

    def sample_theta(self, x: float) -> float:
        if self.kind == _lib.PHASE_TABLE_COS:
            return math.acos(self.table.sample(x))
        return self.table.sample(x)

    def sample_cos(self, x: float) -> float:
        if self.kind == _lib.PHASE_TABLE_COS:
            return self.table.sample(x)
        return math.cos(self.table.sample(x))


class RayleighPhaseFunction(_TablePhase):
    """Table over cos(theta), PDF 3/8 (1 + x^2) (src/phase.cpp:87-98)."""

    kind = _lib.PHASE_TABLE_COS

    def __init__(self, nDiv: int, minVal: float, maxVal: float):
        self._build(nDiv, minVal, maxVal)

    def pdf(self, x):
        return 0.375 * (1.0 + np.asarray(x) ** 2)


class DrainePhaseFunction(_TablePhase):
    """Table over cos(theta) (src/phase.cpp:218-240)."""

    kind = _lib.PHASE_TABLE_COS

    def __init__(self, g: float, a: float, nDiv: int, minVal: float, maxVal: float):
        self.g, self.a = float(g), float(a)
        self._build(nDiv, minVal, maxVal)

    def pdf(self, x):
        x = np.asarray(x, dtype=np.float64)
        g, a = self.g, self.a
        hg = (1.0 - g * g) / np.power(1.0 + g * g - 2.0 * g * x, 1.5)
        dr = (1.0 + a * x * x) / (3 + a * (1 + 2 * g * g))
        return (1.0 / 4.0 * math.pi) * hg * dr


class RayleighDebyePhaseFunction(_TablePhase):
    """Table over theta with PDF F^2/(4 pi) (src/phase.cpp:119-142)."""

    def __init__(self, wavelength: float, radius: float, nDiv: int, minVal: float, maxVal: float):
        self.radius, self.wavelength = float(radius), float(wavelength)
        self.k = 2 * math.pi / self.wavelength
        self._build(nDiv, minVal, maxVal)

    def pdf(self, x):
        F = form_factor(np.asarray(x, dtype=np.float64), self.k, self.radius)
        return (1.0 / (4.0 * math.pi)) * F ** 2


class RayleighDebyeEMCPhaseFunction(_TablePhase):
    """RGD with the (1 + cos^2) dipole factor (src/phase.cpp:146-215)."""

    def __init__(self, wavelength: float, radius: float, n_particle: float, n_medium: float, nDiv: int,
                 minVal: float, maxVal: float):
        self.radius, self.wavelength = float(radius), float(wavelength)
        self.n_particle, self.n_medium = float(n_particle), float(n_medium)
        self.k = 2 * math.pi * self.n_medium / self.wavelength
        self.scattering_cross_section_value = self.scattering_cross_section()
        self._build(nDiv, minVal, maxVal)

    def scattering_efficiency(self) -> float:
        """Simpson's rule with N = 100000 (src/phase.cpp:168-195)."""
        k, r = self.k, self.radius
        x = r * k
        V = 4 * math.pi * r ** 3 / 3.0
        a = (k ** 6 * (self.n_particle / self.n_medium - 1.0) ** 2 * V * V) / (4 * math.pi ** 2)
        c = a / x ** 2
        N = 100000
        h = math.pi / N
        th = np.arange(N + 1) * h
        F = form_factor(th, k, r)
        f = F * F * np.sin(th) * (1.0 + np.cos(th) ** 2)
        w = np.ones(N + 1)
        w[1:N:2] = 4.0
        w[2:N:2] = 2.0
        return c * float(np.sum(w * f)) * h / 3.0

    def scattering_cross_section(self) -> float:
        return self.scattering_efficiency() * math.pi * self.radius ** 2

    def rho_phase_function(self, x):
        m1 = self.n_particle / self.n_medium - 1.0
        v = (4.0 / 3.0) * math.pi * self.radius ** 3
        F = form_factor(np.asarray(x, dtype=np.float64), self.k, self.radius)
        return (self.k ** 4 / (4.0 * math.pi * self.scattering_cross_section_value)) * m1 * m1 * v * v * F * F * (
            1.0 + np.cos(x) ** 2)

    def pdf(self, x):
        return self.rho_phase_function(x) * np.sin(x)


class MiePhaseFunction(_TablePhase):
    """Mie theta sampler at x = k r with k = 2 pi n_medium / wavelength
    (src/phase.cpp:244-304; note the S tables of MieMedium use 2 pi r / wavelength)."""

    def __init__(self, wavelength: float, radius: float, n_particle: float, n_medium: float, nDiv: int,
                 minVal: float, maxVal: float):
        self.radius, self.wavelength = float(radius), float(wavelength)
        self.n_particle, self.n_medium = float(n_particle), float(n_medium)
        self.k = 2 * math.pi * self.n_medium / self.wavelength
        self.m = complex(self.n_particle / self.n_medium, 0.0)
        self.scattering_cross_section_value = self.scattering_cross_section()
        self._build(nDiv, minVal, maxVal)

    def scattering_efficiency(self) -> float:
        return mie_info(self.radius * self.k, self.m)[1]

    def scattering_cross_section(self) -> float:
        return self.scattering_efficiency() * math.pi * self.radius ** 2

    def rho_phase_function(self, x):
        mu = np.clip(np.cos(np.asarray(x, dtype=np.float64)), -1.0, 1.0)
        s1, s2 = mie_amplitudes(self.k * self.radius, self.m, mu)
        val = (math.pi / (self.k * self.k * self.scattering_cross_section_value)) * (np.abs(s1) ** 2 + np.abs(s2) ** 2)
        return val if np.ndim(x) else float(val[0])

    def pdf(self, x):
        return self.rho_phase_function(x) * np.sin(x)
