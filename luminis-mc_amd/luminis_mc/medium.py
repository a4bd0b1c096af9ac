"""Scattering media and the layered sample.

Mirrors include/luminis/core/medium.hpp + src/medium.cpp (RGD and Mie amplitude
tables, free-path field, coefficients) and include/luminis/core/sample.hpp +
src/sample.cpp (homogeneous and mixture layers, contiguity rules).
"""
from __future__ import annotations

import math

import numpy as np

from .linalg import CMatrix
from .phase import PhaseFunction, form_factor, mie_amplitudes


class _UniformTable:
    """DataTable on the uniform grid theta_i = i*dx (src/table.cpp:78-118)."""

    def __init__(self, dx: float, values: np.ndarray):
        if values.size == 0:
            raise ValueError("y_vals must not be empty")
        if dx <= 0.0:
            raise ValueError("dx must be positive")
        self.dx = float(dx)
        self.values = np.ascontiguousarray(values, dtype=np.complex128)

    def sample(self, x: float) -> complex:
        pos = x / self.dx
        if pos <= 0.0:
            return complex(self.values[0])
        last = self.values.size - 1
        if pos >= last:
            return complex(self.values[-1])
        i = int(pos)
        t = pos - i
        return complex(self.values[i] + t * (self.values[i + 1] - self.values[i]))


class ScatteringMedium:
    """Base medium (include/luminis/core/medium.hpp:71-157)."""

    def __init__(self, phase_func: PhaseFunction):
        if phase_func is None:
            raise ValueError("ScatteringMedium: phase function must not be None")
        self.phase_function = phase_func
        self._mu_a = 0.0
        self._mu_s = 0.0
        self._mu_t = 0.0
        self.mean_free_path = float("nan")  # the reference leaves it unset until set_mean_free_path
        self.S1_table: _UniformTable | None = None
        self.S2_table: _UniformTable | None = None

    mu_a = property(lambda self: self._mu_a)
    mu_s = property(lambda self: self._mu_s)
    mu_t = property(lambda self: self._mu_t)

    def set_scattering_coefficient(self, mu_s: float) -> None:
        self._mu_s = float(mu_s)
        self._mu_t = self._mu_a + self._mu_s

    def set_absorption_coefficient(self, mu_a: float) -> None:
        self._mu_a = float(mu_a)
        self._mu_t = self._mu_a + self._mu_s

    def set_mean_free_path(self, mfp: float) -> None:
        self.mean_free_path = float(mfp)

    def scattering_matrix(self, theta: float, phi: float = 0.0) -> CMatrix:
        out = CMatrix(2, 2)
        out.data[0, 0] = self.S2_table.sample(theta)
        out.data[1, 1] = self.S1_table.sample(theta)
        return out

    def scattering_cross_section(self) -> float:
        return self.phase_function.scattering_cross_section() if self.phase_function else 0.0

    def sample_free_path(self, rng) -> float:
        return -1 * self.mean_free_path * math.log(rng.uniform())

    def sample_scattering_angle(self, rng) -> float:
        return self.phase_function.sample_theta(rng.uniform())

    def sample_azimuthal_angle(self, rng) -> float:
        return self.phase_function.sample_phi(rng.uniform())

    def sample_conditional_azimuthal_angle(self, rng, S, E, theta):
        return self.phase_function.sample_phi_conditional(theta, S, E, self.k, rng)

    # flat table export for the device scene
    def s_tables(self) -> tuple[np.ndarray, np.ndarray, float]:
        return self.S1_table.values, self.S2_table.values, self.S1_table.dx


class RGDMedium(ScatteringMedium):
    """Rayleigh-Gans-Debye amplitudes on 2000 theta nodes (src/medium.cpp:77-121)."""

    def __init__(self, phase_func: PhaseFunction, radius: float, n_particle: float, n_medium: float,
                 wavelength: float):
        super().__init__(phase_func)
        self.radius = float(radius)
        self.n_particle, self.n_medium = float(n_particle), float(n_medium)
        self.wavelength = float(wavelength)
        self.k = 2 * math.pi * self.n_medium / self.wavelength
        kkk = self.k ** 3
        volume = 4.0 * math.pi * self.radius ** 3 / 3.0
        mm = self.n_particle / self.n_medium - 1.0
        self.rgd_prefactor = -1.0 * kkk * mm * volume / (2.0 * math.pi)
        self.precompute_scattering_tables(2000)

    def precompute_scattering_tables(self, n_samples: int = 2000) -> None:
        n = max(2, int(n_samples))
        dth = math.pi / (n - 1)
        th = np.arange(n) * dth
        F = form_factor(th, self.k, self.radius)
        self.S1_table = _UniformTable(dth, 1j * (self.rgd_prefactor * F))
        self.S2_table = _UniformTable(dth, 1j * (self.rgd_prefactor * F * np.cos(th)))


class MieMedium(ScatteringMedium):
    """Mie amplitudes at x = 2 pi r / wavelength on ``n_table`` theta nodes
    (src/medium.cpp:160-236; the reference fixes n_table = 1000)."""

    def __init__(self, phase_func: PhaseFunction, radius: float, n_particle: float, n_medium: float,
                 wavelength: float, n_table: int = 1000):
        super().__init__(phase_func)
        self.radius = float(radius)
        self.n_particle, self.n_medium = float(n_particle), float(n_medium)
        self.m = complex(self.n_particle / self.n_medium, 0.0)
        self.wavelength = float(wavelength)
        self.k = 2 * math.pi * self.n_medium / self.wavelength
        self.precompute_scattering_tables(self.wavelength, 2 * math.pi * self.radius / self.wavelength,
                                          int(n_table))

    def precompute_scattering_tables(self, wavelength: float, size_parameter: float, n_samples: int = 1000):
        dth = math.pi / (n_samples - 1)
        mu = np.cos(np.arange(n_samples) * dth)
        s1, s2 = mie_amplitudes(size_parameter, self.m, mu)
        self.S1_table = _UniformTable(dth, s1)
        self.S2_table = _UniformTable(dth, s2)


# ---------------------------------------------------------------- layers --
class Layer:
    """A z-slab [z_min, z_max) (include/luminis/core/sample.hpp:68-114)."""

    def __init__(self, z_min: float, z_max: float):
        self.z_min, self.z_max = float(z_min), float(z_max)

    def contains(self, z: float) -> bool:
        return self.z_min <= z < self.z_max

    def thickness(self) -> float:
        return self.z_max - self.z_min


class SampleLayer(Layer):
    """HomogeneousLayer (sample.hpp:129-148)."""

    def __init__(self, medium: ScatteringMedium, z_min: float, z_max: float):
        super().__init__(z_min, z_max)
        self.medium = medium

    def mu_attenuation(self) -> float:
        return self.medium.mu_t

    def mu_scattering(self) -> float:
        return self.medium.mu_s

    def mu_absorption(self) -> float:
        return self.medium.mu_a

    @property
    def free_path_mean(self) -> float:
        return self.medium.mean_free_path

    @property
    def species(self):
        return [self.medium]


class MixtureLayer(Layer):
    """Co-located species drawn per vertex with probability mu_s^(i)/mu_s
    (src/sample.cpp:41-104)."""

    def __init__(self, species, number_densities, z_min: float, z_max: float):
        super().__init__(z_min, z_max)
        species = list(species)
        number_densities = [float(v) for v in number_densities]
        if not species:
            raise ValueError("MixtureLayer: species list must not be empty.")
        if len(species) != len(number_densities):
            raise ValueError("MixtureLayer: species and number_densities sizes differ.")
        self.species = species
        self.number_densities = number_densities
        self.mu_s_i = []
        self.mu_s_total = 0.0
        self.mu_a_total = 0.0
        for sp, nd in zip(species, number_densities):
            if sp is None:
                raise ValueError("MixtureLayer: species pointer must not be null.")
            mu_s = nd * sp.scattering_cross_section()
            self.mu_s_i.append(mu_s)
            self.mu_s_total += mu_s
            if sp.mu_s > 0.0:
                self.mu_a_total += mu_s * (sp.mu_a / sp.mu_s)
        if self.mu_s_total <= 0.0:
            raise ValueError("MixtureLayer: total μ_s must be positive (check number densities and that the "
                             "species' phase functions expose a scattering cross-section).")
        self.mu_t_total = self.mu_s_total + self.mu_a_total
        self.mfp_total = 1.0 / self.mu_s_total
        acc = 0.0
        self.selection_cdf = []
        for v in self.mu_s_i:
            acc += v / self.mu_s_total
            self.selection_cdf.append(acc)
        self.selection_cdf[-1] = 1.0

    def mu_attenuation(self) -> float:
        return self.mu_t_total

    def mu_scattering(self) -> float:
        return self.mu_s_total

    def mu_absorption(self) -> float:
        return self.mu_a_total

    @property
    def free_path_mean(self) -> float:
        return self.mfp_total


class Sample:
    """Contiguous layer stack with one host refractive index (src/sample.cpp:110-297)."""

    def __init__(self, n_medium: float = 1.0):
        self.refractive_index = float(n_medium)
        self.light_speed = 1.0 / float(n_medium)
        self.layers: list[Layer] = []
        self.interfaces: list[float] = []

    def light_speed_in_medium(self) -> float:
        return self.light_speed

    def _validate_contiguous(self, z_min: float, z_max: float) -> None:
        if z_max <= z_min:
            raise ValueError("Sample::add_layer: z_max must be greater than z_min.")
        if not self.layers:
            if z_min != 0.0:
                raise ValueError("Sample::add_layer: first layer must start at z = 0.")
            return
        prev = self.layers[-1].z_max
        if math.isinf(prev):
            raise ValueError("Sample::add_layer: cannot add a layer after an infinite top layer.")
        if z_min != prev:
            raise ValueError(f"Sample::add_layer: layer z_min ({z_min}) does not match previous z_max ({prev}).")

    def add_layer(self, medium: ScatteringMedium, z_min: float, z_max: float) -> None:
        if medium is None:
            raise ValueError("Sample::add_layer: medium pointer must not be null.")
        z_min, z_max = float(z_min), float(z_max)
        self._validate_contiguous(z_min, z_max)
        self.layers.append(SampleLayer(medium, z_min, z_max))
        if len(self.layers) > 1:
            self.interfaces.append(z_min)

    def add_mixture_layer(self, species, number_densities, z_min: float, z_max: float) -> None:
        z_min, z_max = float(z_min), float(z_max)
        self._validate_contiguous(z_min, z_max)
        layer = MixtureLayer(species, number_densities, z_min, z_max)
        self.layers.append(layer)
        if len(self.layers) > 1:
            self.interfaces.append(z_min)

    def size(self) -> int:
        return len(self.layers)

    def get_layer(self, index: int) -> Layer:
        return self.layers[index]

    def get_layer_index_at(self, z: float) -> int:
        if not self.layers or z < 0.0:
            return len(self.layers)
        top = self.layers[-1].z_max
        if not math.isinf(top) and z >= top:
            return len(self.layers)
        return int(np.searchsorted(np.asarray(self.interfaces), z, side="right"))

    def is_inside(self, position) -> bool:
        if not self.layers:
            return False
        vals = position.as_tuple() if hasattr(position, "as_tuple") else tuple(position)
        if any(math.isnan(v) or math.isinf(v) for v in vals):
            return False
        if vals[2] < 0.0:
            return False
        top = self.layers[-1].z_max
        return math.isinf(top) or vals[2] < top

    def z_top(self) -> float:
        return self.layers[-1].z_max if self.layers else 0.0


# ------------------------------------------------------ polydisperse media --
def lognormal_mie_mixture(wavelength: float, median_radius: float, sigma_ln: float, n_particle: float,
                          n_medium: float, n_bins: int, mu_s: float, mu_a: float, span: float = 3.0,
                          n_div: int = 10000, n_table: int = 1000):
    """A lognormal distribution of sphere radii as K Mie species for
    ``Sample.add_mixture_layer`` (the MixtureLayer route of src/sample.cpp:41-104:
    species drawn per vertex with probability n_i sigma_i / sum n_j sigma_j).

    ln r is cut at median +- span * sigma_ln into n_bins bins of equal width in
    ln r; bin k is one MieMedium at its geometric-mean radius, with number
    density proportional to the lognormal probability of the bin, scaled so
    the mixture's total scattering coefficient is mu_s (1/mm).  Every species
    carries mu_a/mu_s of the mixture, so its absorption totals mu_a.  Each
    species gets a MiePhaseFunction of n_div divisions over [0, pi] and
    n_table-node S1/S2 tables.

    Returns (species, number_densities, weights): weights are the bins'
    lognormal probabilities, renormalised to sum to one."""
    from .phase import MiePhaseFunction
    K = int(n_bins)
    if K < 1:
        raise ValueError("lognormal_mie_mixture: n_bins must be >= 1")
    if not (median_radius > 0.0 and sigma_ln > 0.0 and span > 0.0):
        raise ValueError("lognormal_mie_mixture: need median_radius > 0, sigma_ln > 0 and span > 0")
    if not (mu_s > 0.0 and mu_a >= 0.0):
        raise ValueError("lognormal_mie_mixture: need mu_s > 0 and mu_a >= 0")
    ln_med = math.log(median_radius)
    edges = [ln_med - span * sigma_ln + 2.0 * span * sigma_ln * k / K for k in range(K + 1)]
    cdf = [0.5 * (1.0 + math.erf((e - ln_med) / (sigma_ln * math.sqrt(2.0)))) for e in edges]
    weights = [cdf[k + 1] - cdf[k] for k in range(K)]
    tot = sum(weights)
    weights = [w / tot for w in weights]
    species = []
    for k in range(K):
        r = math.exp(0.5 * (edges[k] + edges[k + 1]))
        ph = MiePhaseFunction(wavelength, r, n_particle, n_medium, int(n_div), 0.0, math.pi)
        m = MieMedium(ph, r, n_particle, n_medium, wavelength, n_table=n_table)
        m.set_scattering_coefficient(mu_s)
        m.set_absorption_coefficient(mu_a)
        m.set_mean_free_path(1.0 / (mu_s + mu_a))
        species.append(m)
    scale = mu_s / sum(w * m.scattering_cross_section() for w, m in zip(weights, species))
    return species, [w * scale for w in weights], weights
