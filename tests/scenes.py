"""Scene builders shared by the tests, bench.py and smoke(): the workloads of
the reference's configurations, built through the luminis_mc public API."""
from __future__ import annotations

import math

import luminis_mc as lmc

WAVELENGTH = 0.633e-3  # mm


def mie_medium(radius_mm: float, n_p: float, n_m: float, mu_s: float, mu_a: float, n_div: int = 10000,
               n_table: int = 1000):
    ph = lmc.MiePhaseFunction(WAVELENGTH, radius_mm, n_p, n_m, n_div, 0.0, math.pi)
    m = lmc.MieMedium(ph, radius_mm, n_p, n_m, WAVELENGTH, n_table=n_table)
    m.set_scattering_coefficient(mu_s)
    m.set_absorption_coefficient(mu_a)
    m.set_mean_free_path(1.0 / (mu_s + mu_a))
    return m, ph


def cbs_sensor(n_theta: int = 100, theta_max: float = 0.1, n_phi: int = 1, len_t: float = 0.0, n_t: int = 0):
    dt = (len_t / n_t) if n_t else 0.0
    return lmc.FarFieldCBSSensor(theta_max, 2 * math.pi if n_phi > 1 else 6.28, len_t, theta_max / n_theta,
                                 (2 * math.pi / n_phi) if n_phi > 1 else 6.28, dt)


def config_plain_mie(n_photons: int, seed: int = 0, n_theta: int = 100):
    """Plain method: one semi-infinite Mie layer (r = 0.5 um, n_p = 1.59 in n = 1.33,
    mu_s = 10/mm, mu_a = 0.01/mm), linearly polarised point beam at normal
    incidence, far-field CBS 100 theta x 1 phi bins over 0..0.1 rad."""
    m, ph = mie_medium(0.5e-3, 1.59, 1.33, 10.0, 0.01)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, WAVELENGTH, 0.0, lmc.LaserSource.Point)
    group = lmc.SensorsGroup()
    det = group.add_detector(lmc.FarFieldCBSSensor(0.1, 6.28, 0.0, 0.1 / n_theta, 6.28, 0.0))
    cfg = lmc.SimConfig()
    cfg.n_photons = n_photons
    cfg.seed = seed
    cfg.track_reverse_paths = True
    cfg.sample, cfg.laser, cfg.detector = sample, laser, group
    cfg._keep = (m, ph)
    return cfg, det


def config_multilayer(n_photons: int, seed: int = 0, small: bool = True, with_field: bool = False):
    """Three media: [0,0.5) mm RGD r = 0.1 um (RDG-EMC theta sampler, mu_s 5, mu_a 0.02),
    [0.5,1.0) mm Mie r = 1.0 um (mu_s 20, mu_a 0.005), [1,inf) HG g = 0.9 on an RGD
    amplitude host (mu_s 10, mu_a 0.01); uniform n = 1.33; circular Gaussian beam
    sigma = 0.05 mm; CBS, planar fluence (x, y, t), far-field fluence and reflectance
    sensors at z = 0.  ``small`` shrinks the grids for CPU-oracle comparisons."""
    n = 1.33
    ph1 = lmc.RayleighDebyeEMCPhaseFunction(WAVELENGTH, 0.1e-3, 1.59, n, 2000 if small else 10000, 0.0, math.pi)
    m1 = lmc.RGDMedium(ph1, 0.1e-3, 1.59, n, WAVELENGTH)
    m1.set_scattering_coefficient(5.0)
    m1.set_absorption_coefficient(0.02)
    m1.set_mean_free_path(1.0 / 5.02)
    m2, ph2 = mie_medium(1.0e-3, 1.59, n, 20.0, 0.005, n_div=2000 if small else 10000)
    ph3 = lmc.HenyeyGreensteinPhaseFunction(0.9)
    m3 = lmc.RGDMedium(ph3, 0.1e-3, 1.59, n, WAVELENGTH)
    m3.set_scattering_coefficient(10.0)
    m3.set_absorption_coefficient(0.01)
    m3.set_mean_free_path(1.0 / 10.01)
    sample = lmc.Sample(n)
    sample.add_layer(m1, 0.0, 0.5)
    sample.add_layer(m2, 0.5, 1.0)
    sample.add_layer(m3, 1.0, math.inf)
    laser = lmc.Laser(1 / math.sqrt(2), 1j / math.sqrt(2), WAVELENGTH, 0.05, lmc.LaserSource.Gaussian)
    group = lmc.SensorsGroup()
    # time axis: 0-200 ps as geometric path length at c/n (0.299792458 mm/ps / 1.33)
    len_t = 200.0 * 0.299792458 / n
    if small:
        cbs = group.add_detector(lmc.FarFieldCBSSensor(0.1, 2 * math.pi, 0.0, 0.1 / 16, 2 * math.pi / 4, 0.0))
        flu = lmc.PlanarFluenceSensor(0.0, 10.0, 10.0, len_t, 10.0 / 16, 10.0 / 16, len_t / 8, absorb=False)
        fff = lmc.FarFieldFluenceSensor(math.pi / 2, 2 * math.pi, (math.pi / 2) / 18, 2 * math.pi / 9, absorb=False)
    else:
        cbs = group.add_detector(lmc.FarFieldCBSSensor(0.1, 2 * math.pi, 0.0, 0.1 / 256, 2 * math.pi / 64, 0.0))
        flu = lmc.PlanarFluenceSensor(0.0, 10.0, 10.0, len_t, 10.0 / 256, 10.0 / 256, len_t / 64, absorb=False)
        fff = lmc.FarFieldFluenceSensor(math.pi / 2, 2 * math.pi, (math.pi / 2) / 180, 2 * math.pi / 90,
                                        absorb=False)
    flu = group.add_detector(flu)
    fff = group.add_detector(fff)
    refl = group.add_detector(lmc.ReflectanceSensor(0.0, absorb=False))
    field = None
    if with_field:
        field = group.add_detector(lmc.PlanarFieldSensor(0.0, 4.0, 4.0, 4.0 / 32, 4.0 / 32, absorb=False))
    cfg = lmc.SimConfig()
    cfg.n_photons = n_photons
    cfg.seed = seed
    cfg.track_reverse_paths = True
    cfg.sample, cfg.laser, cfg.detector = sample, laser, group
    cfg._keep = (m1, m2, m3, ph1, ph2, ph3)
    return cfg, {"cbs": cbs, "fluence": flu, "ff_fluence": fff, "reflectance": refl, "field": field}


def scene_adjoint(n_theta: int = 100):
    """Two media for the gradient workload: [0, 1) mm Mie r = 0.3 um (n_p = 1.59 in
    n = 1.33; mu_s 15, mu_a 0.01 /mm) over a semi-infinite Henyey-Greenstein
    g = 0.8 layer on an RGD r = 0.1 um amplitude host (mu_s 8, mu_a 0.05 /mm);
    linearly polarised point beam; total reflectance and far-field CBS 100 theta x
    1 phi over 0..0.1 rad at z = 0.  Returns (sample, laser, sensors, keep)."""
    n = 1.33
    m1, ph1 = mie_medium(0.3e-3, 1.59, n, 15.0, 0.01)
    ph2 = lmc.HenyeyGreensteinPhaseFunction(0.8)
    m2 = lmc.RGDMedium(ph2, 0.1e-3, 1.59, n, WAVELENGTH)
    m2.set_scattering_coefficient(8.0)
    m2.set_absorption_coefficient(0.05)
    m2.set_mean_free_path(1.0 / 8.05)
    sample = lmc.Sample(n)
    sample.add_layer(m1, 0.0, 1.0)
    sample.add_layer(m2, 1.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, WAVELENGTH, 0.0, lmc.LaserSource.Point)
    group = lmc.SensorsGroup()
    refl = group.add_detector(lmc.ReflectanceSensor(0.0, absorb=False))
    cbs = group.add_detector(cbs_sensor(n_theta, 0.1, 1))
    return sample, laser, group.detectors, (m1, ph1, m2, ph2, refl, cbs)


def scene_polydisperse(K: int = 32, n_table: int = 2048, n_div: int = 10000, n_theta: int = 512, n_phi: int = 128,
                       field_n: int = 1024, field_len: float = 4.0, theta_max: float = 0.1):
    """Polydisperse Mie suspension: a semi-infinite mixture of K radius bins of a
    lognormal distribution (median 0.4 um, sigma_ln 0.3, +-3 sigma, equal widths
    in ln r; bin weight = number fraction), n_p = 1.59 in n = 1.33, each species a
    MieMedium with n_table-node S1/S2 tables; number densities scaled to mu_s =
    30 /mm in total, mu_a = 0.002 /mm; linearly polarised point beam; far-field
    CBS n_theta x n_phi over 0..theta_max and a PlanarFieldSensor field_n x field_n
    over field_len x field_len mm at z = 0.  Returns (sample, laser, sensors, keep)."""
    n = 1.33
    species, dens, _ = lmc.lognormal_mie_mixture(WAVELENGTH, 0.4e-3, 0.3, 1.59, n, K, 30.0, 0.002, span=3.0,
                                                 n_div=n_div, n_table=n_table)
    keep = list(species)
    sample = lmc.Sample(n)
    sample.add_mixture_layer(species, dens, 0.0, math.inf)
    layer = sample.layers[0]
    laser = lmc.Laser(1.0, 0.0, WAVELENGTH, 0.0, lmc.LaserSource.Point)
    group = lmc.SensorsGroup()
    cbs = group.add_detector(lmc.FarFieldCBSSensor(theta_max, 2 * math.pi, 0.0, theta_max / n_theta,
                                                   2 * math.pi / n_phi, 0.0))
    field = group.add_detector(lmc.PlanarFieldSensor(0.0, field_len, field_len, field_len / field_n,
                                                     field_len / field_n, absorb=False))
    keep += [layer, cbs, field]
    return sample, laser, group.detectors, keep


def scene_estimator(n_theta: int = 300):
    """The reference's estimator study (results/estimator_vs_full/sim.py) in mm:
    semi-infinite non-absorbing medium of RGD spheres (r = 0.110 um, n_p = 1.59
    in n = 1.33, lambda = 0.514 um, RDG-EMC theta sampler on 100000 divisions),
    mu_s 10 /mm; circularly polarised Gaussian beam (radius 4 mean free paths);
    one FarFieldCBSSensor in estimator mode, 10 degrees in n_theta x 1 bins,
    time-integrated."""
    n = 1.33
    wl, r = 0.514e-3, 0.110e-3
    ph = lmc.RayleighDebyeEMCPhaseFunction(wl, r, 1.59, n, 100000, 0.0, math.pi)
    m = lmc.RGDMedium(ph, r, 1.59, n, wl)
    mu_s = 10.0
    m.set_scattering_coefficient(mu_s)
    m.set_absorption_coefficient(0.0)
    m.set_mean_free_path(1.0 / mu_s)
    sample = lmc.Sample(n)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1 / math.sqrt(2), -1j / math.sqrt(2), wl, 4.0 / mu_s, lmc.LaserSource.Gaussian)
    group = lmc.SensorsGroup()
    th = math.radians(10.0)
    group.add_detector(lmc.FarFieldCBSSensor(th, 2 * math.pi, 0.0, th / n_theta, 2 * math.pi, 0.0, True))
    return sample, laser, group.detectors, (m, ph)
