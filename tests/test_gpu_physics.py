"""Physics invariants of the GPU transport (no oracle involved).

* Coherent backscattering cone: for a non-absorbing semi-infinite medium of
  small Rayleigh-Gans-Debye scatterers, the coherent/incoherent ratio of the
  backscattered intensity is largest at exact backscatter (towards 2 in the
  co-polarised channel, less in total intensity because single scattering has
  no partner), falls off over ~lambda/(2 pi l*) and tends to 1 in the wings.
* Mixture of identical species (test_gpu_polydisperse.py) and the reverse
  field against brute-force reversed paths (test_oracle_cpu.py) cover the rest.

The scene uses RayleighDebyeEMCPhaseFunction, whose theta density carries the
sin(theta) Jacobian; the reference's RayleighDebyePhaseFunction samples theta
with density F(theta)^2 (no Jacobian, src/phase.cpp:138-142), which sends
single scattering preferentially into exact backscatter and buries the cone
under the unenhanced single-scattering background.
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import scenes
from luminis_mc.simulation import GpuScene, SimConfig, run_desc_from

pytestmark = pytest.mark.gpu


def _cone(pol: str, n_photons: int = 100_000_000):
    n = 1.33
    r = 0.05e-3
    ph = lmc.RayleighDebyeEMCPhaseFunction(scenes.WAVELENGTH, r, 1.59, n, 2000, 0.0, math.pi)
    m = lmc.RGDMedium(ph, r, 1.59, n, scenes.WAVELENGTH)
    m.set_scattering_coefficient(100.0)
    m.set_absorption_coefficient(0.0)
    m.set_mean_free_path(1.0 / 100.0)
    sample = lmc.Sample(n)
    sample.add_layer(m, 0.0, math.inf)
    if pol == "linear":
        laser = lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    else:
        laser = lmc.Laser(1 / math.sqrt(2), 1j / math.sqrt(2), scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    group = lmc.SensorsGroup()
    group.add_detector(lmc.FarFieldCBSSensor(0.1, 6.28, 0.0, 0.002, 6.28, 0.0))
    sc = GpuScene(sample, laser, group.detectors)
    c = SimConfig()
    c.seed, c.track_reverse_paths = 12, True
    sc.run(run_desc_from(c, n_photons=n_photons))
    acc, hits = sc.read()
    sc.close()
    return acc.reshape(8, 50), int(hits[0])


@pytest.mark.parametrize("pol", ["linear", "circular"])
def test_cbs_cone_peaks_at_backscatter(require_gpu, pol):
    G, hits = _cone(pol)
    eta = G[0] / G[4]
    print(pol, "hits", hits, "eta(S0) first bins", np.round(eta[:8], 3), "wings", round(float(eta[25:].mean()), 3))
    assert eta[0] > 1.35
    assert eta[:3].mean() > eta[5:10].mean() > eta[25:].mean()
    assert eta[:3].mean() - eta[25:].mean() > 0.25
    assert 0.98 < eta[25:].mean() < 1.1
    assert np.all(eta < 2.05)
    if pol == "linear":  # co-polarised channel: (S0 + S1) / 2
        eco = (G[0] + G[1]) / (G[4] + G[5])
        print("co-polarised eta first bins", np.round(eco[:8], 3))
        assert eco[0] > eta[0] and eco[0] > 1.5


def test_cbs_cone_with_nanometre_scatterers(require_gpu):
    """Scatterers of r = 0.5 nm have |S1|, |S2| ~ 1e-7: the FP32 path (interior
    product renormalised every other event, reverse-field prefix brought to
    unit scale) must still form the reverse field: the cone is present
    (a zeroed reverse field would give coherent = incoherent)."""
    n = 1.33
    r = 0.5e-6
    ph = lmc.RayleighDebyeEMCPhaseFunction(scenes.WAVELENGTH, r, 1.59, n, 2000, 0.0, math.pi)
    m = lmc.RGDMedium(ph, r, 1.59, n, scenes.WAVELENGTH)
    m.set_scattering_coefficient(100.0)
    m.set_absorption_coefficient(0.0)
    m.set_mean_free_path(1.0 / 100.0)
    sample = lmc.Sample(n)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    group = lmc.SensorsGroup()
    group.add_detector(lmc.FarFieldCBSSensor(0.1, 6.28, 0.0, 0.002, 6.28, 0.0))
    sc = GpuScene(sample, laser, group.detectors)
    c = SimConfig()
    c.seed, c.track_reverse_paths = 13, True
    sc.run(run_desc_from(c, n_photons=20_000_000))
    acc, hits = sc.read()
    sc.close()
    G = acc.reshape(8, 50)
    eta = G[0] / G[4]
    print("nm scatterers: hits", int(hits[0]), "eta(S0) first bins", np.round(eta[:6], 3))
    assert eta[0] > 1.3
