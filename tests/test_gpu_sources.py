"""Laser sources, emission-time profiles and run controls on the GPU against
the oracle (FP64 + xoshiro: the reference's draw order, so the grids agree to
rounding).

* Laser::emit_photon (laser.cpp:77-99) for the uniform disk and the Gaussian
  spot, and sample_emission_time (:42-75) for each temporal profile, seen
  through time-resolved CBS and planar-fluence sensors (launch times shift
  the arrival bins).
* The event cap (simulation.cpp:609-613) with a small MAX_EVENTS.
* A run of zero photons leaves the accumulators untouched.
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import scenes
from luminis_mc.simulation import GpuScene, SimConfig, run_desc_from
from parity import assert_trajectory_parity

pytestmark = pytest.mark.gpu


def _scene_with(source, profile, pulse_ps=0.0, rep_hz=0.0, offset_ps=0.0):
    cfg, _ = scenes.config_plain_mie(1, seed=0)
    cfg.laser = lmc.Laser(1 / math.sqrt(2), 1j / math.sqrt(2), scenes.WAVELENGTH, 0.2, source)
    cfg.laser.set_temporal_profile(profile, pulse_ps, rep_hz, offset_ps)
    len_t = 100.0 * 0.299792458 / 1.33
    group = lmc.SensorsGroup()
    group.add_detector(lmc.FarFieldCBSSensor(0.1, 2 * math.pi, len_t, 0.1 / 16, 2 * math.pi / 4, len_t / 8))
    group.add_detector(lmc.PlanarFluenceSensor(0.0, 10.0, 10.0, len_t, 10.0 / 16, 10.0 / 16, len_t / 8,
                                               absorb=False))
    cfg.detector = group
    return cfg


def _parity(cfg, n, seed, max_events=1000):
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    cfg.precision, cfg.rng, cfg.seed, cfg.MAX_EVENTS = "fp64", "xoshiro", seed, max_events
    st = sc.run(run_desc_from(cfg, n_photons=n))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, n, seed, max_events=max_events)
    assert np.abs(cpu_run[0]).max() > 0
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets, label="sources")
    sc.close()
    return gpu, sc.build


@pytest.mark.parametrize("source,profile,pulse_ps,rep_hz,offset_ps", [
    (lmc.LaserSource.Uniform, lmc.TemporalProfile.TopHat, 5000.0, 0.0, 2000.0),
    (lmc.LaserSource.Gaussian, lmc.TemporalProfile.Gaussian, 5000.0, 0.0, 8000.0),
    (lmc.LaserSource.Point, lmc.TemporalProfile.Exponential, 3000.0, 0.0, 0.0),
    (lmc.LaserSource.Point, lmc.TemporalProfile.PulseTrain, 0.0, 1.0e8, 0.0),
    (lmc.LaserSource.Uniform, lmc.TemporalProfile.CW, 6000.0, 0.0, 1000.0),
])
def test_gpu_fp64_xoshiro_sources_and_profiles(require_gpu, source, profile, pulse_ps, rep_hz, offset_ps):
    cfg = _scene_with(source, profile, pulse_ps, rep_hz, offset_ps)
    gpu, build = _parity(cfg, 6000, 31)
    # launch times spread the hits over the time windows (bin 0 is their sum)
    off, cnt = build.offsets[1]
    g = gpu[off:off + cnt].reshape(4, 9, -1)
    assert (g[0, 1:].sum(axis=1) > 0).sum() >= 3


def test_gpu_fp64_xoshiro_event_cap(require_gpu):
    cfg = _scene_with(lmc.LaserSource.Point, lmc.TemporalProfile.Delta)
    _parity(cfg, 6000, 32, max_events=5)


def test_zero_photons_is_a_no_op(require_gpu):
    cfg = _scene_with(lmc.LaserSource.Point, lmc.TemporalProfile.Delta)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    c = SimConfig()
    st = sc.run(run_desc_from(c, n_photons=0))
    out, hits = sc.read()
    assert st["photons"] == 0 and not out.any() and not hits.any()
    sc.close()
