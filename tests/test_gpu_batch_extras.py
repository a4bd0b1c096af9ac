"""Absorption grids and PhotonRecordSensor records in batched runs.

A batch of four scenes -- two absorbing slabs with different Absorption grids
(src/absortion.cpp:13-130) and two semi-infinite media with a
PhotonRecordSensor (detector.cpp:427-485) at different laser wavelengths --
runs in one launch.  Each scene's grid and records are checked against the
oracle run of that scene alone (FP64 + xoshiro: the same photons), and the
record fields that name the scene (sensor index, k = 2 pi / wavelength)
against the scene they came from.  run_simulation_batched fills each config's
Absorption and recorded_photons as run_simulation_parallel does per config.
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import records_check as rc
import scenes
from luminis_mc.simulation import GpuBatch, SceneBuild, SimConfig, run_desc_from

pytestmark = pytest.mark.gpu


def _rd(n, seed):
    c = SimConfig()
    c.seed, c.track_reverse_paths, c.precision, c.rng = seed, True, "fp64", "xoshiro"
    return run_desc_from(c, n_photons=n)


def _record_parts(wavelength):
    """Semi-infinite Mie (r 0.5 um, mu_s 10, mu_a 0.1 /mm), circular point beam
    of the given wavelength; at z = 0 a CBS sensor (not absorbing), then a
    PhotonRecordSensor and a total-reflectance sensor."""
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 10.0, 0.1)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1 / math.sqrt(2), 1j / math.sqrt(2), wavelength, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.FarFieldCBSSensor(0.4, 2 * math.pi, 0.0, 0.02, 2 * math.pi / 4, 0.0))
    g.detectors[0].absorb_photons = False
    g.add_detector(lmc.PhotonRecordSensor(0.0))
    g.add_detector(lmc.ReflectanceSensor(0.0, absorb=True))
    return sample, laser, g.detectors, (m, ph)


def _slab_parts():
    """1 mm slab of Mie scatterers (mu_s 10, mu_a 1 /mm), point beam, total
    reflectance (z = 0) and transmittance (z = 1, forward) sensors."""
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 10.0, 1.0)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, 1.0)
    laser = lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.ReflectanceSensor(0.0, absorb=True))
    t = lmc.ReflectanceSensor(1.0, absorb=True)
    t.set_direction_limit(lmc.CrossingDirection.Forward)
    g.add_detector(t)
    return sample, laser, g.detectors, (m, ph)


def _build(parts, absorption=None):
    b = SceneBuild(*parts[:3], absorption)
    b._keep = parts[3]
    return b


def _sorted(rows):
    return rows[np.lexsort((rows[:, 15], rows[:, 14]))]


def _check_records(g, c, tag):
    print(tag, "records gpu/cpu", len(g), len(c))
    assert abs(len(g) - len(c)) <= 2, tag
    if len(g) == len(c):
        cols = [i for i in range(34) if i != rc.DEPTH]
        d = np.abs(_sorted(g)[:, cols] - _sorted(c)[:, cols]).max(axis=0)
        assert np.all(d < 1e-8), (tag, np.round(d, 12))


def test_batch_absorption_grids_and_records_per_scene(require_gpu):
    wl2 = 0.5e-3
    builds = [_build(_slab_parts(), lmc.Absorption(20.0, 1.0, 0.5, 0.05, 2.0, 10.0)),
              _build(_record_parts(scenes.WAVELENGTH)),
              _build(_record_parts(wl2)),
              _build(_slab_parts(), lmc.Absorption(10.0, 1.0, 0.25, 0.1))]
    budgets = np.array([6000, 5000, 4000, 5000], dtype=np.uint64)
    seed = 61
    batch = GpuBatch.from_builds(builds)
    st = batch.run(budgets, _rd(0, seed))
    assert st["photons"] == int(budgets.sum()) and st["aux_mode"] == 2
    for k in (0, 3):
        b = builds[k]
        out, hits = batch.read(k)
        cpu, chits, _, _ = oracle.run(b, int(budgets[k]), seed)
        size = b.absorption.buffer.size
        sl = slice(b.abs_offset, b.abs_offset + size)
        g, c = out[sl], cpu[sl]
        rel = np.abs(g - c).max() / np.abs(c).max()
        print("scene", k, "grid cells", size, "max rel diff", rel, "hits", list(hits), list(chits))
        assert np.abs(c).max() > 0 and rel < 2e-3
        assert len(batch.records(k)) == 0
    for k, wl in ((1, scenes.WAVELENGTH), (2, wl2)):
        b = builds[k]
        out, hits = batch.read(k)
        g = batch.records(k)
        oracle.run(b, int(budgets[k]), seed)
        c = oracle.records()
        assert len(g) == int(hits[1]) > 500
        assert np.all(g[:, rc.SI] == 1.0)  # the record sensor's index within its scene
        assert np.all(g[:, rc.K] == 2 * math.pi / wl)
        _check_records(g, c, f"scene {k}")
    batch.zero()
    assert all(len(batch.records(k)) == 0 for k in range(4))
    batch.close()


def test_run_simulation_batched_fills_absorption_and_records(require_gpu):
    def cfg_of(parts, n, absorption=None):
        c = SimConfig()
        c.n_photons, c.seed, c.precision, c.rng = n, 7, "fp64", "xoshiro"
        c.sample, c.laser = parts[0], parts[1]
        g = lmc.SensorsGroup()
        for s in parts[2]:
            g.add_detector(s)
        c.detector, c.absorption = g, absorption
        return c, parts[3]

    def configs():
        return [cfg_of(_slab_parts(), 6000, lmc.Absorption(20.0, 1.0, 0.5, 0.05, 2.0, 10.0)),
                cfg_of(_record_parts(scenes.WAVELENGTH), 4000),
                cfg_of(_slab_parts(), 3000, lmc.Absorption(10.0, 1.0, 0.25, 0.1))]

    bk, sk = configs(), configs()  # (config, media kept alive)
    batched, single = [c for c, _ in bk], [c for c, _ in sk]
    lmc.run_simulation_batched(batched)
    for c in single:
        lmc.run_simulation_parallel(c)
    for k in (0, 2):
        a, b = batched[k].absorption.buffer, single[k].absorption.buffer
        rel = np.abs(a - b).max() / np.abs(b).max()
        print("config", k, "absorption max rel diff batched/single", rel)
        assert b.sum() > 0 and rel < 2e-3
    det_b, det_s = batched[1].detector.detectors[1], single[1].detector.detectors[1]
    assert len(det_b.recorded_photons) == det_b.hits > 500
    _check_records(det_b.records, det_s.records, "batched vs single")
    lmc.run_simulation_batched(batched)  # results add up, records are appended
    assert len(det_b.recorded_photons) == det_b.hits and det_b.hits > len(det_s.recorded_photons)
