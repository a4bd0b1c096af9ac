"""GPU checks of sensor planes away from the usual z = 0 plane, of large
grids, and of the planar field's phase on long paths.

* A non-absorbing planar fluence and a reflectance sensor on a layer
  interface (z = 0.5 mm of the three-media scene): the FP64 check mode equals
  the oracle; the FP32 production path (which places a photon exactly on the
  interface it crosses instead of 1e-12 past it) records each crossing once,
  so its hit counts match the FP64 build on the same streams.
* A far-field CBS grid of 512 KiB (beyond the 48 KiB the sensor kernel keeps
  in LDS, so its deposits take HBM FP64 atomics): FP64 trajectory parity.
* The planar field's phase k0 * (optical path) on paths beyond 1e6 rad, FP32
  production path against the FP64 build photon by photon: the wavenumber and
  the refractive index enter in FP64 and the path accumulates in FP64, so the
  median relative phase difference is what the FP32 step lengths leave
  (measured +3.8e-8 over 413 photons: the hardware log2 behind -mfp ln u is
  biased by about 0.3 ulp), below the +8.2e-8 more an FP32 k0 and n would add
  at this wavelength.
"""
import ctypes
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import scenes
from luminis_mc import _lib
from luminis_mc.simulation import GpuScene, run_desc_from
from parity import assert_trajectory_parity

pytestmark = pytest.mark.gpu


def _rd(cfg, precision, rng, n, off=0):
    cfg.precision, cfg.rng = precision, rng
    return run_desc_from(cfg, n_photons=n, photon_offset=off)


def _interface_cfg(n, seed):
    cfg, _ = scenes.config_multilayer(n, seed=seed, small=True)
    g = lmc.SensorsGroup()
    for s in cfg.detector.detectors:
        g.add_detector(s)
    g.add_detector(lmc.PlanarFluenceSensor(0.5, 10.0, 10.0, 0.0, 10.0 / 16, 10.0 / 16, 0.0, absorb=False))
    refl = lmc.ReflectanceSensor(0.5, absorb=False)
    refl.set_direction_limit(lmc.CrossingDirection.Both)
    g.add_detector(refl)
    cfg.detector = g
    return cfg


def test_interface_plane_fp64_parity(require_gpu):
    cfg = _interface_cfg(20000, 41)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    st = sc.run(_rd(cfg, "fp64", "xoshiro", 20000))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, 20000, 41)
    assert int(cpu_run[1][5]) > 5000  # crossings of the interface plane, both ways
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets,
                             ["cbs", "fluence", "ff", "refl", "fluence@0.5", "refl@0.5"], label="interface plane")
    sc.close()


def test_interface_plane_fp32_counts_each_crossing_once(require_gpu):
    cfg = _interface_cfg(200000, 42)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    sc.run(_rd(cfg, "fp32", "philox", 200000))
    a32, h32 = sc.read()
    sc.zero()
    sc.run(_rd(cfg, "fp64", "philox", 200000))
    a64, h64 = sc.read()
    sc.close()
    print("hits fp32", list(h32), "fp64", list(h64))
    for k in (4, 5):  # the sensors on the interface
        assert abs(int(h32[k]) - int(h64[k])) <= 0.003 * int(h64[k]) + 10, k
    off, n = sc.build.offsets[5]
    assert abs(a32[off] - a64[off]) <= 0.003 * a64[off]


def test_cbs_grid_beyond_lds_fp64_parity(require_gpu):
    cfg, _ = scenes.config_plain_mie(1, seed=43)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.FarFieldCBSSensor(0.5, 2 * math.pi, 0.0, 0.5 / 128, 2 * math.pi / 64, 0.0))
    cfg.detector = g
    sc = GpuScene(cfg.sample, cfg.laser, g.detectors)
    assert sc.build.offsets[0][1] * 8 > 48 * 1024  # accumulators in HBM
    st = sc.run(_rd(cfg, "fp64", "xoshiro", 30000))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, 30000, 43)
    assert int(cpu_run[1][0]) > 3000
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets, ["cbs 128x64"], label="cbs beyond LDS")
    sc.close()


# wavelength where FP32 rounds k0 = 2 pi / lambda by +4.9e-8 relative (and n =
# 1.33 by +3.2e-8): an FP32 k0 and n would shift every phase by +8.2e-8 of it
WL_FIELD = 0.63592947e-3


def test_field_phase_on_long_paths_fp32_vs_fp64(require_gpu):
    n = 1.33
    m, ph = scenes.mie_medium(0.5e-3, 1.59, n, 0.5, 0.0, n_div=4000)
    sample = lmc.Sample(n)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, WL_FIELD, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.PlanarFieldSensor(0.0, 1000.0, 1000.0, 1000.0, 1000.0, absorb=False))
    g.add_detector(lmc.PhotonRecordSensor(0.0, absorb=True))
    sc = GpuScene(sample, laser, g.detectors)
    cfg = lmc.SimConfig()
    cfg.seed, cfg.track_reverse_paths, cfg.MAX_EVENTS = 44, True, 100000
    k0 = 2 * math.pi / WL_FIELD
    rel, phases = [], []
    for i in range(600):
        res = []
        for prec in ("fp32", "fp64"):
            sc.zero()
            sc.run(_rd(cfg, prec, "philox", 1, off=i))
            out, hits = sc.read()
            if int(hits[1]) != 1:
                break
            rec = sc.records()
            res.append((complex(out[0], out[1]), rec[0, 5], int(rec[0, 1])))
        if len(res) != 2 or res[0][2] != res[1][2]:
            continue
        phase = k0 * n * res[1][1]
        if phase < 1e6 or abs(res[1][0]) == 0.0:
            continue
        d = np.angle(res[0][0] * np.conj(res[1][0]))
        rel.append(d / phase)
        phases.append(phase)
    sc.close()
    rel = np.array(rel)
    print(f"{rel.size} photons beyond 1e6 rad (max {max(phases, default=0):.3g}): relative phase difference "
          f"mean {rel.mean():.3e} median {np.median(rel):.3e} max|.| {np.abs(rel).max():.3e}")
    assert rel.size >= 15
    assert abs(np.median(rel)) < 7e-8
