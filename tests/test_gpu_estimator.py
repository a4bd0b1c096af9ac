"""GPU CBS estimator: FarFieldCBSSensor in estimator mode (process_estimation,
src/detector.cpp:1315-1566) against the oracle.

* FP64 + xoshiro: the device follows the reference's per-photon streams, so the
  estimator grids (every vertex forced into every bin) and the estimator hit
  counts match the oracle to rounding -- one Mie layer with a time-resolved
  grid beside analog sensors, the three-media stack (optical depth summed over
  the layers a forced ray crosses), and explicit phi slices with a theta cap and
  a scattering-order filter.
* FP32 production path: per-bin z-scores against the oracle from batch means.
* Runs the estimator cannot take (batches, gradients, no reverse-path tracking)
  are refused with the library's errors.
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import scenes
import stats
from luminis_mc.simulation import GpuBatch, GpuScene, SimConfig, run_desc_from
from parity import assert_trajectory_parity

pytestmark = pytest.mark.gpu


def _cfg(sample, laser, detectors, keep, seed=0):
    c = SimConfig()
    c.seed, c.track_reverse_paths = seed, True
    g = lmc.SensorsGroup()
    for d in detectors:
        g.add_detector(d)
    c.sample, c.laser, c.detector = sample, laser, g
    c._keep = keep
    return c


def _mie_estimator(time_resolved=True):
    """Semi-infinite Mie r = 0.3 um (mu_s 10, mu_a 0.5 /mm), circular point beam;
    estimator CBS 12 theta x 4 phi over 0.3 rad (4 time windows when
    time-resolved), with an analog planar fluence and reflectance beside it."""
    n = 1.33
    m, ph = scenes.mie_medium(0.3e-3, 1.59, n, 10.0, 0.5, n_div=4000)
    sample = lmc.Sample(n)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1 / math.sqrt(2), 1j / math.sqrt(2), scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    len_t = 40.0 * 0.299792458 / n if time_resolved else 0.0
    est = lmc.FarFieldCBSSensor(0.3, 2 * math.pi, len_t, 0.3 / 12, 2 * math.pi / 4, len_t / 4, True)
    flu = lmc.PlanarFluenceSensor(0.0, 4.0, 4.0, 0.0, 0.5, 0.5, 0.0, absorb=False)
    refl = lmc.ReflectanceSensor(0.0, absorb=False)
    return _cfg(sample, laser, [est, flu, refl], (m, ph))


def _multilayer_estimator():
    """The three media of the multilayer scene with an estimator CBS sensor
    (10 theta x 2 phi over 0.2 rad) and analog far-field fluence."""
    cfg, _ = scenes.config_multilayer(1, small=True)
    est = lmc.FarFieldCBSSensor(0.2, 2 * math.pi, 0.0, 0.02, math.pi, 0.0, True)
    fff = lmc.FarFieldFluenceSensor(math.pi / 2, 2 * math.pi, (math.pi / 2) / 18, 2 * math.pi / 9, absorb=False)
    return _cfg(cfg.sample, cfg.laser, [est, fff], cfg._keep)


def _sliced_estimator():
    """Linear polarisation, explicit phi slices, theta cap and order filter."""
    n = 1.33
    m, ph = scenes.mie_medium(0.2e-3, 1.59, n, 8.0, 0.2, n_div=4000)
    sample = lmc.Sample(n)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    est = lmc.FarFieldCBSSensor(0.4, 2 * math.pi, 0.0, 0.4 / 16, 2 * math.pi / 8, 0.0, True)
    est.set_phi_slices([4.0, 0.3, 1.0, 2.5, 1.0])
    est.theta_pp_max = 0.26
    est.set_events_limit(1, 40)
    return _cfg(sample, laser, [est], (m, ph))


def _run_desc(cfg, precision, rng, n, off=0):
    cfg.precision, cfg.rng = precision, rng
    return run_desc_from(cfg, n_photons=n, photon_offset=off)


def _fp64_parity(cfg, n, seed):
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    st = sc.run(_run_desc(cfg, "fp64", "xoshiro", n))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, n, seed)
    for i, (off, cnt) in enumerate(sc.build.offsets):
        assert np.abs(cpu_run[0][off:off + cnt]).max() > 0, i
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets, label="estimator")
    sc.close()
    return gpu, sc.build


def test_gpu_fp64_xoshiro_estimator_parity_mie(require_gpu):
    cfg = _mie_estimator()
    cfg.seed = 21
    gpu, build = _fp64_parity(cfg, 3000, 21)
    off, cnt = build.offsets[0]
    g = gpu[off:off + cnt].reshape(8, 5, 12, 4)
    # time-integrated row 0 is the sum of the time windows
    assert np.allclose(g[:, 0], g[:, 1:].sum(axis=1), rtol=1e-9, atol=1e-300)


def test_gpu_fp64_xoshiro_estimator_parity_multilayer(require_gpu):
    cfg = _multilayer_estimator()
    cfg.seed = 22
    _fp64_parity(cfg, 2000, 22)


def test_gpu_fp64_xoshiro_estimator_phi_slices(require_gpu):
    cfg = _sliced_estimator()
    cfg.seed = 23
    gpu, build = _fp64_parity(cfg, 3000, 23)
    off, cnt = build.offsets[0]
    g = gpu[off:off + cnt].reshape(8, 1, 16, 4)
    assert np.all(g[:, :, int(0.26 / (0.4 / 16)):] == 0)  # theta_pp_max caps the rows
    assert np.abs(g[0, 0, :5]).min() > 0


def test_gpu_fp64_xoshiro_estimator_study_scene(require_gpu):
    """The reference's estimator study (300 theta bins, RDG-EMC table on 100001
    nodes: a guide table whose entries are stored shifted)."""
    sample, laser, sensors, keep = scenes.scene_estimator()
    cfg = _cfg(sample, laser, sensors, keep)
    cfg.seed = 24
    _fp64_parity(cfg, 300, 24)


def test_gpu_fp32_estimator_statistical_parity(require_gpu):
    N, B, seed = 32000, 16, 31
    cfg = _mie_estimator(time_resolved=False)
    cfg.seed = seed
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    g = stats.gpu_batches(sc, lambda n, off: _run_desc(cfg, "fp32", "philox", n, off), N, B)
    c = stats.oracle_batches(sc.build, N, B, seed)
    off, cnt = sc.build.offsets[0]
    z, mask = stats.zscores(g[:, off:off + cnt], c[:, off:off + cnt])
    zz = z[mask]
    print("estimator bins compared", zz.size, "max|z|", np.abs(zz).max(), "frac |z|<3", np.mean(np.abs(zz) < 3))
    assert zz.size > 300
    assert np.abs(zz).max() < 5.0
    assert np.mean(np.abs(zz) < 3) > 0.98
    sc.close()


def test_gpu_estimator_refusals(require_gpu):
    cfg = _mie_estimator()
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    r = _run_desc(cfg, "fp32", "philox", 1000)
    r.track_reverse_paths = 0
    with pytest.raises(ValueError):
        sc.run(r)
    with pytest.raises(NotImplementedError):
        sc.run_adjoint(_run_desc(cfg, "fp32", "philox", 1000))
    sc.close()
    b = GpuBatch([(cfg.sample, cfg.laser, cfg.detector.detectors)])
    with pytest.raises(NotImplementedError):
        b.run(np.array([1000], dtype=np.uint64), _run_desc(cfg, "fp32", "philox", 0))
    b.close()
