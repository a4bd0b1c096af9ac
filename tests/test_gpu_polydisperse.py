"""GPU parity on the polydisperse Mie mixture with the planar field sensor.

* FP64 + xoshiro: trajectory-level agreement with the oracle (mixture species
  selection, 2048-node S tables, CBS and complex field accumulators).
* FP32 production path: per-bin z-scores against the oracle from batch means.
* A mixture of identical species is the homogeneous layer (same random
  streams: the species draw is part of every event's four words).
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import scenes
import stats
from luminis_mc.simulation import GpuScene, SimConfig, run_desc_from
from parity import assert_trajectory_parity

pytestmark = pytest.mark.gpu


def _small():
    return scenes.scene_polydisperse(K=8, n_table=2048, n_div=2000, n_theta=32, n_phi=8, field_n=64)


def _rd(precision, rng, n, seed, off=0):
    c = SimConfig()
    c.seed, c.track_reverse_paths, c.precision, c.rng = seed, True, precision, rng
    return run_desc_from(c, n_photons=n, photon_offset=off)


def test_polydisperse_fp64_xoshiro_trajectory_parity(require_gpu):
    sample, laser, sensors, keep = _small()
    sc = GpuScene(sample, laser, sensors)
    n = 3000
    st = sc.run(_rd("fp64", "xoshiro", n, seed=5))
    gpu, hits = sc.read()
    assert_trajectory_parity(st, gpu, hits, oracle.run(sc.build, n, 5), sc.build.offsets, ["cbs", "field"],
                             label="polydisperse")
    sc.close()


def test_polydisperse_fp32_statistical_parity(require_gpu):
    sample, laser, sensors, keep = _small()
    sc = GpuScene(sample, laser, sensors)
    N, B = 64000, 16
    g = stats.gpu_batches(sc, lambda n, off: _rd("fp32", "philox", n, 17, off), N, B)
    c = stats.oracle_batches(sc.build, N, B, seed=17)
    (o_cbs, n_cbs), (o_f, n_f) = sc.build.offsets
    for name, sl in (("cbs", slice(o_cbs, o_cbs + n_cbs)), ("field", slice(o_f, o_f + n_f))):
        z, mask = stats.zscores(g[:, sl], c[:, sl])
        zz = np.abs(z[mask])
        print(name, "bins", zz.size, "max|z|", zz.max(), "frac<3", np.mean(zz < 3))
        assert zz.size > 50
        assert zz.max() < 5.5 and np.mean(zz < 3) > 0.97
    sc.close()


def test_mixture_of_identical_species_is_homogeneous(require_gpu):
    m, ph = scenes.mie_medium(0.4e-3, 1.59, 1.33, 30.0, 0.002, n_div=2000)
    sigma = m.scattering_cross_section()
    mix = lmc.Sample(1.33)
    mix.add_mixture_layer([m, m, m, m], [30.0 / (4 * sigma)] * 4, 0.0, math.inf)
    L = mix.layers[0]
    hom_m, hom_ph = scenes.mie_medium(0.4e-3, 1.59, 1.33, L.mu_s_total, L.mu_a_total, n_div=2000)
    hom_m.set_mean_free_path(L.mfp_total)  # a mixture's free path mean is 1/mu_s (sample.cpp:77)
    hom = lmc.Sample(1.33)
    hom.add_layer(hom_m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    outs = []
    for sample in (mix, hom):
        group = lmc.SensorsGroup()
        group.add_detector(scenes.cbs_sensor(50, 0.1, 1))
        group.add_detector(lmc.ReflectanceSensor(0.0, absorb=False))
        sc = GpuScene(sample, laser, group.detectors)
        sc.run(_rd("fp32", "philox", 200000, seed=3))
        outs.append(sc.read())
        sc.close()
    (a, ha), (b, hb) = outs
    rel = np.abs(a - b).sum() / np.abs(b).sum()
    print("mixture vs homogeneous: rel L1", rel, "hits", ha, hb)
    assert rel < 1e-3
    assert np.all(np.abs(ha.astype(np.int64) - hb.astype(np.int64)) <= np.maximum(2, hb.astype(np.int64) // 1000))
