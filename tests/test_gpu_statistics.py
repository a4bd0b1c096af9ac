"""GPU StatisticsSensor (detector.cpp:1571-1888) against the oracle.

* FP64 + xoshiro: the counts of all six histograms (events, exit theta and
  phi, penetration depth -- each with time windows --, arrival time, weight)
  equal the oracle's; the multilayer stack tracks the penetration depth
  across layers.
* FP32 production path: totals equal the hit count, histograms statistically
  equal to the oracle's (batch-mean z-scores).
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import scenes
import stats
from luminis_mc.simulation import GpuScene, run_desc_from

pytestmark = pytest.mark.gpu


def _stats(absorb=True):
    st = lmc.StatisticsSensor(0.0, absorb=absorb)
    st.set_time_resolution(60.0, 15.0)
    st.set_events_histogram_bins(1200)
    st.set_theta_histogram_bins(0.0, math.pi / 2, 18)
    st.set_phi_histogram_bins(0.0, 2 * math.pi, 12)
    st.set_depth_histogram_bins(60.0, 25)
    st.set_time_histogram_bins(100.0, 20)
    st.set_weight_histogram_bins(1.0, 10)
    return st


def _scene(cfg, sensors):
    g = lmc.SensorsGroup()
    for s in sensors:
        g.add_detector(s)
    cfg.detector = g
    return GpuScene(cfg.sample, cfg.laser, g.detectors)


def _rd(cfg, precision, rng, n, off=0):
    cfg.precision, cfg.rng = precision, rng
    return run_desc_from(cfg, n_photons=n, photon_offset=off)


@pytest.mark.parametrize("which", ["plain", "multilayer"])
def test_gpu_fp64_xoshiro_statistics_parity(require_gpu, which):
    if which == "plain":
        cfg, _ = scenes.config_plain_mie(1, seed=41)
    else:
        cfg, _ = scenes.config_multilayer(1, seed=41, small=True)
    sc = _scene(cfg, [_stats(), lmc.ReflectanceSensor(0.0, absorb=False)])
    n = 20000
    sc.run(_rd(cfg, "fp64", "xoshiro", n))
    gpu, hits = sc.read()
    cpu, chits, _, _ = oracle.run(sc.build, n, 41)
    off, cnt = sc.build.offsets[0]
    g, c = gpu[off:off + cnt], cpu[off:off + cnt]
    print(which, "hits", list(hits), list(chits), "count differences", int(np.abs(g - c).sum()))
    assert int(hits[0]) == int(chits[0]) > 1000
    assert np.array_equal(np.rint(g), g)
    # a handful of photons may branch differently on rounding
    assert np.abs(g - c).sum() <= 8
    sc.close()


def test_gpu_fp32_statistics(require_gpu):
    N, B, seed = 160000, 16, 43
    cfg, _ = scenes.config_plain_mie(N, seed=seed)
    st = _stats()
    sc = _scene(cfg, [st])
    g = stats.gpu_batches(sc, lambda n, off: _rd(cfg, "fp32", "philox", n, off), N, B)
    c = stats.oracle_batches(sc.build, N, B, seed)
    z, mask = stats.zscores(g, c)
    zz = z[mask]
    print("statistics bins compared", zz.size, "max|z|", np.abs(zz).max(), "frac |z|<3", np.mean(np.abs(zz) < 3))
    assert zz.size > 150
    assert np.abs(zz).max() < 5.0
    assert np.mean(np.abs(zz) < 3) > 0.98
    sc.zero()
    sc.run(_rd(cfg, "fp32", "philox", 50000))
    out, hits = sc.read()
    sc.build.scatter_results(out, hits)
    s = sc.build.sensors[0]
    h = int(hits[0])
    for name in ("events", "theta", "phi", "depth"):
        assert np.array(getattr(s, f"{name}_histogram"))[0].sum() == h, name
    assert sum(s.weight_histogram) == h
    sc.close()
