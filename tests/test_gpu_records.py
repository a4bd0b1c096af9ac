"""GPU PhotonRecordSensor (detector.cpp:427-485) against the oracle.

* FP64 + xoshiro: the records equal the oracle's (same photons; compared
  after sorting by hit position), penetration depth to FP32 (it is queued as
  FP32 bits), and the CBS grid rebuilt from the GPU records equals the GPU's
  CBS sensor grid (the records carry the reverse Jones vector the sensor used).
* FP32 production path: counts, weight sums against a reflectance sensor,
  frames, the rebuilt CBS grid, and a run long enough to span several
  transport launches (record lists gathered per launch).
* run_simulation_parallel fills recorded_photons; gradient runs keep the same
  records as forward runs (batched runs keep them per scene:
  test_gpu_batch_extras.py).
"""
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import records_check as rc
import scenes
from luminis_mc.simulation import GpuScene, SceneBuild, SimConfig, run_desc_from
from test_oracle_cpu import _record_scene

pytestmark = pytest.mark.gpu


def _rd(precision, rng, n, seed, off=0):
    c = SimConfig()
    c.seed, c.track_reverse_paths, c.precision, c.rng = seed, True, precision, rng
    return run_desc_from(c, n_photons=n, photon_offset=off)


def _sorted(rows):
    return rows[np.lexsort((rows[:, 15], rows[:, 14]))]


def test_gpu_fp64_xoshiro_records_parity(require_gpu):
    b = _record_scene()
    sc = GpuScene(None, None, None, build=b)
    n, seed = 6000, 17
    sc.run(_rd("fp64", "xoshiro", n, seed))
    out, hits = sc.read()
    g = _sorted(sc.records())
    cpu_out, chits, _, _ = oracle.run(b, n, seed)
    c = _sorted(oracle.records())
    print("records gpu/cpu", len(g), len(c), "hits", list(hits), list(chits))
    assert len(g) == int(hits[1]) and abs(len(g) - len(c)) <= 2
    if len(g) == len(c):
        cols = [i for i in range(34) if i != rc.DEPTH]
        d = np.abs(g[:, cols] - c[:, cols]).max(axis=0)
        print("max |gpu - cpu| per field", np.round(d, 12))
        assert np.all(d < 1e-8)
        assert np.allclose(g[:, rc.DEPTH], c[:, rc.DEPTH], rtol=1e-6, atol=1e-7)
    off, cnt = b.offsets[0]
    rebuilt = rc.cbs_from_records(g, 20, 4, 0.4 / 20, 2 * math.pi / 4)
    cbs = out[off:off + cnt].reshape(8, 20, 4)
    assert np.allclose(rebuilt, cbs, rtol=1e-9, atol=1e-9 * np.abs(cbs).max())
    sc.close()


def test_gpu_fp32_records_consistency(require_gpu):
    b = _record_scene()
    sc = GpuScene(None, None, None, build=b)
    n = 200000
    st = sc.run(_rd("fp32", "philox", n, 5))
    out, hits = sc.read()
    rows = sc.records()
    refl = out[b.offsets[2][0]]
    print("records", len(rows), "hits", list(hits), "launches", st["transport_launches"])
    assert len(rows) == int(hits[1]) > 0.3 * n
    assert abs(rows[:, rc.W].sum() - refl) <= 1e-6 * refl
    fe, je = rc.frame_errors(rows)
    assert fe < 1e-4 and je < 1e-4
    assert np.abs(rows[:, 16]).max() == 0.0
    assert np.allclose(rows[:, rc.K], 2 * math.pi / scenes.WAVELENGTH)
    assert np.all(rows[:, rc.EV] >= 1) and np.all(rows[:, rc.DEPTH] >= 0)
    off, cnt = b.offsets[0]
    rebuilt = rc.cbs_from_records(rows, 20, 4, 0.4 / 20, 2 * math.pi / 4)
    cbs = out[off:off + cnt].reshape(8, 20, 4)
    # FP32 binning (exit angle) and FP32 path phase (k0 |r_n - r_1| ~ 1e4 rad)
    # move a few records between neighbouring bins and blur the coherent part
    err = np.abs(rebuilt - cbs).max() / np.abs(cbs).max()
    inc = abs(rebuilt[4].sum() - cbs[4].sum()) / cbs[4].sum()
    coh = abs(rebuilt[0].sum() - cbs[0].sum()) / cbs[0].sum()
    print("rebuilt CBS grid max rel diff", err, "S0 totals: incoherent", inc, "coherent", coh)
    assert err < 2e-2 and inc < 1e-4 and coh < 2e-3
    sc.zero()
    assert len(sc.records()) == 0
    sc.close()


def test_gpu_records_over_several_launches(require_gpu):
    """A run of 1.5e8 photons in a thin slab spans several transport launches;
    a record sensor behind a 0.1 rad cone keeps few of them, and its weight
    sum equals a reflectance sensor with the same filter."""
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 2.0, 0.1)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, 0.2)
    laser = lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    rec = lmc.PhotonRecordSensor(0.0)
    rec.set_theta_limit(0.0, 0.1)
    ref = lmc.ReflectanceSensor(0.0, absorb=True)
    ref.set_theta_limit(0.0, 0.1)
    b = SceneBuild(sample, laser, [rec, ref])
    b._keep = (m, ph)
    sc = GpuScene(None, None, None, build=b)
    st = sc.run(_rd("fp32", "philox", 150_000_000, 3))
    out, hits = sc.read()
    rows = sc.records()
    print("launches", st["transport_launches"], "records", len(rows), "hits", list(hits))
    assert st["transport_launches"] >= 2
    assert len(rows) == int(hits[0]) == int(hits[1]) > 1000
    assert abs(rows[:, rc.W].sum() - out[b.offsets[1][0]]) <= 1e-6 * out[b.offsets[1][0]]
    assert np.all(np.arccos(np.clip(-rows[:, 19], -1, 1)) <= 0.1 + 1e-6)
    sc.close()


def test_run_simulation_parallel_fills_recorded_photons(require_gpu):
    b = _record_scene()
    cfg = SimConfig()
    cfg.n_photons, cfg.seed = 5000, 4
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 10.0, 0.1)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, math.inf)
    cfg.sample, cfg.laser = sample, lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    det = g.add_detector(lmc.PhotonRecordSensor(0.0))
    cfg.detector = g
    lmc.run_simulation_parallel(cfg)
    recs = det.recorded_photons
    assert len(recs) == det.hits > 1000
    r = recs[0]
    assert isinstance(r, lmc.PhotonRecord) and r.events >= 1 and r.weight > 0
    assert abs(r.position_detector.z) == 0.0 and r.k == pytest.approx(2 * math.pi / scenes.WAVELENGTH)
    assert r.arrival_time == pytest.approx(r.launch_time + r.opticalpath, rel=1e-6)
    lmc.run_simulation_parallel(cfg)  # results add up, as merge_from appends
    assert len(det.recorded_photons) == det.hits and det.hits > len(recs)
    del b


def test_records_in_adjoint_runs_equal_forward(require_gpu):
    """A gradient run keeps the PhotonRecordSensor's records as a forward run
    does: on FP64 xoshiro streams the same crossings, field for field (their
    adjoint scores travel in separate quads of the crossing record)."""
    b = _record_scene()
    sc = GpuScene(None, None, None, build=b)
    n, seed = 6000, 19
    grad, _ = sc.run_adjoint(_rd("fp64", "xoshiro", n, seed))
    a_out, a_hits = sc.read()
    a = _sorted(sc.records())
    sc.zero()
    sc.run(_rd("fp64", "xoshiro", n, seed))
    f_out, f_hits = sc.read()
    f = _sorted(sc.records())
    print("records adjoint/forward", len(a), len(f), "hits", list(a_hits), list(f_hits))
    assert len(a) == len(f) == int(f_hits[1]) > 0 and list(a_hits) == list(f_hits)
    assert np.abs(a - f).max() <= 1e-12 * max(1.0, np.abs(f).max())
    assert np.abs(a_out - f_out).max() <= 1e-12 * np.abs(f_out).max()
    assert grad.any()
    sc.close()
