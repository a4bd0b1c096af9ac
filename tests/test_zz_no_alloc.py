"""lmc_scene_run_noalloc (the run without device allocation) against
lmc_scene_run, both through the C ABI: the same photons give the same
counts and accumulators, on two calls in a row (the second smaller, so
stale queue contents of the first would show), and a run whose crossing
records cannot fit the static queue is refused with an error."""
import ctypes

import numpy as np
import pytest

import luminis_mc as lmc
import scenes
from luminis_mc import _lib
from luminis_mc.simulation import GpuScene, run_desc_from

pytestmark = pytest.mark.gpu


def _noalloc(sc, rd):
    st = _lib.Stats()
    rc = _lib.lib().lmc_scene_run_noalloc(sc._h, ctypes.byref(rd), ctypes.byref(st))
    return rc, st.as_dict()


def _run_pair(cfg, precision, rng, n, off):
    cfg.precision, cfg.rng = precision, rng
    rd = run_desc_from(cfg, n_photons=n, photon_offset=off)
    ref = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    st_ref = ref.run(rd)
    want, hits_ref = ref.read()
    ref.close()
    return rd, st_ref, want, hits_ref


def _check(sc, rd, st_ref, want, hits_ref, cbs_tol, label):
    sc.zero()
    rc, st = _noalloc(sc, rd)
    assert rc == _lib.LMC_OK, (label, _lib.lib().lmc_last_error())
    got, hits = sc.read()
    print(label, "launches", st["transport_launches"], "vs", st_ref["transport_launches"], "events", st["events"])
    for k in ("photons", "events", "phi_trials", "crossings", "cbs_detections"):
        assert st[k] == st_ref[k], (label, k)
    assert np.array_equal(hits, hits_ref), label
    for i, (off, m) in enumerate(sc.build.offsets):
        g, w = got[off:off + m], want[off:off + m]
        scale = max(np.abs(w).max(), 1e-300)
        tol = cbs_tol if sc.build.sensors[i].kind == _lib.SENSOR_FARFIELD_CBS else 1e-12
        assert np.abs(g - w).max() <= tol * scale, (label, i, np.abs(g - w).max() / scale)
    return st


def test_noalloc_equals_scene_run_fp32_two_calls(require_gpu):
    """FP32 production path, the multilayer scene with four sensors: a first
    call of 6e5 photons (several launches of the static queue) and a second
    of 1.5e5; CBS grids agree to the FP32 renormalisation of the interior
    product (its timing depends on the launch split), the others to FP64
    summation order."""
    cfg, _ = scenes.config_multilayer(1, seed=71, small=True)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    for n, off in ((600000, 0), (150000, 10 ** 6)):
        rd, st_ref, want, hits_ref = _run_pair(cfg, "fp32", "philox", n, off)
        st = _check(sc, rd, st_ref, want, hits_ref, 1e-5, f"fp32 n={n}")
        if n == 600000:
            assert st["transport_launches"] > 1
    sc.close()


def test_noalloc_equals_scene_run_fp64_two_calls(require_gpu):
    """FP64 check mode (xoshiro streams): identical branches, accumulators to
    summation order.  The FP64 tables come with the scene's first ordinary
    run; before it the no-alloc call refuses to upload them."""
    cfg, _ = scenes.config_plain_mie(1, seed=72)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    cfg.precision, cfg.rng = "fp64", "xoshiro"
    rc, _ = _noalloc(sc, run_desc_from(cfg, n_photons=10, photon_offset=0))
    assert rc == _lib.LMC_ERR_INVALID and b"tables" in _lib.lib().lmc_last_error()
    sc.run(run_desc_from(cfg, n_photons=1, photon_offset=0))
    for n, off in ((60000, 0), (20000, 500000)):
        rd, st_ref, want, hits_ref = _run_pair(cfg, "fp64", "xoshiro", n, off)
        _check(sc, rd, st_ref, want, hits_ref, 1e-12, f"fp64 n={n}")
    sc.close()


def test_noalloc_refuses_what_does_not_fit(require_gpu):
    """An interior non-absorbing plane with a large event cap needs more
    queue per photon than the static array holds: LMC_ERR_INVALID naming the
    sizes.  A PhotonRecordSensor scene: LMC_ERR_UNSUPPORTED."""
    cfg, _ = scenes.config_plain_mie(1, seed=73)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.ReflectanceSensor(0.3, absorb=False))
    g.add_detector(lmc.FarFieldCBSSensor(0.1, 6.28, 0.0, 0.001, 6.28, 0.0))
    sc = GpuScene(cfg.sample, cfg.laser, g.detectors)
    cfg.MAX_EVENTS = 400000
    rc, _ = _noalloc(sc, run_desc_from(cfg, n_photons=1000))
    assert rc == _lib.LMC_ERR_INVALID and b"scratch" in _lib.lib().lmc_last_error()
    sc.close()
    g2 = lmc.SensorsGroup()
    g2.add_detector(lmc.PhotonRecordSensor(0.0, absorb=True))
    sc2 = GpuScene(cfg.sample, cfg.laser, g2.detectors)
    cfg.MAX_EVENTS = 1000
    rc, _ = _noalloc(sc2, run_desc_from(cfg, n_photons=1000))
    assert rc == _lib.LMC_ERR_UNSUPPORTED
    sc2.close()
