"""GPU parity of the batched launch (lmc_batch_*).

* FP64 + xoshiro: every scene of a batch reproduces the oracle's run of that
  scene alone (trajectory-level parity, as for single scenes).
* FP32 production path: a batched scene reproduces its single-scene run on the same
  per-photon random streams.
* hipGraph replay: each replay adds exactly what an eager run adds, with
  other device work in between.
"""
import numpy as np
import pytest

import batch_scenes
import oracle
import scenes
from luminis_mc.simulation import GpuBatch, GpuScene, SimConfig, run_desc_from
from parity import assert_trajectory_parity

pytestmark = pytest.mark.gpu


def _batch(specs, extra_sensors=False):
    built = [batch_scenes.build_scene(s, n_div=4000, extra_sensors=extra_sensors) for s in specs]
    b = GpuBatch([(sa, la, g.detectors) for sa, la, g, _, _ in built])
    b._keep = built
    return b


def _rd(precision, rng, seed, offset=0, **kw):
    c = SimConfig()
    c.seed = seed
    c.MAX_EVENTS = 1000
    c.track_reverse_paths = True
    c.precision, c.rng = precision, rng
    c.photon_offset = offset
    for k, v in kw.items():
        setattr(c, k, v)
    return run_desc_from(c, n_photons=0)


def test_batch_fp64_xoshiro_matches_oracle_per_scene(require_gpu):
    """Each scene of a batch against the oracle's run of that scene alone:
    equal event, trial and hit counts, accumulators to rounding.  Besides its
    CBS sensor each scene has a reflectance and a far-field fluence sensor, so
    every scene's check rests on more than a thousand detected photons."""
    specs = batch_scenes.load()[:6]
    budgets = np.array([3000, 5000, 2500, 6000, 4000, 3500], dtype=np.uint64)
    b = _batch(specs, extra_sensors=True)
    st = b.run(budgets, _rd("fp64", "xoshiro", seed=11))
    print("batch stats", st)
    assert st["photons"] == int(budgets.sum())
    ev_cpu = tr_cpu = 0
    for k in range(len(specs)):
        gpu, hits = b.read(k)
        cpu_run = oracle.run(b.builds[k], int(budgets[k]), 11)
        ev_cpu += cpu_run[2]
        tr_cpu += cpu_run[3]
        assert int(cpu_run[1][1]) >= 1000, k  # photons reaching the reflectance sensor
        # per-scene event counts are not separated by the batch: compare the totals below
        one = {"events": cpu_run[2], "phi_trials": cpu_run[3]}
        assert_trajectory_parity(one, gpu, hits, cpu_run, b.builds[k].offsets, ["cbs", "refl", "ff"],
                                 label=f"scene {k} ({len(specs[k]['layers'])} layers)")
    assert int(st["events"]) == ev_cpu and int(st["phi_trials"]) == tr_cpu
    b.close()


def test_batch_fp32_matches_single_scene_runs(require_gpu):
    specs = batch_scenes.load()[6:18]
    rng = np.random.default_rng(5)
    budgets = rng.integers(10000, 40000, size=len(specs)).astype(np.uint64)
    b = _batch(specs)
    st = b.run(budgets, _rd("fp32", "philox", seed=99))
    assert st["photons"] == int(budgets.sum())
    tot_l1 = tot_ref = 0.0
    for k in range(len(specs)):
        gb, hb = b.read(k)
        sa, la, g, _, _ = b._keep[k]
        sc = GpuScene(sa, la, g.detectors)
        r = _rd("fp32", "philox", seed=99)
        r.n_photons = int(budgets[k])
        sc.run(r)
        gs, hs = sc.read()
        sc.close()
        l1 = np.abs(gb - gs).sum()
        tot_l1 += l1
        tot_ref += np.abs(gs).sum()
        print(f"scene {k}: hits batch={int(hb[0])} single={int(hs[0])} L1 diff / L1 = {l1 / max(np.abs(gs).sum(), 1e-300):.2e}")
        assert abs(int(hb[0]) - int(hs[0])) <= max(2, int(1e-3 * int(hs[0])))
    assert tot_l1 <= 2e-3 * tot_ref
    b.close()


def test_batch_graph_replay_matches_eager(require_gpu):
    specs = batch_scenes.load()[18:26]
    budgets = np.array([20000, 5000, 30000, 1, 0, 15000, 25000, 8000], dtype=np.uint64)
    b = _batch(specs)
    r = _rd("fp32", "philox", seed=7)
    b.run(budgets, r, use_graph=False)
    eager = [b.read(k) for k in range(len(specs))]
    other_cfg, _ = scenes.config_plain_mie(50000, seed=1)
    other = GpuScene(other_cfg.sample, other_cfg.laser, other_cfg.detector.detectors)
    prev = [(np.zeros_like(o), np.zeros_like(h)) for o, h in eager]
    b.zero()
    for rep in range(3):
        st = b.run(budgets, r, use_graph=True)
        assert st["photons"] == int(budgets.sum())
        other.run(run_desc_from(other_cfg))  # unrelated device work between replays
        for k in range(len(specs)):
            o, h = b.read(k)
            inc_o, inc_h = o - prev[k][0], h - prev[k][1]
            eo, eh = eager[k]
            scale = max(np.abs(eo).max(), 1e-300)
            assert np.array_equal(inc_h, eh), (rep, k)
            assert np.abs(inc_o - eo).max() <= 1e-9 * scale, (rep, k)
            prev[k] = (o, h)
    other.close()
    b.close()


def test_batch_rejects_bad_budgets(require_gpu):
    specs = batch_scenes.load()[:2]
    b = _batch(specs)
    with pytest.raises(ValueError):
        b.run([1, 2, 3], _rd("fp32", "philox", seed=1))
    with pytest.raises(ValueError):
        b.run([2 ** 33, 1], _rd("fp32", "philox", seed=1))
    b.close()
