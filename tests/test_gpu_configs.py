"""The benchmark configurations at their configured sizes on the GPU.

* Multilayer, many detectors (CBS 256 x 64, planar fluence 256 x 256 x 65,
  far-field fluence 180 x 90, reflectance; n_div 10000): FP64 trajectory
  parity with the oracle, and FP32 per-bin z-scores from batch means on
  block-summed grids (16 x 16 pixel and angle blocks, every Stokes plane).
* Polydisperse (32 lognormal Mie species with 2048-node tables, CBS
  512 x 128, complex planar field 1024 x 1024): the same two checks.
* The 256-scene batch in one launch: FP64 parity with the oracle for every
  scene, and the FP32 batch at the autotuner's 1024-thread shape against each
  scene run alone on the same streams.
"""
import math

import numpy as np
import pytest

import batch_scenes
import oracle
import scenes
import stats
from luminis_mc.simulation import GpuBatch, GpuScene, SimConfig, run_desc_from
from luminis_mc.tuning import DEFAULT_CANDIDATES, autotune
from parity import assert_trajectory_parity

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 8


def _rd(seed, precision, rng, n=0, off=0, **kw):
    c = SimConfig()
    c.seed, c.track_reverse_paths, c.precision, c.rng = seed, True, precision, rng
    for k, v in kw.items():
        setattr(c, k, v)
    return run_desc_from(c, n_photons=n, photon_offset=off)


def _blocks(a, shape, f):
    """Sum a grid over (f0 x f1) blocks of its last two axes."""
    a = a.reshape(shape)
    s0, s1 = shape[-2] // f[0], shape[-1] // f[1]
    return a.reshape(shape[:-2] + (s0, f[0], s1, f[1])).sum(axis=(-3, -1))


def _zcheck(g, c, name, zmax=5.5, frac=0.97):
    z, mask = stats.zscores(g, c)
    zz = z[mask]
    print(f"{name}: bins {zz.size} max|z| {np.abs(zz).max():.2f} frac<3 {np.mean(np.abs(zz) < 3):.4f}")
    assert zz.size > 50, name
    assert np.abs(zz).max() < zmax, name
    assert np.mean(np.abs(zz) < 3) > frac, name


# ---------------------------------------------------------- multilayer ----
def _multilayer_reduce(build):
    (oc, nc), (of, nf), (off, nff), (orf, nr) = build.offsets

    def red(acc):
        cbs = _blocks(acc[oc:oc + nc], (8, 256, 64), (16, 16)).ravel()
        flu = _blocks(acc[of:of + nf], (4, 65, 256, 256), (16, 16))[:, 0].ravel()  # time-integrated row
        ff = _blocks(acc[off:off + nff], (4, 180, 90), (10, 10)).ravel()
        return np.concatenate([cbs, flu, ff, acc[orf:orf + nr]])

    return red, (8 * 16 * 4, 4 * 16 * 16, 4 * 18 * 9, 1)


def test_multilayer_full_size_fp64_parity(require_gpu):
    cfg, d = scenes.config_multilayer(20000, seed=61, small=False)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    assert sc.build.n_out > 17_000_000  # 256 x 256 x 65 x 4 fluence doubles beside the rest
    cfg.precision, cfg.rng = "fp64", "xoshiro"
    st = sc.run(run_desc_from(cfg))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, 20000, 61, n_threads=ORACLE_THREADS)
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets, ["cbs", "fluence", "ff", "refl"],
                             label="multilayer full size")
    sc.close()


def test_multilayer_full_size_fp32_zscores(require_gpu):
    N, B, seed = 320000, 16, 62
    cfg, d = scenes.config_multilayer(N, seed=seed, small=False)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    red, sizes = _multilayer_reduce(sc.build)
    per = N // B
    g, c = [], []
    for b in range(B):
        sc.zero()
        cfg.precision, cfg.rng = "fp32", "philox"
        sc.run(run_desc_from(cfg, n_photons=per, photon_offset=b * per))
        g.append(red(sc.read()[0]))
        c.append(red(oracle.run(sc.build, per, seed, n_threads=ORACLE_THREADS, photon_offset=b * per)[0]))
    sc.close()
    g, c = np.array(g), np.array(c)
    o = np.cumsum((0,) + sizes)
    for k, name in enumerate(("cbs", "fluence", "ff", "refl")):
        if name == "refl":
            z = (g[:, -1].mean() - c[:, -1].mean()) / math.sqrt(g[:, -1].var(ddof=1) / B + c[:, -1].var(ddof=1) / B)
            print("refl z", z)
            assert abs(z) < 4.5
            continue
        _zcheck(g[:, o[k]:o[k + 1]], c[:, o[k]:o[k + 1]], f"multilayer {name}")


# ------------------------------------------------------- polydisperse ----
def test_polydisperse_full_size_fp64_parity(require_gpu):
    sample, laser, sensors, keep = scenes.scene_polydisperse()
    sc = GpuScene(sample, laser, sensors)
    assert sc.build.desc.n_species == 32 and sc.build.offsets[0][1] == 8 * 512 * 128
    assert sc.build.offsets[1][1] == 4 * 1024 * 1024
    st = sc.run(_rd(63, "fp64", "xoshiro", n=4000))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, 4000, 63, n_threads=ORACLE_THREADS)
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets, ["cbs", "field"],
                             label="polydisperse full size")
    sc.close()


def test_polydisperse_full_size_fp32_zscores(require_gpu):
    sample, laser, sensors, keep = scenes.scene_polydisperse()
    sc = GpuScene(sample, laser, sensors)
    (oc, nc), (of, nf) = sc.build.offsets
    N, B, seed = 192000, 16, 64
    per = N // B

    def red(acc):
        cbs = _blocks(acc[oc:oc + nc], (8, 512, 128), (32, 8)).ravel()
        f = acc[of:of + nf].reshape(2, 1024, 1024, 2)  # [Ex, Ey][x][y][re, im]
        field = _blocks(np.moveaxis(f, -1, 1), (2, 2, 1024, 1024), (64, 64)).ravel()
        return np.concatenate([cbs, field])

    g, c = [], []
    for b in range(B):
        sc.zero()
        sc.run(_rd(seed, "fp32", "philox", n=per, off=b * per))
        g.append(red(sc.read()[0]))
        c.append(red(oracle.run(sc.build, per, seed, n_threads=ORACLE_THREADS, photon_offset=b * per)[0]))
    sc.close()
    g, c = np.array(g), np.array(c)
    _zcheck(g[:, :8 * 16 * 16], c[:, :8 * 16 * 16], "polydisperse cbs")
    _zcheck(g[:, 8 * 16 * 16:], c[:, 8 * 16 * 16:], "polydisperse field")


# ------------------------------------------------------------ the batch ----
def _all_scenes(n_div=4000):
    specs = batch_scenes.load()
    built = [batch_scenes.build_scene(s, n_div=n_div) for s in specs]
    return specs, built


def test_batch_of_256_fp64_parity_per_scene(require_gpu):
    specs, built = _all_scenes()
    assert len(specs) == 256
    budgets = np.full(256, 200, dtype=np.uint64)
    b = GpuBatch([(sa, la, g.detectors) for sa, la, g, _, _ in built])
    st = b.run(budgets, _rd(65, "fp64", "xoshiro", MAX_EVENTS=1000))
    assert st["photons"] == 200 * 256
    ev = tr = 0
    for k in range(256):
        gpu, hits = b.read(k)
        cpu_run = oracle.run(b.builds[k], 200, 65, n_threads=4)
        ev, tr = ev + cpu_run[2], tr + cpu_run[3]
        one = {"events": cpu_run[2], "phi_trials": cpu_run[3]}
        assert_trajectory_parity(one, gpu, hits, cpu_run, b.builds[k].offsets, ["cbs"], label=f"scene {k}")
    assert int(st["events"]) == ev and int(st["phi_trials"]) == tr
    b.close()


def test_batch_of_256_at_tuned_shape_equals_single_scenes(require_gpu):
    """All 256 scenes in one FP32 launch at 1/100 of their budgets on the
    autotuner's shape for this batch (1024-thread workgroups, one per CU,
    256-photon claims): each scene equals its single-scene run on the same
    streams up to FP32 rounding flips."""
    specs, built = _all_scenes()
    budgets = np.maximum(1, np.array([s["budget"] for s in specs]) // 100).astype(np.uint64)
    b = GpuBatch([(sa, la, g.detectors) for sa, la, g, _, _ in built])
    tune = autotune(b, _rd(66, "fp32", "philox"), budgets=budgets, fraction=0.2, min_photons=200000,
                    key="test:batch256")
    print("tuned", {k: tune[k] for k in ("block_size", "blocks_per_cu", "claim", "ms")})
    assert (tune["block_size"], tune["blocks_per_cu"], tune["claim"]) in DEFAULT_CANDIDATES
    r = _rd(66, "fp32", "philox", block_size=1024, blocks_per_cu=1, claim=256)
    st = b.run(budgets, r)
    assert st["block_size"] == 1024 and st["photons"] == int(budgets.sum())
    l1 = ref = 0.0
    for k in range(256):
        gb, hb = b.read(k)
        sa, la, g, _, _ = built[k]
        sc = GpuScene(sa, la, g.detectors)
        sc.run(_rd(66, "fp32", "philox", n=int(budgets[k])))
        gs, hs = sc.read()
        sc.close()
        assert abs(int(hb[0]) - int(hs[0])) <= max(2, int(1e-3 * int(hs[0]))), k
        l1 += np.abs(gb - gs).sum()
        ref += np.abs(gs).sum()
    print("batch vs single L1 / L1:", l1 / ref)
    assert l1 <= 2e-3 * ref
    b.close()
