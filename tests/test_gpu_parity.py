"""GPU parity: the HIP transport path against the oracle.

* RNG: device Philox4x32-10 against Random123's known answers.
* FP64 + xoshiro mode: the device follows the reference's per-photon streams
  and draw order, so every accumulator must match the oracle to rounding
  (trajectory-level parity) -- plain Mie CBS and the three-media multilayer
  scene with fluence, far-field fluence, reflectance and field sensors.
* FP32 production path (Philox-seeded xoshiro128** streams): statistically identical to the oracle,
  per-bin z-scores from batch means (|z| < 5 everywhere, >= 98 % within 3 sigma),
  and close to the FP64 build on the same streams.
"""
import ctypes
import math

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import scenes
import stats
from luminis_mc import _lib
from luminis_mc.simulation import GpuScene, SceneBuild, run_desc_from
from parity import assert_trajectory_parity
from test_oracle_cpu import PHILOX_KAT

pytestmark = pytest.mark.gpu


def _scene(cfg):
    return GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)


def _run_desc(cfg, precision, rng, n, off=0):
    cfg.precision, cfg.rng = precision, rng
    return run_desc_from(cfg, n_photons=n, photon_offset=off)


def test_device_info(require_gpu):
    info = lmc.device_info()
    print("device:", info)
    assert info["n_cu"] > 0 and info["name"].startswith("gfx")


def test_gpu_philox_known_answers(require_gpu):
    rng = np.random.default_rng(1)
    ctrs = [c for c, _, _ in PHILOX_KAT]
    keys = [k for _, k, _ in PHILOX_KAT]
    L = _lib.lib()
    for (ctr, key, expect) in PHILOX_KAT:
        c = (ctypes.c_uint32 * 4)(*ctr)
        o = (ctypes.c_uint32 * 4)()
        _lib.check(L.lmc_probe_philox(1, c, key[0], key[1], o))
        assert tuple(o) == expect
    n = 1024
    cs = rng.integers(0, 2 ** 32, size=(n, 4), dtype=np.uint64).astype(np.uint32)
    out = np.zeros((n, 4), dtype=np.uint32)
    P = ctypes.POINTER(ctypes.c_uint32)
    _lib.check(L.lmc_probe_philox(n, cs.ctypes.data_as(P), 0xDEADBEEF, 0x12345678, out.ctypes.data_as(P)))
    for i in range(0, n, 97):
        assert list(out[i]) == oracle.philox(cs[i], 0xDEADBEEF, 0x12345678)


def _device_stream(k0, k1, photon, n):
    out = np.zeros(n, dtype=np.uint32)
    _lib.check(_lib.lib().lmc_probe_stream(k0, k1, photon, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
               "lmc_probe_stream")
    return out


def test_gpu_stream_matches_restatement(require_gpu):
    """A photon's production stream: Philox4x32-10 of (photon low, photon high,
    0xFFFFFFFE, 0) under the seed as the xoshiro128** state, then its outputs;
    the device against the host Philox and the host restatement."""
    import rng_check
    for (k0, k1) in ((0, 0), (0xDEADBEEF, 0x12345678)):
        for photon in (0, 1, 12345, (1 << 32) + 7):
            st = oracle.philox(np.array([photon & 0xFFFFFFFF, photon >> 32, 0xFFFFFFFE, 0], dtype=np.uint32), k0, k1)
            if not any(st):
                st[0] = 1
            want = rng_check.xoshiro128ss(st, 64)
            assert list(_device_stream(k0, k1, photon, 64)) == want, (k0, k1, photon)


def test_gpu_stream_uniformity(require_gpu):
    """2^20 outputs of one photon's stream: top bytes and 24-bit uniforms
    uniform (chi-square over 256 bins within 5 sigma), each bit set half the
    time, and no lag-1 correlation."""
    x = _device_stream(0x9E3779B9, 7, 42, 1 << 20).astype(np.int64)
    n = x.size
    counts = np.bincount(x >> 24, minlength=256)
    chi2 = ((counts - n / 256) ** 2 / (n / 256)).sum()
    assert abs(chi2 - 255) < 5 * np.sqrt(2 * 255), chi2
    bits = ((x[:, None] >> np.arange(32)) & 1).mean(axis=0)
    assert np.all(np.abs(bits - 0.5) < 5 * 0.5 / np.sqrt(n)), bits
    u = ((x >> 8) + 0.5) / 2 ** 24
    r = np.corrcoef(u[:-1], u[1:])[0, 1]
    assert abs(r) < 5 / np.sqrt(n), r


def test_gpu_fp64_xoshiro_trajectory_parity_plain_mie(require_gpu):
    cfg, det = scenes.config_plain_mie(20000, seed=3)
    sc = _scene(cfg)
    st = sc.run(_run_desc(cfg, "fp64", "xoshiro", 20000))
    gpu, hits = sc.read()
    assert st["photons"] == 20000
    assert_trajectory_parity(st, gpu, hits, oracle.run(sc.build, 20000, 3), sc.build.offsets, label="plain")


def _time_resolved_scene(n_photons, seed):
    """Plain Mie medium with a time-resolved CBS sensor (16 x 4 angles, 8 time bins)
    and a time-resolved planar fluence (16 x 16, 8 time bins) at z = 0."""
    cfg, _ = scenes.config_plain_mie(n_photons, seed=seed)
    len_t = 100.0 * 0.299792458 / 1.33
    group = lmc.SensorsGroup()
    group.add_detector(lmc.FarFieldCBSSensor(0.1, 2 * math.pi, len_t, 0.1 / 16, 2 * math.pi / 4, len_t / 8))
    group.add_detector(lmc.PlanarFluenceSensor(0.0, 10.0, 10.0, len_t, 10.0 / 16, 10.0 / 16, len_t / 8,
                                               absorb=False))
    cfg.detector = group
    return cfg


def _bin0_minus_sum(block, n_q, n_t):
    g = block.reshape(n_q, n_t, -1)
    return np.abs(g[:, 0] - g[:, 1:].sum(axis=1)).max() / max(np.abs(g[:, 0]).max(), 1e-300)


def test_gpu_fp64_xoshiro_parity_time_resolved(require_gpu):
    """Time-resolved CBS and planar fluence: the device deposits each hit in its
    time bin and completes the time-integrated bin 0 at read-out; the oracle
    deposits twice as the reference does (detector.cpp:1223-1225, :829-840)."""
    cfg = _time_resolved_scene(20000, 5)
    sc = _scene(cfg)
    st = sc.run(_run_desc(cfg, "fp64", "xoshiro", 20000))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, 20000, 5)
    cpu = cpu_run[0]
    (o_cbs, n_cbs), (o_flu, n_flu) = sc.build.offsets[:2]
    n_t = n_cbs // (8 * 16 * 4)
    assert n_t == n_flu // (4 * 16 * 16) and n_t > 2
    for arr in (gpu, cpu):
        assert _bin0_minus_sum(arr[o_cbs:o_cbs + n_cbs], 8, n_t) < 1e-9
        assert _bin0_minus_sum(arr[o_flu:o_flu + n_flu], 4, n_t) < 1e-9
    assert np.abs(gpu[o_cbs:o_cbs + n_cbs]).max() > 0 and np.abs(gpu[o_flu:o_flu + n_flu]).max() > 0
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets, ["cbs", "fluence"], label="time-resolved")
    # a second run accumulates: bin 0 of the read-out stays the sum of the time bins
    sc.run(_run_desc(cfg, "fp64", "xoshiro", 20000, off=20000))
    gpu2, _ = sc.read()
    assert _bin0_minus_sum(gpu2[o_cbs:o_cbs + n_cbs], 8, n_t) < 1e-9
    assert _bin0_minus_sum(gpu2[o_flu:o_flu + n_flu], 4, n_t) < 1e-9
    sc.close()


def test_gpu_fp64_xoshiro_trajectory_parity_multilayer(require_gpu):
    cfg, d = scenes.config_multilayer(20000, seed=11, small=True, with_field=True)
    sc = _scene(cfg)
    st = sc.run(_run_desc(cfg, "fp64", "xoshiro", 20000))
    gpu, hits = sc.read()
    assert_trajectory_parity(st, gpu, hits, oracle.run(sc.build, 20000, 11), sc.build.offsets,
                             ["cbs", "fluence", "ff", "refl", "field"], label="multilayer")


def test_gpu_fp32_statistical_parity_plain_mie(require_gpu):
    N, B, seed = 400000, 16, 5
    cfg, det = scenes.config_plain_mie(N, seed=seed)
    sc = _scene(cfg)
    g = stats.gpu_batches(sc, lambda n, off: _run_desc(cfg, "fp32", "philox", n, off), N, B)
    c = stats.oracle_batches(sc.build, N, B, seed)
    z, mask = stats.zscores(g, c)
    zz = z[mask]
    frac3 = np.mean(np.abs(zz) < 3)
    print("bins compared", zz.size, "max|z|", np.abs(zz).max(), "frac |z|<3", frac3)
    assert zz.size > 400
    assert np.abs(zz).max() < 5.0
    assert frac3 > 0.98


def test_gpu_fp32_statistical_parity_multilayer(require_gpu):
    N, B, seed = 160000, 16, 9
    cfg, d = scenes.config_multilayer(N, seed=seed, small=True, with_field=False)
    sc = _scene(cfg)
    g = stats.gpu_batches(sc, lambda n, off: _run_desc(cfg, "fp32", "philox", n, off), N, B)
    c = stats.oracle_batches(sc.build, N, B, seed)
    for (off, n), name in zip(sc.build.offsets, ["cbs", "fluence", "ff", "refl"]):
        z, mask = stats.zscores(g[:, off:off + n], c[:, off:off + n])
        zz = z[mask]
        print(name, "bins", zz.size, "max|z|", np.abs(zz).max() if zz.size else 0,
              "frac<3", np.mean(np.abs(zz) < 3) if zz.size else 1)
        if zz.size:
            assert np.abs(zz).max() < 5.5, name
            assert np.mean(np.abs(zz) < 3) > 0.97, name


def test_gpu_fp32_close_to_fp64_on_same_streams(require_gpu):
    N = 200000
    cfg, det = scenes.config_plain_mie(N, seed=21)
    sc = _scene(cfg)
    sc.run(_run_desc(cfg, "fp32", "philox", N))
    a32, _ = sc.read()
    sc.zero()
    sc.run(_run_desc(cfg, "fp64", "philox", N))
    a64, _ = sc.read()
    g32 = a32.reshape(8, -1)
    g64 = a64.reshape(8, -1)
    # same random numbers: only rounding-driven branch flips separate the two
    rel = np.abs(g32[[0, 4]] - g64[[0, 4]]).sum() / np.abs(g64[[0, 4]]).sum()
    print("FP32 vs FP64 (same streams) relative L1 difference of S0 coh/incoh:", rel)
    assert rel < 0.05


def test_run_simulation_parallel_accumulates(require_gpu):
    cfg, det = scenes.config_plain_mie(20000, seed=1)
    lmc.run_simulation_parallel(cfg)
    first = np.array(det.S0_incoh[0]).copy()
    h1 = det.hits
    lmc.run_simulation_parallel(cfg)
    assert det.hits == 2 * h1
    assert np.allclose(np.array(det.S0_incoh[0]), 2 * first, rtol=1e-12)
    res = lmc.postprocess_farfield_cbs(det, 40000)
    assert np.all(np.isfinite(np.array(res.coherent[0].S0)))
    assert cfg.last_stats["photons"] == 20000


def test_chunked_runs_equal_one_run(require_gpu):
    cfg, det = scenes.config_plain_mie(10000, seed=4)
    sc = _scene(cfg)
    sc.run(_run_desc(cfg, "fp64", "xoshiro", 10000, 0))
    one, h1 = sc.read()
    sc.zero()
    sc.run(_run_desc(cfg, "fp64", "xoshiro", 4000, 0))
    sc.run(_run_desc(cfg, "fp64", "xoshiro", 6000, 4000))
    two, h2 = sc.read()
    assert int(h1[0]) == int(h2[0])
    assert np.allclose(one, two, rtol=1e-12, atol=1e-15)


def test_gpu_fp64_xoshiro_parity_steady_fluence_first(require_gpu):
    """A steady-state planar fluence as the first sensor (accumulators at offset
    0: its grid row must be row 0, nothing before it) beside an analog CBS."""
    cfg, _ = scenes.config_plain_mie(1, seed=13)
    group = lmc.SensorsGroup()
    group.add_detector(lmc.PlanarFluenceSensor(0.0, 4.0, 4.0, 0.0, 0.25, 0.25, 0.0, absorb=False))
    group.add_detector(lmc.FarFieldCBSSensor(0.1, 6.28, 0.0, 0.1 / 20, 6.28, 0.0))
    cfg.detector = group
    sc = _scene(cfg)
    st = sc.run(_run_desc(cfg, "fp64", "xoshiro", 20000))
    gpu, hits = sc.read()
    cpu_run = oracle.run(sc.build, 20000, 13)
    assert np.abs(cpu_run[0][:sc.build.offsets[0][1]]).max() > 0
    assert_trajectory_parity(st, gpu, hits, cpu_run, sc.build.offsets, ["fluence", "cbs"], label="steady")
    sc.close()
