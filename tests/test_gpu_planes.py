"""GPU checks of sensor planes away from the usual z = 0 plane, of large
grids, and of the planar field's phase on long paths.

* A non-absorbing planar fluence and a reflectance sensor on a layer
  interface (z = 0.5 mm of the three-media scene): the FP64 check mode equals
  the oracle; the FP32 production path (which places a photon exactly on the
  interface it crosses instead of 1e-12 past it) records each crossing once,
  so its hit counts match the FP64 build on the same streams.
* A far-field CBS grid of 512 KiB (beyond the 48 KiB the sensor kernel keeps
  in LDS, so its deposits take HBM FP64 atomics): FP64 trajectory parity.
* The planar field's phase k0 * (optical path) on paths of 1e6 - 2e7 rad, FP32
  production path against the FP64 build photon by photon: the wavenumber and
  the refractive index enter in FP64 and the path accumulates in FP64, so the
  median relative phase difference is what the FP32 step lengths leave
  (a few 1e-8: the hardware log2 behind -mfp ln u is biased by a fraction of
  an ulp), below the +8.2e-8 more an FP32 k0 and n would add at this
  wavelength.
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
    st32 = sc.run(_rd(cfg, "fp32", "philox", 200000))
    a32, h32 = sc.read()
    sc.zero()
    sc.run(_rd(cfg, "fp64", "philox", 200000))
    a64, h64 = sc.read()
    sc.close()
    print("hits fp32", list(h32), "fp64", list(h64))
    assert st32["aux_mode"] == 1  # a plane inside the sample: in-flight recording
    for k in (4, 5):  # the sensors on the interface
        assert abs(int(h32[k]) - int(h64[k])) <= 0.003 * int(h64[k]) + 10, k
    off, n = sc.build.offsets[5]
    assert abs(a32[off] - a64[off]) <= 0.003 * a64[off]


def test_surface_planes_launch_and_exit_crossings(require_gpu):
    """With every sensor plane on the surface (z = 0) and the laser at
    z = 1e-9 (laser.cpp:29-40) a flight crosses a plane only when it leaves
    the sample: the FP32 forward run takes the kernel with no plane code in
    its flight loop (aux_mode 0) and records exits at the wave's flush, and
    its hit counts, first time bin and crossing count equal the FP64 build's
    (in-flight recording) on the same streams.  A laser moved onto the plane
    (z = 0) makes the first segment cross it (the reference records that
    crossing, t = 0; here the absorbing CBS sensor, which has no direction
    filter, takes every photon at launch): such a scene keeps the in-flight
    recording kernel (aux_mode 1) in FP32 too."""
    N = 200000
    from luminis_mc.simulation import SceneBuild
    for z0 in (None, 0.0):
        cfg, _ = scenes.config_multilayer(N, seed=45, small=True)
        b = SceneBuild(cfg.sample, cfg.laser, cfg.detector.detectors)
        if z0 is not None:
            b.desc.laser.pos[2] = z0
        sc = GpuScene.from_build(b)
        st32 = sc.run(_rd(cfg, "fp32", "philox", N))
        a32, h32 = sc.read()
        sc.zero()
        st64 = sc.run(_rd(cfg, "fp64", "philox", N))
        a64, h64 = sc.read()
        sc.close()
        print("laser z", b.desc.laser.pos[2], "aux", st32["aux_mode"], st64["aux_mode"], "hits fp32", list(h32),
              "fp64", list(h64), "crossings", st32["crossings"], st64["crossings"])
        assert st64["aux_mode"] == 1
        if z0 is None:
            assert st32["aux_mode"] == 0
            assert 0 < int(h64[1]) < N  # exits only
        else:
            # every photon absorbed at launch (the fluence sensor's time bin of
            # a launch, path length ~ +-1e-17, is a rounding artifact)
            assert st32["aux_mode"] == 1
            for st, h in ((st32, h32), (st64, h64)):
                assert int(st["crossings"]) == N and int(st["events"]) == 0 and int(h[2]) == int(h[3]) == 0
            continue
        for k in range(4):
            assert abs(int(h32[k]) - int(h64[k])) <= 0.003 * int(h64[k]) + 10, k
        off, n = b.offsets[1]
        g32, g64 = a32[off:off + n].reshape(4, 9, 16, 16), a64[off:off + n].reshape(4, 9, 16, 16)
        assert abs(g32[0, 1].sum() - g64[0, 1].sum()) <= 0.003 * g64[0, 1].sum()
        assert abs(int(st32["crossings"]) - int(st64["crossings"])) <= 0.003 * int(st64["crossings"])


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


# wavelength (17 nm) where FP32 rounds k0 = 2 pi / lambda up by 4.2e-8 (and
# n = 1.33 up by 3.2e-8): an FP32 k0 and n would add 7.4e-8 of every phase
WL_FIELD = 1.6994495e-05


def test_field_phase_fp32_vs_fp64_on_single_scattering(require_gpu):
    """Photons scattered once (HG g = 0.5, mu_s 1 /mm) carry phases k0 n L of
    about 1e6 rad at this wavelength and their FP32 and FP64 trajectories stay
    close (one free path, one FP32 polar angle), so the per-photon phase
    difference FP32 - FP64 is beta * phase plus small noise.  8 x 50000 photons,
    one launch per batch and precision; a 1024^2 field over 8 mm and a
    PhotonRecordSensor pick out the pixels that hold exactly one photon in
    both runs (the same photon: one event, same first-scatter point).  With k0
    and n in FP64, beta is what the FP32 free-path logarithm leaves (the
    hardware log2 lengthens paths by ~4e-8 of their length); an FP32 k0 and n
    would add 7.4e-8 (measured on 12614 photons: 3.3e-8 against 1.04e-7, each
    +- 7.5e-9)."""
    n = 1.33
    hg = lmc.HenyeyGreensteinPhaseFunction(0.5)
    m = lmc.RGDMedium(hg, 0.1e-3, 1.59, n, WL_FIELD)
    m.set_scattering_coefficient(1.0)
    m.set_absorption_coefficient(0.0)
    m.set_mean_free_path(1.0)
    sample = lmc.Sample(n)
    sample.add_layer(m, 0.0, math.inf)
    laser = lmc.Laser(1.0, 0.0, WL_FIELD, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    N_PIX, LEN = 1024, 8.0
    g.add_detector(lmc.PlanarFieldSensor(0.0, LEN, LEN, LEN / N_PIX, LEN / N_PIX, absorb=False))
    g.add_detector(lmc.PhotonRecordSensor(0.0, absorb=True))
    sc = GpuScene(sample, laser, g.detectors)
    cfg = lmc.SimConfig()
    cfg.seed, cfg.track_reverse_paths = 44, True
    k0 = 2 * math.pi / WL_FIELD
    d, phi = [], []
    for batch in range(8):
        runs = {}
        for prec in ("fp32", "fp64"):
            sc.zero()
            sc.run(_rd(cfg, prec, "philox", 50000, off=batch * 50000))
            out, hits = sc.read()
            rec = sc.records()
            f = out[:4 * N_PIX * N_PIX].reshape(2, N_PIX, N_PIX, 2)
            pix = np.floor((rec[:, 14:16] + 0.5 * LEN) / (LEN / N_PIX)).astype(np.int64)
            inside = np.all((pix >= 0) & (pix < N_PIX), axis=1)
            key = np.where(inside, pix[:, 0] * N_PIX + pix[:, 1], -1 - np.arange(len(rec)))  # outside: never shared
            runs[prec] = (f[0, :, :, 0] + 1j * f[0, :, :, 1], rec, key)
        (f32, r32, k32), (f64, r64, k64) = runs["fp32"], runs["fp64"]
        u32, c32 = np.unique(k32, return_counts=True)
        u64, c64 = np.unique(k64, return_counts=True)
        single = np.intersect1d(u32[c32 == 1], u64[c64 == 1])
        single = single[single >= 0]
        i32 = {int(k): i for i, k in enumerate(k32)}
        i64 = {int(k): i for i, k in enumerate(k64)}
        for k in single:
            a, b = r32[i32[int(k)]], r64[i64[int(k)]]
            if a[1] != 1 or b[1] != 1 or np.abs(a[8:11] - b[8:11]).max() > 1e-5:
                continue
            x, y = divmod(int(k), N_PIX)
            if f64[x, y] == 0:
                continue
            d.append(np.angle(f32[x, y] * np.conj(f64[x, y])))
            phi.append(k0 * n * b[5])
    sc.close()
    d, phi = np.array(d), np.array(phi)
    beta = (d * phi).sum() / (phi * phi).sum()
    se = math.sqrt(((d - beta * phi) ** 2 * phi ** 2).sum()) / (phi * phi).sum()
    print(f"{d.size} single-scattered photons alone in their pixel, phase {np.median(phi):.3g} (median): "
          f"slope {beta:.3e} +- {se:.1e}; |d| median {np.median(np.abs(d)):.3f} rad")
    assert d.size > 10000
    assert se < 1e-8 and beta < 6.5e-8
