"""GPU Absorption grid (src/absortion.cpp:13-130, recorded at every vertex,
simulation.cpp:577-592) against the oracle: FP64 + xoshiro equal to rounding on
a time-resolved grid; FP32 production path conserves weight (absorbed +
reflected + transmitted = 1 per photon) and matches the oracle's grid
statistically; run_simulation_parallel fills SimConfig.absorption; gradient
runs fill it as forward runs do."""
import numpy as np
import pytest

import luminis_mc as lmc
import oracle
import stats
from luminis_mc.simulation import GpuScene, SimConfig, run_desc_from
from test_oracle_cpu import _absorbing_slab

pytestmark = pytest.mark.gpu


def _scene():
    b, ab = _absorbing_slab()
    return GpuScene(None, None, None, build=b), b, ab


def _rd(precision, rng, n, seed, off=0):
    c = SimConfig()
    c.seed, c.track_reverse_paths, c.precision, c.rng = seed, True, precision, rng
    return run_desc_from(c, n_photons=n, photon_offset=off)


def test_gpu_fp64_xoshiro_absorption_parity(require_gpu):
    sc, b, ab = _scene()
    n = 20000
    sc.run(_rd("fp64", "xoshiro", n, 51))
    gpu, hits = sc.read()
    cpu, chits, _, _ = oracle.run(b, n, 51)
    sl = slice(b.abs_offset, b.abs_offset + ab.buffer.size)
    g, c = gpu[sl], cpu[sl]
    rel = np.abs(g - c).max() / np.abs(c).max()
    print("absorption grid max rel diff", rel, "hits", list(hits), list(chits))
    assert np.abs(c).max() > 0
    assert rel < 2e-3
    sc.close()


def test_gpu_fp32_absorption_balance_and_statistics(require_gpu):
    sc, b, ab = _scene()
    N, B, seed = 160000, 16, 53
    g = stats.gpu_batches(sc, lambda n, off: _rd("fp32", "philox", n, seed, off), N, B)
    c = stats.oracle_batches(b, N, B, seed)
    sl = slice(b.abs_offset, b.abs_offset + ab.n_r * ab.n_z)  # time-integrated slice
    z, mask = stats.zscores(g[:, sl], c[:, sl])
    zz = z[mask]
    print("absorption bins compared", zz.size, "max|z|", np.abs(zz).max(), "frac |z|<3", np.mean(np.abs(zz) < 3))
    assert zz.size > 100
    assert np.abs(zz).max() < 5.0
    assert np.mean(np.abs(zz) < 3) > 0.98
    tot = g.sum(axis=0) / N
    R, T, A = tot[b.offsets[0][0]], tot[b.offsets[1][0]], tot[sl].sum()
    print("R", R, "T", T, "A", A, "sum", R + T + A)
    assert abs(R + T + A - 1.0) < 0.005
    sc.close()


def test_run_simulation_parallel_fills_absorption(require_gpu):
    b, ab = _absorbing_slab()
    import math
    import scenes
    m, ph = scenes.mie_medium(0.5e-3, 1.59, 1.33, 10.0, 1.0)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, 1.0)
    cfg = SimConfig()
    cfg.n_photons, cfg.seed = 20000, 3
    cfg.sample, cfg.laser = sample, lmc.Laser(1.0, 0.0, scenes.WAVELENGTH, 0.0, lmc.LaserSource.Point)
    g = lmc.SensorsGroup()
    g.add_detector(lmc.ReflectanceSensor(0.0, absorb=True))
    cfg.detector = g
    cfg.absorption = lmc.Absorption(20.0, 1.0, 0.5, 0.05)
    lmc.run_simulation_parallel(cfg)
    A = cfg.absorption.get_total_image(cfg.n_photons)
    assert A.rows == 2 * cfg.absorption.n_r and cfg.absorption.buffer.sum() > 0.5 * cfg.n_photons
    assert math.isfinite(float(np.asarray(A.data).sum()))


def test_absorption_in_adjoint_runs_equals_forward(require_gpu):
    """A gradient run fills the Absorption grid as a forward run does (the
    reference fills it for any configuration): FP64 xoshiro streams give the
    same photons in both, so every accumulator agrees to rounding; the grid's
    gradient rows stay zero (gradients cover the sensor grids)."""
    sc, b, ab = _scene()
    n, seed = 20000, 57
    grad, _ = sc.run_adjoint(_rd("fp64", "xoshiro", n, seed))
    adj, ahits = sc.read()
    sc.zero()
    sc.run(_rd("fp64", "xoshiro", n, seed))
    fwd, fhits = sc.read()
    sl = slice(b.abs_offset, b.abs_offset + ab.buffer.size)
    rel = np.abs(adj - fwd).max() / np.abs(fwd).max()
    print("adjoint vs forward max rel diff", rel, "hits", list(ahits), list(fhits))
    assert np.abs(fwd[sl]).max() > 0 and list(ahits) == list(fhits)
    assert rel < 1e-12
    assert not grad[:, sl].any() and grad[:, :sl.start].any()
    # the production path too: the FP32 gradient kernel records the grid (same
    # streams; its own compilation may round a few photons apart, so totals)
    sc.zero()
    sc.run_adjoint(_rd("fp32", "philox", n, seed))
    a32, _ = sc.read()
    sc.zero()
    sc.run(_rd("fp32", "philox", n, seed))
    f32, _ = sc.read()
    print("FP32 absorbed weight adjoint / forward", a32[sl].sum(), f32[sl].sum())
    assert abs(a32[sl].sum() - f32[sl].sum()) <= 1e-3 * f32[sl].sum()
    sc.close()
