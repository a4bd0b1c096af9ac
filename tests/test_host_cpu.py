"""Host-side logic of the library (CPU only): C ABI exports, table builders
against the oracle and the reference MIEV0 build, API error behaviour."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import luminis_mc as lmc
import oracle
from luminis_mc import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "luminis_hip.h")).read()
    names = set(re.findall(r"\b(lmc_[a-z0-9_]+)\s*\(", hdr))
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert L.lmc_abi_version() == 2


def test_ctypes_structs_match_c_layout(tmp_path):
    """Compile a probe against include/luminis_hip.h and compare sizeof/offsetof
    with the ctypes mirrors field by field."""
    structs = {"lmc_species_desc": _lib.SpeciesDesc, "lmc_layer_desc": _lib.LayerDesc,
               "lmc_sensor_desc": _lib.SensorDesc, "lmc_laser_desc": _lib.LaserDesc,
               "lmc_scene_desc": _lib.SceneDesc, "lmc_run_desc": _lib.RunDesc, "lmc_stats": _lib.Stats,
               "lmc_device_info": _lib.DeviceInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/luminis_hip.h"',
             'int main(void) {']
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('return 0; }')
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    import subprocess
    subprocess.check_call(["gcc", "-O0", "-o", str(exe), str(src)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, cls in structs.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"


@pytest.mark.parametrize("x,m", [(10.0, complex(1.5, -0.1)), (4.96, complex(1.59 / 1.33, 0)),
                                 (6.60, complex(1.59 / 1.33, 0)), (19.8, complex(1.2, 0)),
                                 (0.5, complex(1.1, 0)), (30.0, complex(1.33, 0.01)), (2.0, complex(1.45 / 1.33, 0))])
def test_library_mie_matches_oracle_and_reference(x, m):
    mu = np.cos(np.linspace(0, np.pi, 37))
    s1, s2 = lmc.mie_amplitudes(x, m, mu)
    o1, o2, qe, qs, gq = oracle.mie(x, m, mu)
    scale = max(np.abs(o1).max(), np.abs(o2).max())
    assert np.max(np.abs(s1 - o1)) < 1e-10 * scale
    assert np.max(np.abs(s2 - o2)) < 1e-10 * scale
    lqe, lqs, lg = lmc.mie_info(x, m)
    assert math.isclose(lqe, qe, rel_tol=1e-10) and math.isclose(lqs, qs, rel_tol=1e-10)
    assert math.isclose(lg, gq / qs, rel_tol=1e-9)
    if oracle.ref_available():
        r1, r2 = oracle.ref_miev(x, m, mu)
        assert np.max(np.abs(s1 - r1)) < 1e-10 * scale
        assert np.max(np.abs(s2 - r2)) < 1e-10 * scale


def test_library_mie_testmi_known_answers():
    s1, s2 = lmc.mie_amplitudes(10.0, complex(1.5, -0.1), [-0.7660444])
    assert abs(s1[0] - complex(-0.1548380, -1.128972)) < 1e-4 * abs(complex(-0.1548380, -1.128972))
    assert abs(s2[0] - complex(0.05669755, 0.5425681)) < 1e-4 * abs(complex(0.05669755, 0.5425681))
    qe, qs, g = lmc.mie_info(10.0, complex(1.5, -0.1))
    assert math.isclose(qe, 2.459791, rel_tol=1e-4) and math.isclose(qs, 1.235144, rel_tol=1e-4)
    assert math.isclose(g * qs, 1.139235, rel_tol=1e-4)


def test_sampling_table_matches_oracle():
    ph = lmc.MiePhaseFunction(0.633e-3, 0.5e-3, 1.59, 1.33, 2000, 0.0, math.pi)
    nodes = ph.table.values
    pdf = ph.pdf(nodes)
    assert np.array_equal(ph.table.cdf, oracle.sampling_cdf(2000, 0.0, math.pi, pdf))
    assert ph.table.cdf[-1] == 1.0 and np.all(np.diff(ph.table.cdf) >= 0)


def test_hg_inverse_cdf_moment():
    # <cos theta> = g for Henyey-Greenstein (closed form of phase.cpp:108-111)
    hg = lmc.HenyeyGreensteinPhaseFunction(0.8)
    u = (np.arange(200000) + 0.5) / 200000
    c = np.array([hg.sample_cos(v) for v in u[::50]])
    assert abs(c.mean() - 0.8) < 2e-3


def test_form_factor_limits():
    assert lmc.form_factor(0.0, 1000.0, 1e-4) == 1.0
    # small-u series 3(sin u - u cos u)/u^3 = 1 - u^2/10 + u^4/280 - ...
    k, r, th = 10.0, 1e-3, 1.0
    u = 2 * k * r * math.sin(th / 2)
    assert math.isclose(lmc.form_factor(th, k, r), 1 - u * u / 10 + u ** 4 / 280, rel_tol=1e-9)


def test_sensor_bin_counts_follow_widths():
    d = lmc.FarFieldCBSSensor(0.1, 6.28, 0.0, 0.001, 6.28, 0.0)
    assert (d.N_theta, d.N_phi, d.N_t) == (100, 1, 1)
    f = lmc.PlanarFluenceSensor(0.0, 10, 10, 1.0, 0.5, 0.5, 0.25)
    assert (f.N_x, f.N_y, f.N_t) == (20, 20, 5)


def test_config_validation_errors():
    cfg = lmc.SimConfig()
    with pytest.raises(ValueError, match="sample"):
        lmc.run_simulation_parallel(cfg)
    cfg.sample = lmc.Sample(1.33)
    with pytest.raises(ValueError, match="laser"):
        lmc.run_simulation_parallel(cfg)
    cfg.laser = lmc.Laser(1, 0, 0.633e-3, 0.0, lmc.LaserSource.Point)
    with pytest.raises(ValueError, match="detector"):
        lmc.run_simulation_parallel(cfg)


def test_sample_contiguity_errors():
    ph = lmc.HenyeyGreensteinPhaseFunction(0.5)
    m = lmc.RGDMedium(ph, 1e-4, 1.59, 1.33, 0.633e-3)
    s = lmc.Sample(1.33)
    with pytest.raises(ValueError, match="start at z = 0"):
        s.add_layer(m, 0.1, 1.0)
    s.add_layer(m, 0.0, 1.0)
    with pytest.raises(ValueError, match="does not match"):
        s.add_layer(m, 1.5, 2.0)
    s.add_layer(m, 1.0, math.inf)
    with pytest.raises(ValueError, match="infinite"):
        s.add_layer(m, 5.0, 6.0)


def test_mixture_validation_errors():
    ph = lmc.HenyeyGreensteinPhaseFunction(0.5)
    m = lmc.RGDMedium(ph, 1e-4, 1.59, 1.33, 0.633e-3)
    with pytest.raises(ValueError):
        lmc.MixtureLayer([], [], 0.0, 1.0)
    with pytest.raises(ValueError):
        lmc.MixtureLayer([m], [1.0, 2.0], 0.0, 1.0)
    with pytest.raises(NotImplementedError):
        lmc.MixtureLayer([m], [1.0], 0.0, 1.0)  # HG exposes no cross-section


def test_unset_mean_free_path_is_an_error():
    import scenes
    cfg, det = scenes.config_plain_mie(10)
    cfg.sample.layers[0].medium.mean_free_path = float("nan")
    from luminis_mc.simulation import SceneBuild
    with pytest.raises(ValueError, match="mean free path"):
        SceneBuild(cfg.sample, cfg.laser, cfg.detector.detectors)


def test_oracle_runs_plain_mie_and_is_thread_count_invariant():
    import scenes
    from luminis_mc.simulation import SceneBuild
    cfg, det = scenes.config_plain_mie(2000)
    b = SceneBuild(cfg.sample, cfg.laser, cfg.detector.detectors)
    a1, h1, e1, t1 = oracle.run(b, 2000, 0, n_threads=1)
    a4, h4, e4, t4 = oracle.run(b, 2000, 0, n_threads=4)
    # per-photon reseeding: identical streams; sums differ only by merge order
    assert e1 == e4 and t1 == t4 and int(h1[0]) == int(h4[0])
    assert np.allclose(a1, a4, rtol=1e-12, atol=0)
    g = a1.reshape(8, 1, det.N_theta, det.N_phi)
    assert np.all(g[0] >= 0) and np.all(g[4] >= 0)
    # single scatter photons deposit coh == incoh; multiple ones bound coh <= 2 incoh
    assert np.all(g[0] <= 2 * g[4] + 1e-12)


def test_autograd_function_chain_rule_without_device(monkeypatch):
    """The autograd wrapper returns J^T g (CPU check with a stand-in evaluate)."""
    import torch
    import scenes
    sample, laser, sensors, keep = scenes.scene_adjoint()
    tr = lmc.DifferentiableTransport(sample, laser, sensors, n_photons=10)
    rng = np.random.default_rng(3)
    vals = rng.normal(size=tr.n_out)
    jac = rng.normal(size=(tr.n_layers, 3, tr.n_out))
    monkeypatch.setattr(tr, "evaluate", lambda th, with_grad=True, **kw: (vals, jac if with_grad else None))
    theta = torch.tensor(tr.theta0(), requires_grad=True)
    out = tr(theta)
    g = torch.tensor(rng.normal(size=tr.n_out))
    (out * g).sum().backward()
    assert np.allclose(theta.grad.numpy(), np.einsum("lpd,d->lp", jac, g.numpy()))
    assert np.allclose(out.detach().numpy(), vals)


def test_param_build_rewrites_coefficients():
    import scenes
    from luminis_mc.adjoint import _ParamBuild
    sample, laser, sensors, keep = scenes.scene_adjoint()
    tr = lmc.DifferentiableTransport(sample, laser, sensors, n_photons=10)
    th = tr.theta0()
    assert np.allclose(th, [[15.0, 0.01, 0.0], [8.0, 0.05, 0.8]])
    th2 = th * np.array([[1.1, 2.0, 1.0], [0.9, 0.5, 1.05]])
    b = _ParamBuild(tr.base, th2)
    for l in range(2):
        assert b.layers[l].mu_s == th2[l, 0] and b.layers[l].mu_a == th2[l, 1]
        assert b.layers[l].mfp == pytest.approx(1.0 / (th2[l, 0] + th2[l, 1]))
    assert b.species[1].hg_g == pytest.approx(0.84)
    assert tr.base.layers[0].mu_s == 15.0  # the template is untouched
    with pytest.raises(ValueError):
        _ParamBuild(tr.base, np.array([[15.0, 0.01, 0.0], [8.0, 0.05, 1.5]]))


def _azimuth_numpy(abc, words, mistake=None):
    """Restatement of the device azimuth sampler with optional planted mistakes."""
    a, b, c = abc
    M, A, B = 0.5 * (a + b), 0.5 * (a - b), 0.5 * c
    Rm = np.hypot(A, B)
    t = Rm + abs(A)
    c0, s0 = (1.0, 0.0) if Rm == 0 else ((t, B) if A >= 0 else (B, t)) / np.sqrt(2 * Rm * t)
    if mistake == "axis":  # minor instead of principal axis
        c0, s0 = -s0, c0
    v = ((words[:, 0] >> 8).astype(np.float64) + 0.5) / 2.0 ** 24
    lobe = v * M < (Rm if mistake != "weight" else Rm * Rm / M)
    with np.errstate(divide="ignore", invalid="ignore"):
        r2 = np.where(lobe, v * M / Rm, 1.0)
    r = r2 if mistake == "radius" else np.sqrt(r2)
    ang = 2 * np.pi * ((words[:, 1] >> 9).astype(np.float64) + 0.5) / 2.0 ** 23
    px, py = np.where(lobe, 1.0, 0.0) + r * np.cos(ang), r * np.sin(ang)
    nrm = 1 / np.hypot(px, py)
    if mistake != "half":
        nrm = np.where(words[:, 1] & 0x100, -nrm, nrm)
    cx, sx = px * nrm, py * nrm
    return c0 * cx - s0 * sx, s0 * cx + c0 * sx


def test_azimuth_lattice_check_is_sensitive():
    """The lattice CDF check applied to the device sampler passes the exact
    mixture sampler and fails each plausible slip (wrong axis, wrong mixture
    weight, radius not square-rooted, half period ignored)."""
    import azimuth_check as ac
    forms = ac.forms(np.random.default_rng(11), 12)
    words = ac.lattice_words()
    worst = max(ac.cdf_discrepancy(f, *_azimuth_numpy(f, words)) for f in forms)
    assert worst < ac.CDF_TOL / 4
    for mistake in ("axis", "weight", "radius", "half"):
        bad = max(ac.cdf_discrepancy(f, *_azimuth_numpy(f, words, mistake)) for f in forms)
        assert bad > 5 * ac.CDF_TOL, mistake



def test_param_build_gives_shared_medium_one_species_per_layer():
    """Two layers built from one medium object share a species in the scene
    build; the gradient build gives each layer its own copy so each layer's g
    (and mu_s, mu_a) is its own."""
    from luminis_mc.adjoint import _ParamBuild
    ph = lmc.HenyeyGreensteinPhaseFunction(0.7)
    m = lmc.RGDMedium(ph, 0.1e-3, 1.59, 1.33, 0.633e-3)
    m.set_scattering_coefficient(10.0)
    m.set_absorption_coefficient(0.01)
    m.set_mean_free_path(1.0 / 10.01)
    sample = lmc.Sample(1.33)
    sample.add_layer(m, 0.0, 0.5)
    sample.add_layer(m, 0.5, math.inf)
    laser = lmc.Laser(1.0, 0.0, 0.633e-3, 0.0, lmc.LaserSource.Point)
    group = lmc.SensorsGroup()
    group.add_detector(lmc.ReflectanceSensor(0.0, absorb=False))
    tr = lmc.DifferentiableTransport(sample, laser, group.detectors, n_photons=10)
    assert tr.base.layers[0].species_first == tr.base.layers[1].species_first
    th = np.array([[10.0, 0.01, 0.6], [12.0, 0.02, 0.8]])
    b = _ParamBuild(tr.base, th)
    s0, s1 = b.layers[0].species_first, b.layers[1].species_first
    assert s0 != s1 and b.desc.n_species == tr.base.desc.n_species + 1
    assert b.species[s0].hg_g == 0.6 and b.species[s1].hg_g == 0.8
    assert b.species[s0].mu_s == 10.0 and b.species[s1].mu_s == 12.0
    assert b.species[s1].s1_offset == b.species[s0].s1_offset  # tables shared
    assert tr.base.species[s0].hg_g == 0.7  # the template is untouched


def test_lognormal_mixture_weights_and_total_scattering():
    """Bin weights are the lognormal probabilities of equal ln r bins over
    +-span sigma (closed form by erf), renormalised; number densities give
    the requested total mu_s and mu_a through the MixtureLayer sums."""
    K, med, s_ln, span = 8, 0.4e-3, 0.3, 3.0
    sp, dens, w = lmc.lognormal_mie_mixture(0.633e-3, med, s_ln, 1.59, 1.33, K, 30.0, 0.002, span=span,
                                            n_div=400, n_table=200)
    assert len(sp) == len(dens) == len(w) == K
    assert sum(w) == pytest.approx(1.0, abs=1e-15)
    # symmetric about the median in ln r, central bins heaviest
    assert np.allclose(w, w[::-1], rtol=1e-12)
    z = np.linspace(-span, span, K + 1)
    from scipy.stats import norm as _norm
    pk = np.diff(_norm.cdf(z))
    assert np.allclose(w, pk / pk.sum(), rtol=1e-12)
    radii = np.array([m.radius for m in sp])
    assert np.allclose(np.log(radii), np.log(med) + 0.5 * (z[:-1] + z[1:]) * s_ln, rtol=0, atol=1e-12)
    layer = lmc.MixtureLayer(sp, dens, 0.0, math.inf)
    assert layer.mu_scattering() == pytest.approx(30.0, rel=1e-12)
    assert layer.mu_absorption() == pytest.approx(0.002, rel=1e-12)
    with pytest.raises(ValueError):
        lmc.lognormal_mie_mixture(0.633e-3, med, 0.0, 1.59, 1.33, K, 30.0, 0.002)


def test_reference_public_names_are_exported():
    """Every transport-surface name of the reference package
    (python/luminis_mc/__init__.py:1-76) and its persistence names (:77-89,
    the HDF5 manager, typed records and schema registry) import from
    luminis_mc; the MCMC mean-free-path helpers and the sweep manifest are not
    on the transport path."""
    names = """Rng Vec3 Vec2 CVec2 Matrix CMatrix dot cross norm matcmul matcmulscalar calculate_rotation_angle
    PhaseFunction UniformPhaseFunction RayleighPhaseFunction HenyeyGreensteinPhaseFunction RayleighDebyePhaseFunction
    RayleighDebyeEMCPhaseFunction DrainePhaseFunction MiePhaseFunction form_factor Photon PhotonRecord LaserSource
    TemporalProfile Laser uniform_distribution gaussian_distribution Sensor PhotonRecordSensor PlanarFieldSensor
    PlanarFluenceSensor FarFieldCBSSensor StatisticsSensor SensorsGroup CrossingDirection StokesMatrixProcessed
    StokesRadialProcessed FarFieldCBSProcessed FarFieldCBSRadialProcessed PlanarFluenceProcessed
    PlanarFieldProcessed postprocess_farfield_cbs postprocess_planar_fluence postprocess_planar_field
    ScatteringMedium RGDMedium MieMedium Layer SampleLayer MixtureLayer Sample Absorption SimConfig
    run_simulation_parallel LogLevel set_log_level Experiment ResultsLoader capture_params derived_quantities
    derived_quantities_mixture SimParams RunParams LaserParams MediumParams LayerParams SensorMeta FarFieldCBSResult
    PlanarFluenceResult PlanarFieldResult StatisticsResult PhotonRecordsResult AbsorptionResult
    FarFieldCBSProcessedResult PlanarFluenceProcessedResult PlanarFieldProcessedResult SensorSchema
    SENSOR_SCHEMAS""".split()
    missing = [n for n in names if not hasattr(lmc, n)]
    assert not missing, missing
    assert [int(lmc.LogLevel.debug), int(lmc.LogLevel.info), int(lmc.LogLevel.warn), int(lmc.LogLevel.error),
            int(lmc.LogLevel.off)] == [1, 2, 3, 4, 6]
    lmc.set_log_level(lmc.LogLevel.warn)
    assert lmc.get_log_level() == lmc.LogLevel.warn
    lmc.set_log_level(lmc.LogLevel.info)


def test_photon_and_source_distributions():
    """Photon's constructor frame and Stokes vector (src/photon.cpp:24-90);
    uniform_distribution / gaussian_distribution against the same draws taken
    by hand from the same xoshiro stream (src/laser.cpp:11-27), and their
    moments: disk radius^2 mean sigma^2/2, Gaussian per-axis variance sigma^2."""
    V = lmc.Vec3
    p = lmc.Photon(V(1, 2, 3), V(0, 0, 1), V(1, 0, 0), V(0, 1, 0), 0.5)
    assert p.pos.as_tuple() == (1, 2, 3) and p.prev_pos.as_tuple() == (1, 2, 3)
    assert p.k == pytest.approx(4 * math.pi)
    assert np.array_equal(np.asarray(p.P_local), np.eye(3))
    p.set_polarization(lmc.CVec2(1 / math.sqrt(2), 1j / math.sqrt(2)))
    I, Q, U, Vs = p.get_stokes_parameters()
    assert (I, Q, U) == pytest.approx((1.0, 0.0, 0.0), abs=1e-15) and Vs == pytest.approx(1.0)
    p.polarized = False
    assert p.get_stokes_parameters() == [1.0, 0.0, 0.0, 0.0]
    r1, r2 = lmc.Rng(77), lmc.Rng(77)
    c = V(0.5, -0.25, 7.0)
    q = lmc.uniform_distribution(r1, c, 2.0)
    th, rr = r2.uniform() * 2 * math.pi, math.sqrt(r2.uniform()) * 2.0
    assert (q.x, q.y, q.z) == (0.5 + rr * math.cos(th), -0.25 + rr * math.sin(th), 7.0)
    q = lmc.gaussian_distribution(r1, c, 3.0)
    assert (q.x, q.y) == (r2.normal(0.5, 3.0), r2.normal(-0.25, 3.0))
    rng = lmc.Rng(5)
    n = 20000
    d = np.array([lmc.uniform_distribution(rng, V(), 1.0).as_tuple() for _ in range(n)])
    r2m = (d[:, 0] ** 2 + d[:, 1] ** 2).mean()
    assert abs(r2m - 0.5) < 5 * math.sqrt(1 / 12 / n) and (d[:, 0] ** 2 + d[:, 1] ** 2).max() <= 1.0
    g = np.array([lmc.gaussian_distribution(rng, V(), 2.0).as_tuple() for _ in range(n)])
    assert abs(g[:, 0].var() - 4.0) < 5 * 4.0 * math.sqrt(2 / n) and abs(g[:, 1].mean()) < 5 * 2.0 / math.sqrt(n)


def test_autotuner_picks_the_fastest_usable_shape(monkeypatch, tmp_path):
    """The tuner times every candidate on a scaled-down copy of the run (the
    same scenes at `fraction` of their budgets, every non-empty scene at least
    one photon), skips shapes the run refuses, zeroes the target, returns the
    fastest and caches it in memory and on disk."""
    from luminis_mc import tuning
    monkeypatch.setattr(tuning, "_device_name", lambda: "gfx950:256")
    monkeypatch.setattr(tuning, "_memo", {})
    cost = {(256, 4, 256): 9.0, (512, 2, 256): 7.5, (1024, 1, 256): 6.0, (1024, 1, 64): 5.5, (256, 4, 8): 8.0}

    class FakeBatch:
        def __init__(self):
            self.calls, self.zeroed = [], 0

        def __len__(self):
            return 3

        def run(self, budgets, r, use_graph=False):
            shape = (r.block_size, r.blocks_per_cu, r.claim)
            self.calls.append((shape, [int(b) for b in budgets]))
            if shape == (1024, 1, 64):
                raise NotImplementedError("refused shape")
            return {"kernel_ms": cost.get(shape, 10.0)}

        def zero(self):
            self.zeroed += 1

    t = FakeBatch()
    budgets = np.array([10_000_000, 0, 3], dtype=np.uint64)
    cache = str(tmp_path / "tune.json")
    res = tuning.autotune(t, _lib.RunDesc(), budgets=budgets, fraction=0.05, min_photons=1000, cache_path=cache)
    assert (res["block_size"], res["blocks_per_cu"], res["claim"]) == (1024, 1, 256)
    assert len(res["table"]) == len(tuning.DEFAULT_CANDIDATES) - 1  # the refused shape is skipped
    assert {c for c, _ in t.calls} == set(tuning.DEFAULT_CANDIDATES)
    assert all(b == [500_000, 0, 1] for _, b in t.calls) and t.zeroed == 1
    assert (512, 2, 64) in tuning.DEFAULT_CANDIDATES and (256, 4, 8) in tuning.DEFAULT_CANDIDATES
    n_calls = len(t.calls)
    assert tuning.autotune(t, _lib.RunDesc(), budgets=budgets, cache_path=cache)["ms"] == 6.0 and len(t.calls) == n_calls
    monkeypatch.setattr(tuning, "_memo", {})
    assert tuning.autotune(t, _lib.RunDesc(), budgets=budgets, cache_path=cache)["claim"] == 256  # from disk
    assert len(t.calls) == n_calls
