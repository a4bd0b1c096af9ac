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

