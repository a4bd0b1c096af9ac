"""CPU checks of the batched workload's scene list and its host-side build."""
import math

import numpy as np

import batch_scenes


def test_committed_scene_list_matches_generator():
    fresh = batch_scenes.generate()
    committed = batch_scenes.load()
    assert len(fresh) == len(committed) == batch_scenes.N_SCENES
    for a, b in zip(fresh, committed):
        assert a["budget"] == b["budget"]
        assert len(a["layers"]) == len(b["layers"])
        for la, lb in zip(a["layers"], b["layers"]):
            for k in ("z_min", "z_max", "mu_s", "mu_a"):
                assert la[k] == lb[k]
            assert la["phase"] == lb["phase"]


def test_scene_list_ranges():
    specs = batch_scenes.load()
    budgets = np.array([s["budget"] for s in specs])
    assert budgets.min() >= 1e5 and budgets.max() <= 1e7
    for s in specs:
        L = s["layers"]
        assert 1 <= len(L) <= 4
        assert L[0]["z_min"] == 0.0 and math.isinf(L[-1]["z_max"])
        for a, b in zip(L[:-1], L[1:]):
            assert a["z_max"] == b["z_min"] and 0.1 <= a["z_max"] - a["z_min"] <= 2.0
        for l in L:
            assert 1.0 <= l["mu_s"] <= 50.0 and 1e-3 <= l["mu_a"] <= 1e-1
            ph = l["phase"]
            if ph["kind"] == "hg":
                assert 0.5 <= ph["g"] <= 0.95
            else:
                assert 0.05 <= ph["r_um"] <= 1.5 and 1.45 <= ph["n_p"] <= 1.6


def test_scene_build_flattens():
    from luminis_mc.simulation import SceneBuild
    for spec in batch_scenes.load()[:8]:
        sa, la, g, det, keep = batch_scenes.build_scene(spec, n_div=1000)
        b = SceneBuild(sa, la, g.detectors)
        assert b.desc.n_layers == len(spec["layers"])
        assert b.desc.n_sensors == 1 and b.n_out == 8 * 100
