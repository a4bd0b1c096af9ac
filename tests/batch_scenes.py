"""The batched workload: 256 independent scenes drawn by a fixed-seed generator.

Each scene: 1-4 layers (the last semi-infinite), mu_s log-uniform in [1, 50]
/mm, mu_a log-uniform in [1e-3, 1e-1] /mm, finite thicknesses uniform in
[0.1, 2] mm; each layer's phase function is Henyey-Greenstein (g uniform in
[0.5, 0.95], on an RGD amplitude host r = 0.1 um) or Mie (r log-uniform in
[0.05, 1.5] um, n_p uniform in [1.45, 1.6]); n = 1.33, lambda = 0.633 um, a
linearly polarised point beam, one far-field CBS sensor 100 theta x 1 phi over
0..0.1 rad; photon budgets log-uniform in [1e5, 1e7].

The realised list is committed as tests/data/batch256.json; `python
tests/batch_scenes.py --write` regenerates it (numpy.random.default_rng with
the seed below, draws in the order of `generate`).
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

SEED = 20261203
N_SCENES = 256
HERE = os.path.dirname(os.path.abspath(__file__))
JSON_PATH = os.path.join(HERE, "data", "batch256.json")


def generate(n: int = N_SCENES, seed: int = SEED) -> list[dict]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        n_layers = int(rng.integers(1, 5))
        layers = []
        z = 0.0
        for l in range(n_layers):
            mu_s = float(math.exp(rng.uniform(math.log(1.0), math.log(50.0))))
            mu_a = float(math.exp(rng.uniform(math.log(1e-3), math.log(1e-1))))
            thick = float(rng.uniform(0.1, 2.0))
            if rng.uniform() < 0.5:
                phase = {"kind": "hg", "g": float(rng.uniform(0.5, 0.95))}
            else:
                phase = {"kind": "mie", "r_um": float(math.exp(rng.uniform(math.log(0.05), math.log(1.5)))),
                         "n_p": float(rng.uniform(1.45, 1.6))}
            z_max = math.inf if l == n_layers - 1 else z + thick
            layers.append({"z_min": z, "z_max": z_max, "mu_s": mu_s, "mu_a": mu_a, "phase": phase})
            z = z_max
        budget = int(round(math.exp(rng.uniform(math.log(1e5), math.log(1e7)))))
        out.append({"id": i, "layers": layers, "budget": budget})
    return out


def _json_safe(specs):
    def fix(v):
        return "inf" if isinstance(v, float) and math.isinf(v) else v
    return [{**s, "layers": [{k: fix(v) for k, v in L.items()} for L in s["layers"]]} for s in specs]


def load() -> list[dict]:
    with open(JSON_PATH) as f:
        specs = json.load(f)
    for s in specs:
        for L in s["layers"]:
            if L["z_max"] == "inf":
                L["z_max"] = math.inf
    return specs


def write() -> None:
    os.makedirs(os.path.dirname(JSON_PATH), exist_ok=True)
    with open(JSON_PATH, "w") as f:
        f.write("[\n" + ",\n".join(json.dumps(s) for s in _json_safe(generate())) + "\n]\n")


def build_scene(spec: dict, n_div: int = 10000, extra_sensors: bool = False):
    """(sample, laser, sensors group, cbs sensor, keep-alive objects) of one
    scene.  extra_sensors: also a total reflectance and an 18 x 9 far-field
    fluence over the backward hemisphere (non-absorbing, beside the CBS
    sensor), so a parity check sees most exiting photons, not only the few in
    the CBS cone."""
    import luminis_mc as lmc
    from scenes import WAVELENGTH, cbs_sensor
    n = 1.33
    sample = lmc.Sample(n)
    keep = []
    for L in spec["layers"]:
        ph = L["phase"]
        if ph["kind"] == "hg":
            pf = lmc.HenyeyGreensteinPhaseFunction(ph["g"])
            m = lmc.RGDMedium(pf, 0.1e-3, 1.59, n, WAVELENGTH)
        else:
            r = ph["r_um"] * 1e-3
            pf = lmc.MiePhaseFunction(WAVELENGTH, r, ph["n_p"], n, n_div, 0.0, math.pi)
            m = lmc.MieMedium(pf, r, ph["n_p"], n, WAVELENGTH)
        m.set_scattering_coefficient(L["mu_s"])
        m.set_absorption_coefficient(L["mu_a"])
        m.set_mean_free_path(1.0 / (L["mu_s"] + L["mu_a"]))
        sample.add_layer(m, L["z_min"], L["z_max"])
        keep += [pf, m]
    laser = lmc.Laser(1.0, 0.0, WAVELENGTH, 0.0, lmc.LaserSource.Point)
    group = lmc.SensorsGroup()
    det = group.add_detector(cbs_sensor(100, 0.1, 1))
    if extra_sensors:
        group.add_detector(lmc.ReflectanceSensor(0.0, absorb=False))
        group.add_detector(lmc.FarFieldFluenceSensor(math.pi / 2, 2 * math.pi, (math.pi / 2) / 18, 2 * math.pi / 9,
                                                     absorb=False))
    return sample, laser, group, det, keep


if __name__ == "__main__":
    if "--write" in sys.argv:
        write()
        print(JSON_PATH)
    else:
        specs = generate()
        print(len(specs), "scenes,", sum(s["budget"] for s in specs), "photons")
