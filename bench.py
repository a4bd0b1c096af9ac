#!/usr/bin/env python3
"""Throughput of the MI355X photon-transport path.

Workload (the multilayer, many-detector configuration): three media -- RGD
r = 0.1 um (mu_s 5, mu_a 0.02 /mm, 0.5 mm), Mie r = 1.0 um (mu_s 20, mu_a
0.005 /mm, 0.5 mm) over a semi-infinite HG g = 0.9 base (mu_s 10, mu_a 0.01
/mm) -- uniform n = 1.33, circularly polarised Gaussian beam (sigma 0.05 mm),
sensors at z = 0: far-field CBS 256 theta x 64 phi, planar fluence 256 x 256 x
(64 + 1) time bins over +-5 mm and 0-200 ps, far-field fluence 180 x 90 and
total reflectance; 1e8 photons per step, FP32 transport with FP64
accumulation.  One step = one lmc_scene_run over the step's photon budget,
with the scene (tables, accumulators) resident in HBM.

Prints one JSON line (see the repository's benchmark contract).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "luminis-mc_amd"), ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# Algorithmic work per unit (frozen model): one scattering event with CBS
# tracking = 266 FLOP (free path + move 14, theta 14, S(theta) 14, phi setup 33,
# phi trial 12 x 1.4, frame 51, Jones 41, T_raw + renorm 78, weight 4); each
# extra azimuth trial 12; light sensor deposit (fluence/field/far-field) 40;
# CBS reverse-path reconstruction 700 (counted in the records kernel).
F_EVENT = 266.0
F_TRIAL = 12.0
F_DEPOSIT = 40.0
F_CBS = 700.0
PEAK_FP32_TFLOPS = 157.3       # MI355X FP32 vector, dual-issue/packed
PEAK_FP32_NONPACKED = 78.6     # single-issue VALU FP32 (this build emits no packed FP32)
# HBM bytes per photon of the transport kernel on this workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE over one 4e6-photon launch, scaled by what the same
# counters read for a device-to-device copy of known size (`bench.py
# --copy-probe 256`): fetch x2.00, write x1.00.  Table reads are cache
# resident; the writes are the crossing records handed to the sensor kernel.
HBM_FETCH_B_PER_PHOTON = 705.4 * 1024 * 2.00 / 4e6
HBM_WRITE_B_PER_PHOTON = 6.381e5 * 1024 * 1.00 / 4e6


def metric_name():
    return "photons/s/GPU at matched CBS-bin σ: plain vs batched+adjoint; speedup vs CPU"


def flops_of(st):
    extra_trials = max(0, st["phi_trials"] - st["events"])
    transport = st["events"] * F_EVENT + extra_trials * F_TRIAL + st["crossings"] * F_DEPOSIT
    return transport, st["cbs_detections"] * F_CBS


def cpu_baseline(build, seconds: float):
    import oracle
    threads = os.cpu_count() or 1
    # two calibration passes: the first is dominated by per-run set-up (one
    # accumulator copy per thread, as in the reference's clone/merge), so the
    # second, ~1/3 of the target, sets the size of the timed sample
    n = 20000
    for frac in (1.0 / 3.0, 1.0):
        t = time.perf_counter()
        oracle.run(build, n, 12345, n_threads=threads)
        rate = n / (time.perf_counter() - t)
        n = int(min(max(rate * seconds * frac, 20000), 5e7))
    t = time.perf_counter()
    oracle.run(build, n, 777, n_threads=threads, photon_offset=10 ** 9)
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{n} photons of the same multilayer workload on the FP64 C++ restatement "
                      f"(xoshiro per-photon streams, std::thread clone/merge), {dt:.1f} s wall"}


def cpu_rate_scenes(builds, budgets, seconds: float, seed: int = 777, calib_photons: int = 20000):
    """Oracle photons/s over several scenes, each run with photons in
    proportion to its budget (one std::thread pool per scene, all host cores),
    sized to about `seconds` of CPU work by a short calibration pass."""
    import numpy as np
    import oracle
    threads = os.cpu_count() or 1
    w = np.asarray(budgets, dtype=np.float64)
    w = w / w.sum()

    def run(total):
        t = time.perf_counter()
        n_tot = 0
        for b, f in zip(builds, w):
            n = max(threads, int(round(total * f)))
            oracle.run(b, n, seed, n_threads=threads)
            n_tot += n
        return n_tot, time.perf_counter() - t

    n0, t0 = run(calib_photons * len(builds))
    target = n0 / t0 * seconds
    n, dt = run(target)
    return {"value": round(n / dt, 1), "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{n} photons over {len(builds)} scenes in proportion to their budgets, {dt:.1f} s wall"}


def copy_probe(mib: float):
    """One device-to-device copy of a known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = int(mib * 2 ** 20 // 4)
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "ok": bool(b[-1].item() == 1.0)}), flush=True)


def plain_mie_rate(n: int):
    import scenes
    from luminis_mc.simulation import GpuScene, run_desc_from
    cfg, det = scenes.config_plain_mie(n, seed=0)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    sc.run(run_desc_from(cfg, n_photons=min(n, 10 ** 6)))
    t = time.perf_counter()
    st = sc.run(run_desc_from(cfg))
    dt = time.perf_counter() - t
    sc.close()
    return n / dt, st


def batched_leg(frac_plain: float = 0.25, cpu_seconds: float = 0.0):
    """The 256-scene batched workload in one persistent launch (hipGraph
    replay, autotuned launch shape) against the same scenes run one after
    another as single-scene launches."""
    import numpy as np
    import batch_scenes
    from luminis_mc.simulation import GpuBatch, GpuScene, SimConfig, run_desc_from
    from luminis_mc.tuning import autotune
    specs = batch_scenes.load()
    built = [batch_scenes.build_scene(s) for s in specs]
    budgets = np.array([s["budget"] for s in specs], dtype=np.uint64)
    c = SimConfig()
    c.seed, c.track_reverse_paths = 4242, True
    b = GpuBatch([(sa, la, g.detectors) for sa, la, g, _, _ in built])
    tune = autotune(b, run_desc_from(c, n_photons=0), budgets=budgets, fraction=0.05)
    c.block_size, c.blocks_per_cu, c.claim = tune["block_size"], tune["blocks_per_cu"], tune["claim"]
    r = run_desc_from(c, n_photons=0)
    b.run(budgets, r, use_graph=True)  # capture + first replay
    b.zero()
    t = time.perf_counter()
    st = b.run(budgets, r, use_graph=True)
    dt_full = time.perf_counter() - t
    # same scenes at a fraction of the budgets: batched vs one launch per scene
    small = np.maximum(1, np.round(budgets * frac_plain)).astype(np.uint64)
    b.run(small, r)
    t = time.perf_counter()
    st_small = b.run(small, r)
    dt_e = time.perf_counter() - t  # eager launch sequence
    b.run(small, r, use_graph=True)  # capture + first replay
    t = time.perf_counter()
    b.run(small, r, use_graph=True)
    dt_b = time.perf_counter() - t  # the batched form: hipGraph replay
    b.close()
    c2 = SimConfig()
    c2.seed, c2.track_reverse_paths = 4242, True
    scs = []
    for k, (sa, la, g, _, _) in enumerate(built):
        sc = GpuScene(sa, la, g.detectors)
        sc.run(run_desc_from(c2, n_photons=int(small[k])))
        scs.append(sc)
    t = time.perf_counter()
    for k, sc in enumerate(scs):
        sc.run(run_desc_from(c2, n_photons=int(small[k])))
    dt_p = time.perf_counter() - t
    for sc in scs:
        sc.close()
    cpu = None
    if cpu_seconds > 0:  # every 16th scene
        idx = list(range(0, len(built), 16))
        cpu = cpu_rate_scenes([_build_of(built[k]) for k in idx], [budgets[k] for k in idx], cpu_seconds)
    res = {"workload": "256 independent scenes, 1-4 layers, HG/Mie, CBS 100x1 each (tests/data/batch256.json)",
            "photons": int(budgets.sum()), "photons_per_s": round(float(budgets.sum()) / dt_full, 1),
            "events_per_photon": round(st["events"] / max(1, st["photons"]), 2), "graph": True,
            "launch_shape": {k: tune[k] for k in ("block_size", "blocks_per_cu", "claim")},
            "vs_plain": {"budget_fraction": frac_plain, "photons": int(small.sum()),
                         "batched_photons_per_s": round(float(small.sum()) / dt_b, 1),
                         "batched_eager_photons_per_s": round(float(small.sum()) / dt_e, 1),
                         "plain_photons_per_s": round(float(small.sum()) / dt_p, 1),
                         "speedup": round(dt_p / dt_b, 2), "speedup_eager": round(dt_p / dt_e, 2)}}
    if cpu is not None:
        res["cpu_baseline"] = cpu
        res["speedup_vs_cpu"] = round(res["photons_per_s"] / cpu["value"], 2)
    return res


def _build_of(scene_tuple):
    from luminis_mc.simulation import SceneBuild
    sa, la, g, _, _ = scene_tuple
    return SceneBuild(sa, la, g.detectors)


def adjoint_leg(n: int = 20_000_000):
    """Forward-only vs forward + score gradients (5 layer parameters) on the
    two-media gradient workload."""
    import scenes
    import luminis_mc as lmc
    sample, laser, sensors, keep = scenes.scene_adjoint()
    tr = lmc.DifferentiableTransport(sample, laser, sensors, n_photons=n, seed=99)
    th = tr.theta0()
    sc = tr.scene(th)
    tr.evaluate(th, with_grad=False, n_photons=min(n, 10 ** 6), scene=sc)
    tr.evaluate(th, with_grad=True, n_photons=min(n, 10 ** 6), scene=sc)
    t = time.perf_counter()
    tr.evaluate(th, with_grad=False, scene=sc)
    dt_f = time.perf_counter() - t
    t = time.perf_counter()
    vals, jac = tr.evaluate(th, with_grad=True, scene=sc)
    dt_a = time.perf_counter() - t
    sc.close()
    return {"workload": "Mie r=0.3um [0,1) mm over HG g=0.8; reflectance + CBS 100x1; d/d(mu_s, mu_a) both "
                        "layers, d/dg HG layer", "photons": n,
            "forward_photons_per_s": round(n / dt_f, 1), "adjoint_photons_per_s": round(n / dt_a, 1),
            "reflectance": float(vals[0]), "dR_dtheta": [[float(x) for x in row] for row in jac[:, :, 0]]}


def polydisperse_leg(n: int = 200_000_000, cpu_seconds: float = 0.0):
    """The polydisperse Mie mixture (32 lognormal radius bins, 2048-node S tables)
    with far-field CBS 512 x 128 and a 1024 x 1024 planar field (FP64 phase)."""
    import scenes
    from luminis_mc.simulation import GpuScene, SimConfig, run_desc_from
    sample, laser, sensors, keep = scenes.scene_polydisperse()
    sc = GpuScene(sample, laser, sensors)
    c = SimConfig()
    c.seed, c.track_reverse_paths = 5150, True
    sc.run(run_desc_from(c, n_photons=min(n, 10 ** 7)))
    t = time.perf_counter()
    st = sc.run(run_desc_from(c, n_photons=n, photon_offset=10 ** 7))
    dt = time.perf_counter() - t
    build = sc.build
    sc.close()
    res = {"workload": "semi-infinite mixture of 32 Mie radius bins (lognormal, median 0.4 um, sigma_ln 0.3), "
                        "mu_s 30, mu_a 0.002 /mm; CBS 512x128 + planar field 1024x1024", "photons": n,
            "photons_per_s": round(n / dt, 1), "events_per_photon": round(st["events"] / max(1, st["photons"]), 2),
            "sensor_kernel_ms": round(st["sensor_ms"], 3), "transport_ms": round(st["transport_ms"], 3)}
    if cpu_seconds > 0:
        res["cpu_baseline"] = cpu_rate_scenes([build], [1], cpu_seconds)
        res["speedup_vs_cpu"] = round(res["photons_per_s"] / res["cpu_baseline"]["value"], 2)
    return res


def estimator_leg(n: int = 1_000_000, cpu_seconds: float = 0.0):
    """The CBS estimator (every vertex forced into each of 300 theta bins) on
    the reference's estimator-study medium; the vertex-bin rate is the work unit
    of the per-bin loop."""
    import scenes
    from luminis_mc.simulation import GpuScene, SimConfig, run_desc_from
    sample, laser, sensors, keep = scenes.scene_estimator()
    sc = GpuScene(sample, laser, sensors)
    c = SimConfig()
    c.seed, c.track_reverse_paths = 8080, True
    sc.run(run_desc_from(c, n_photons=min(n, 20000)))
    t = time.perf_counter()
    st = sc.run(run_desc_from(c, n_photons=n, photon_offset=10 ** 7))
    dt = time.perf_counter() - t
    build = sc.build
    sc.close()
    bins = 300
    res = {"workload": "estimator-mode FarFieldCBSSensor, 10 deg in 300 x 1 bins, semi-infinite RGD r=0.110 um, "
                       "mu_s 10, mu_a 0 /mm, Gaussian beam (the reference's estimator study)",
           "photons": n, "photons_per_s": round(n / dt, 1),
           "events_per_photon": round(st["events"] / max(1, st["photons"]), 2),
           "vertex_bins_per_s": round(st["events"] * bins / dt, 1), "kernel_ms": round(st["kernel_ms"], 3),
           "launch": {"block": st["block_size"], "blocks_per_cu": st["blocks_per_cu"],
                      "lds_tables": bool(st["lds_tables"])}}
    if cpu_seconds > 0:
        res["cpu_baseline"] = cpu_rate_scenes([build], [1], cpu_seconds, calib_photons=64)
        res["speedup_vs_cpu"] = round(res["photons_per_s"] / res["cpu_baseline"]["value"], 2)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8, help="photons per step")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-poly", action="store_true")
    ap.add_argument("--no-estimator", action="store_true")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only copy this many MiB device-to-device (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the transport path has no multi-GPU mode in this build "
              "(photon ranges would shard across GPUs with one final FP64 reduction)", file=sys.stderr)
        sys.exit(2)

    import luminis_mc as lmc
    import scenes
    from luminis_mc.simulation import GpuScene, run_desc_from

    n = int(args.photons)
    cfg, sensors = scenes.config_multilayer(n, seed=2026, small=False)
    scene = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    for w in range(args.warmup):
        scene.run(run_desc_from(cfg, n_photons=n, photon_offset=(10 ** 12) + w * n))
    scene.zero()
    tot = {"photons": 0, "events": 0, "phi_trials": 0, "crossings": 0, "cbs_detections": 0,
           "transport_ms": 0.0, "sensor_ms": 0.0, "kernel_ms": 0.0, "transport_launches": 0}
    t0 = time.perf_counter()
    st = None
    for k in range(args.steps):
        st = scene.run(run_desc_from(cfg, n_photons=n, photon_offset=k * n))
        for key in tot:
            tot[key] += st[key]
    elapsed = time.perf_counter() - t0
    value = args.steps * n / elapsed
    acc, hits = scene.read()
    f_transport, f_cbs = flops_of(tot)
    avg_launch_s = tot["transport_ms"] * 1e-3 / max(1, tot["transport_launches"])
    per_launch = f_transport / max(1, tot["transport_launches"])
    achieved = per_launch / avg_launch_s / 1e12
    photons_per_launch = tot["photons"] / max(1, tot["transport_launches"])
    traffic = (HBM_FETCH_B_PER_PHOTON + HBM_WRITE_B_PER_PHOTON) * photons_per_launch
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "frac_of_nonpacked_peak": round(achieved / PEAK_FP32_NONPACKED, 4), "traffic": round(traffic),
            "traffic_gb_s": round(traffic / avg_launch_s / 1e9, 1),
            "kernel": "k_transport_wide" if st and st["block_size"] > 512 else "k_transport",
            "launch": {"block": st["block_size"], "blocks_per_cu": st["blocks_per_cu"],
                       "lds_tables": bool(st["lds_tables"])} if st else None,
            "avg_launch_ms": round(avg_launch_s * 1e3, 3),
            "flop_per_launch": per_launch}
    out = {
        "metric": metric_name(), "value": round(value, 1), "unit": "photons/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (photon transport from a seeded Philox stream; no input data)",
        "config": {"workload": "multilayer 3-media CBS 256x64 + planar fluence 256x256x65 + far-field "
                               "fluence 180x90 + reflectance, 1e8 photons/step",
                   "photons_per_step": n, "events_per_photon": round(tot["events"] / max(1, tot["photons"]), 2),
                   "kernel_ms_per_step": round(tot["kernel_ms"] / args.steps, 3),
                   "transport_ms_per_step": round(tot["transport_ms"] / args.steps, 3),
                   "sensor_kernel_ms_per_step": round(tot["sensor_ms"] / args.steps, 3),
                   "reflectance": float(acc[scene.build.offsets[3][0]]) / (args.steps * n)},
        "roofline": roof,
    }
    if not args.no_plain:
        rate, pst = plain_mie_rate(10 ** 7)
        out["plain_mie_cbs"] = {"photons_per_s": round(rate, 1), "photons": 10 ** 7,
                                "events_per_photon": round(pst["events"] / max(1, pst["photons"]), 2),
                                "kernel_ms": round(pst["kernel_ms"], 3)}
    if not args.no_batch:
        out["batched"] = batched_leg(cpu_seconds=0.0 if args.no_cpu else 10.0)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
    if not args.no_poly:
        out["polydisperse"] = polydisperse_leg(cpu_seconds=0.0 if args.no_cpu else 8.0)
    if not args.no_estimator:
        out["estimator"] = estimator_leg(cpu_seconds=0.0 if args.no_cpu else 10.0)
    if not args.no_cpu:
        cb = cpu_baseline(scene.build, args.cpu_seconds)
        out["cpu_baseline"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in cb.items()}
        out["speedup_vs_cpu"] = round(value / cb["value"], 2)
    scene.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
