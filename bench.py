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

Beside the timed steps: σ of the CBS coherent S0 (theta-bin 0 and the whole
cone, phi summed) from B = 32 batch means on the GPU and on the CPU
baseline, and the other configurations' legs (plain Mie CBS, the 256-scene
batch against one launch per scene, the adjoint, the polydisperse mixture,
the CBS estimator), each with its CPU baseline where one is timed.  The CPU
baseline is the oracle's release build (liblmc_oracle_fast.so) on all host
cores: one warm-up, three repetitions, the median.

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
# resident; the writes are the light sensors' FP64 deposits (planar fluence,
# far-field fluence: 6.4 per photon, one 32-byte sector each) and the few CBS
# crossing records handed to the sensor kernel.
HBM_FETCH_B_PER_PHOTON = 1663.1 * 1024 * 2.00 / 4e6
HBM_WRITE_B_PER_PHOTON = 8.0712e5 * 1024 * 1.00 / 4e6


def log(msg: str) -> None:
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def metric_name():
    return "photons/s/GPU at matched CBS-bin σ: plain vs batched+adjoint; speedup vs CPU"


def flops_of(st):
    extra_trials = max(0, st["phi_trials"] - st["events"])
    transport = st["events"] * F_EVENT + extra_trials * F_TRIAL + st["crossings"] * F_DEPOSIT
    return transport, st["cbs_detections"] * F_CBS


# ------------------------------------------------ σ from batch means ----
SIGMA_BATCHES = 32


def sigma_record(values, photons: int) -> dict:
    """Relative σ of a bin's total over `photons` photons from B equal batch
    sums (the batch-means estimator: σ(total) = sqrt(B) * sd(batch sums)), and
    σ_rel * sqrt(photons), the per-photon figure two runs of different sizes
    compare at (equal figures = equal σ at equal photon counts)."""
    import numpy as np
    v = np.asarray(values, dtype=np.float64)
    tot = float(v.sum())
    rec = {"B": int(v.size), "photons": int(photons), "total": tot, "nonzero_batches": int(np.count_nonzero(v))}
    if tot != 0.0 and v.size > 1:
        rel = float(v.std(ddof=1) * math.sqrt(v.size) / abs(tot))
        rec["rel_sigma"] = round(rel, 6)
        rec["rel_sigma_sqrt_photons"] = round(rel * math.sqrt(photons), 3)
    return rec


def cbs_s0_probes(acc, off: int, n_theta: int, n_phi: int):
    """(theta-bin 0, whole cone) of a CBS sensor's coherent S0 grid (grid 0 of
    its eight [S0c..S3c, S0i..S3i] [theta][phi] grids), summed over phi."""
    g = acc[off:off + n_theta * n_phi]
    return float(g[:n_phi].sum()), float(g.sum())


def sigma_pair(b0, cone, photons):
    return {"theta_bin0": sigma_record(b0, photons), "cone": sigma_record(cone, photons)}


def gpu_sigma(scene, rd_of, photons: int, probe, offset: int = 10 ** 13, stride: int = 0):
    """B launches of photons/B on a GpuScene (batch b from photon index
    offset + b * stride, stride photons/B by default), read and zeroed after
    each: the per-batch CBS S0 probes."""
    per = photons // SIGMA_BATCHES
    stride = stride or per
    b0, cone = [], []
    scene.zero()
    for b in range(SIGMA_BATCHES):
        scene.run(rd_of(per, offset + b * stride))
        x0, xc = probe(scene.read()[0])
        b0.append(x0)
        cone.append(xc)
        scene.zero()
    return sigma_pair(b0, cone, per * SIGMA_BATCHES)


# --------------------------------------------------------- CPU baseline ----
def cpu_threads() -> int:
    return os.cpu_count() or 1


def cpu_rate(builds, weights, seconds: float, reps: int = 3, seed: int = 777, calib_photons: int = 20000) -> dict:
    """Photons/s of the oracle's release build (liblmc_oracle_fast.so:
    -O3 -flto, FMA contraction, x86-64-v3) on all host cores over scenes run
    with photons in proportion to `weights`: one warm-up/calibration run,
    then `reps` timed repetitions of about seconds/reps each; the median.  A
    run's fixed cost (the per-thread accumulator clones and their merge, as in
    the reference's run_simulation_parallel) is timed on a run of one photon
    per thread and taken off each repetition, which can only raise the rate."""
    import numpy as np
    import oracle
    threads = cpu_threads()
    w = np.asarray(weights, dtype=np.float64)
    w = w / w.sum()

    def run(total, off):
        t = time.perf_counter()
        n_tot = 0
        for b, f in zip(builds, w):
            n = max(threads, int(round(total * f)))
            oracle.run(b, n, seed, n_threads=threads, photon_offset=off, variant="fast")
            n_tot += n
        return n_tot, time.perf_counter() - t

    n0, t0 = run(calib_photons * len(builds), 10 ** 9)  # warm-up and calibration
    n1, t1 = run(threads * len(builds), 2 * 10 ** 9)  # fixed cost: one photon per thread
    target = n0 / max(1e-9, t0 - t1) * seconds / reps
    rates, walls = [], []
    for r in range(reps):
        n, dt = run(target, (3 + r) * 10 ** 9)
        walls.append(n / dt)
        rates.append(n / max(1e-9, dt - t1))
        last_n = n
    return {"value": round(float(np.median(rates)), 1), "unit": "photons/s", "cores": threads, "kind": "port",
            "build": "oracle/liblmc_oracle_fast.so (-O3 -flto -march=x86-64-v3, FMA contraction, std::thread)",
            "reps": reps, "rates": [round(x, 1) for x in rates], "median_wall_rate": round(float(np.median(walls)), 1),
            "fixed_cost_s": round(t1, 3),
            "sample": f"{last_n} photons per repetition over {len(builds)} scene(s), 1 warm-up + {reps} repetitions, "
                      f"median; the FP64 C++ restatement (xoshiro per-photon streams, std::thread clone/merge)"}


def cpu_sigma(build, photons: int, probe, seed: int = 991) -> dict:
    """The oracle's release build over B batches of photons/B (all host cores)."""
    import oracle
    per = max(cpu_threads(), photons // SIGMA_BATCHES)
    b0, cone = [], []
    for b in range(SIGMA_BATCHES):
        acc = oracle.run(build, per, seed, n_threads=cpu_threads(), photon_offset=5 * 10 ** 9 + b * per,
                         variant="fast")[0]
        x0, xc = probe(acc)
        b0.append(x0)
        cone.append(xc)
    return sigma_pair(b0, cone, per * SIGMA_BATCHES)


def copy_probe(mib: float):
    """One device-to-device copy of a known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = int(mib * 2 ** 20 // 4)
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "ok": bool(b[-1].item() == 1.0)}), flush=True)


def plain_mie_leg(n: int = 10 ** 7, cpu_seconds: float = 0.0) -> dict:
    """The plain configuration (one semi-infinite Mie layer, CBS 100 x 1):
    photons/s, and σ of the CBS S0 at theta-bin 0 and over the cone from B
    batch means, on the GPU and on the CPU baseline."""
    import scenes
    from luminis_mc.simulation import GpuScene, SceneBuild, run_desc_from
    cfg, det = scenes.config_plain_mie(n, seed=0)
    sc = GpuScene(cfg.sample, cfg.laser, cfg.detector.detectors)
    sc.run(run_desc_from(cfg, n_photons=min(n, 10 ** 6), photon_offset=10 ** 12))
    sc.zero()
    t = time.perf_counter()
    st = sc.run(run_desc_from(cfg))
    dt = time.perf_counter() - t
    off = sc.build.offsets[0][0]
    probe = lambda acc: cbs_s0_probes(acc, off, 100, 1)  # noqa: E731
    res = {"workload": "semi-infinite Mie r=0.5um (mu_s 10, mu_a 0.01 /mm), point beam, CBS 100x1 over 0.1 rad",
           "photons": n, "photons_per_s": round(n / dt, 1),
           "events_per_photon": round(st["events"] / max(1, st["photons"]), 2), "kernel_ms": round(st["kernel_ms"], 3),
           "sigma_cbs_s0": gpu_sigma(sc, lambda k, o: run_desc_from(cfg, n_photons=k, photon_offset=o), n, probe)}
    sc.close()
    if cpu_seconds > 0:
        b = SceneBuild(cfg.sample, cfg.laser, cfg.detector.detectors)
        cpu = cpu_rate([b], [1], cpu_seconds)
        cpu["sigma_cbs_s0"] = cpu_sigma(b, int(cpu["value"] * cpu_seconds / 3), probe)
        res["cpu_baseline"] = cpu
        res["speedup_vs_cpu"] = round(res["photons_per_s"] / cpu["value"], 2)
    return res


def batched_leg(frac_plain: float = 0.25, cpu_seconds: float = 0.0, sigma_every: int = 16):
    """The 256-scene batched workload in one persistent launch (hipGraph
    replay, autotuned launch shape) against the same scenes run one after
    another as single-scene launches; σ of every 16th scene's CBS S0 from B
    batch means in the batched form, the plain form and on the CPU."""
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
    b.run(small, r)
    dt_e = time.perf_counter() - t  # eager launch sequence
    b.run(small, r, use_graph=True)  # capture + first replay
    t = time.perf_counter()
    b.run(small, r, use_graph=True)
    dt_b = time.perf_counter() - t  # the batched form: hipGraph replay
    # σ: B batched runs of budgets / B; the sampled scenes' CBS S0 probes
    idx = list(range(0, len(built), sigma_every))
    per = np.maximum(1, budgets // SIGMA_BATCHES).astype(np.uint64)
    probe = lambda acc: cbs_s0_probes(acc, 0, 100, 1)  # noqa: E731
    pb = {k: ([], []) for k in idx}
    b.zero()
    for j in range(SIGMA_BATCHES):
        # batched runs index a scene's photons below 2^32
        rj = run_desc_from(c, n_photons=0, photon_offset=10 ** 9 + j * int(budgets.max()))
        b.run(per, rj)
        for k in idx:
            x0, xc = probe(b.read(k)[0])
            pb[k][0].append(x0)
            pb[k][1].append(xc)
        b.zero()
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
    pp = {}
    for k in idx:  # the plain form on the same photon streams as the batched σ runs
        pp[k] = gpu_sigma(scs[k], lambda n_, o_: run_desc_from(c2, n_photons=n_, photon_offset=o_),
                          int(per[k]) * SIGMA_BATCHES, probe, offset=10 ** 9, stride=int(budgets.max()))
    for sc in scs:
        sc.close()

    def med(recs, key):
        vals = [x[key]["rel_sigma_sqrt_photons"] for x in recs if "rel_sigma_sqrt_photons" in x[key]]
        return round(float(np.median(vals)), 3) if vals else None

    sb = {k: sigma_pair(pb[k][0], pb[k][1], int(per[k]) * SIGMA_BATCHES) for k in idx}
    res = {"workload": "256 independent scenes, 1-4 layers, HG/Mie, CBS 100x1 each (tests/data/batch256.json)",
           "photons": int(budgets.sum()), "photons_per_s": round(float(budgets.sum()) / dt_full, 1),
           "events_per_photon": round(st["events"] / max(1, st["photons"]), 2), "graph": True,
           "launch_shape": {k: tune[k] for k in ("block_size", "blocks_per_cu", "claim")},
           "vs_plain": {"budget_fraction": frac_plain, "photons": int(small.sum()),
                        "batched_photons_per_s": round(float(small.sum()) / dt_b, 1),
                        "batched_eager_photons_per_s": round(float(small.sum()) / dt_e, 1),
                        "plain_photons_per_s": round(float(small.sum()) / dt_p, 1),
                        "speedup": round(dt_p / dt_b, 2), "speedup_eager": round(dt_p / dt_e, 2)},
           "sigma_cbs_s0": {"scenes": idx, "B": SIGMA_BATCHES,
                            "median_rel_sigma_sqrt_photons": {
                                "batched": {"theta_bin0": med(sb.values(), "theta_bin0"),
                                            "cone": med(sb.values(), "cone")},
                                "plain": {"theta_bin0": med(pp.values(), "theta_bin0"),
                                          "cone": med(pp.values(), "cone")}},
                            "batched": {str(k): sb[k] for k in idx[:4]}, "plain": {str(k): pp[k] for k in idx[:4]}}}
    if cpu_seconds > 0:
        builds = [_build_of(built[k]) for k in idx]
        cpu = cpu_rate(builds, [budgets[k] for k in idx], cpu_seconds)
        tot = cpu["value"] * cpu_seconds / 3
        w = np.array([float(budgets[k]) for k in idx])
        cs = {k: cpu_sigma(bb, max(SIGMA_BATCHES * 64, int(tot * wk / w.sum())), probe)
              for k, bb, wk in zip(idx, builds, w)}
        cpu["sigma_cbs_s0_median_rel_sigma_sqrt_photons"] = {"theta_bin0": med(cs.values(), "theta_bin0"),
                                                             "cone": med(cs.values(), "cone")}
        res["cpu_baseline"] = cpu
        res["speedup_vs_cpu"] = round(res["photons_per_s"] / cpu["value"], 2)
    return res


def _build_of(scene_tuple):
    from luminis_mc.simulation import SceneBuild
    sa, la, g, _, _ = scene_tuple
    return SceneBuild(sa, la, g.detectors)


def adjoint_leg(n: int = 200_000_000):
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
                        "layers, d/dg HG layer", "photons": n, "configured_photons": 200_000_000,
            "forward_photons_per_s": round(n / dt_f, 1), "adjoint_photons_per_s": round(n / dt_a, 1),
            "reflectance": float(vals[0]), "dR_dtheta": [[float(x) for x in row] for row in jac[:, :, 0]]}


def polydisperse_leg(n: int = 200_000_000, cpu_seconds: float = 0.0):
    """The polydisperse Mie mixture (32 lognormal radius bins, 2048-node S tables)
    with far-field CBS 512 x 128 and a 1024 x 1024 planar field (FP64 phase).
    The configuration's budget is 2e9 photons; the rate is measured on 2e8
    (a tenth: per-photon work does not depend on the count, and the run
    stays within seconds)."""
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
            "configured_photons": 2_000_000_000,
            "photons_per_s": round(n / dt, 1), "events_per_photon": round(st["events"] / max(1, st["photons"]), 2),
            "sensor_kernel_ms": round(st["sensor_ms"], 3), "transport_ms": round(st["transport_ms"], 3)}
    if cpu_seconds > 0:
        res["cpu_baseline"] = cpu_rate([build], [1], cpu_seconds)
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
        res["cpu_baseline"] = cpu_rate([build], [1], cpu_seconds, calib_photons=64)
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
    ap.add_argument("--no-sigma", action="store_true", help="skip the σ passes (profiling runs)")
    ap.add_argument("--sigma-photons", type=float, default=2e8, help="photons of the GPU σ pass (B batches)")
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
    cbs_off = scene.build.offsets[0][0]
    probe = lambda a: cbs_s0_probes(a, cbs_off, 256, 64)  # noqa: E731
    log(f"timed steps done: {value:.4g} photons/s")
    if not args.no_sigma:
        out["sigma_cbs_s0"] = gpu_sigma(scene, lambda k, o: run_desc_from(cfg, n_photons=k, photon_offset=o),
                                        int(args.sigma_photons), probe)
    if not args.no_plain:
        log("plain Mie leg")
        out["plain_mie_cbs"] = plain_mie_leg(cpu_seconds=0.0 if args.no_cpu else 6.0)
    if not args.no_batch:
        log("batched leg")
        out["batched"] = batched_leg(cpu_seconds=0.0 if args.no_cpu else 10.0)
    if not args.no_adjoint:
        log("adjoint leg")
        out["adjoint"] = adjoint_leg()
    if not args.no_poly:
        log("polydisperse leg")
        out["polydisperse"] = polydisperse_leg(cpu_seconds=0.0 if args.no_cpu else 8.0)
    if not args.no_estimator:
        log("estimator leg")
        out["estimator"] = estimator_leg(cpu_seconds=0.0 if args.no_cpu else 10.0)
    if not args.no_cpu:
        from luminis_mc.simulation import SceneBuild
        log("CPU baseline of the multilayer workload")
        cb = cpu_rate([scene.build], [1], args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = round(value / cb["value"], 2)
        if not args.no_sigma:
            # σ of the CBS bins on the CPU from a twin scene with the CBS
            # sensor alone: the other sensors absorb nothing, so trajectories
            # and CBS deposits are the same, without 17M accumulators per run
            twin = SceneBuild(cfg.sample, cfg.laser, [cfg.detector.detectors[0]])
            cs = cpu_sigma(twin, int(cb["value"] * args.cpu_seconds / 3),
                           lambda a: cbs_s0_probes(a, 0, 256, 64))
            cb["sigma_cbs_s0"] = cs
            g, c = out["sigma_cbs_s0"]["cone"], cs["cone"]
            if "rel_sigma_sqrt_photons" in g and "rel_sigma_sqrt_photons" in c:
                r = c["rel_sigma_sqrt_photons"] / g["rel_sigma_sqrt_photons"]
                out["matched_sigma"] = {
                    "quantity": "CBS coherent S0 over the cone (theta-bin 0 holds too few CPU-sample hits)",
                    "gpu_rel_sigma_sqrt_photons": g["rel_sigma_sqrt_photons"],
                    "cpu_rel_sigma_sqrt_photons": c["rel_sigma_sqrt_photons"],
                    "speedup_vs_cpu_at_matched_sigma": round(value / cb["value"] * r * r, 2)}
    scene.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
