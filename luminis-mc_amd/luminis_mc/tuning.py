"""Launch-shape autotuner for the persistent transport kernel.

The transport kernel's launch has three knobs (lmc_run_desc): workgroup size,
resident workgroups per CU and the number of photon indices a wave claims per
refill.  Which shape is fastest depends on the scene (tables in LDS or not,
path-length distribution, crossings per photon) and on the device, so the
tuner times a few candidates on a scaled-down copy of the real run (same
scenes, same relative budgets, fewer photons) with the device's own event
timers, and returns the fastest.  Results are cached per (device, workload
key) in memory and, if `cache_path` is given, in a JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import _lib

# (block_size, blocks_per_cu, claim): 16 resident waves per CU is the target
# occupancy at <= 128 VGPRs; the candidates vary how they are grouped and how
# often a wave returns to the global index counter.  Claims of 64 (the
# library's single-scene default) and 8 (its estimator default) keep short
# runs' tails short; 256 keeps a batch's waves inside one scene.
DEFAULT_CANDIDATES = (
    (256, 4, 256), (512, 2, 256), (1024, 1, 256), (256, 4, 128), (256, 4, 512), (128, 8, 256), (256, 3, 256),
    (512, 2, 64), (1024, 1, 64), (256, 4, 64), (256, 4, 8),
)

_memo: dict = {}


def _device_name() -> str:
    info = _lib.DeviceInfo()
    _lib.check(_lib.lib().lmc_get_device_info(info), "lmc_get_device_info")
    return f"{info.name.decode()}:{info.n_cu}"


def _load(cache_path):
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            return json.load(f)
    return {}


def autotune(target, run: _lib.RunDesc, budgets=None, fraction: float = 0.05, min_photons: int = 200_000,
             candidates=DEFAULT_CANDIDATES, key: str | None = None, cache_path: str | None = None,
             repeats: int = 1) -> dict:
    """Time each launch shape on a scaled-down run of `target` (a GpuScene, or
    a GpuBatch with `budgets`) and return {"block_size", "blocks_per_cu",
    "claim", "ms", "table"}.  The target's accumulators are zeroed afterwards.

    fraction: share of the real photon budget(s) used per trial (at least
    min_photons photons in total)."""
    batch = budgets is not None
    dev = _device_name()
    if key is None:
        total = int(np.sum(budgets)) if batch else int(run.n_photons)
        key = f"{type(target).__name__}:{len(target) if batch else 1}:{total}:{run.precision}:{run.track_reverse_paths}"
    memo_key = f"{dev}|{key}"
    if memo_key in _memo:
        return _memo[memo_key]
    disk = _load(cache_path)
    if memo_key in disk:
        _memo[memo_key] = disk[memo_key]
        return disk[memo_key]

    if batch:
        b = np.asarray(budgets, dtype=np.float64)
        scale = max(fraction, min_photons / max(1.0, b.sum()))
        trial_budgets = np.maximum(np.round(b * min(1.0, scale)), np.where(b > 0, 1, 0)).astype(np.uint64)
    else:
        trial_n = int(max(min(run.n_photons, min_photons), run.n_photons * fraction))
    r = _lib.RunDesc()
    ctypes_fields = [f for f, _ in _lib.RunDesc._fields_]
    for f in ctypes_fields:
        setattr(r, f, getattr(run, f))
    table = []
    for (block, bpc, claim) in candidates:
        r.block_size, r.blocks_per_cu, r.claim = int(block), int(bpc), int(claim)
        best = None
        try:
            for _ in range(max(1, repeats)):
                if batch:
                    st = target.run(trial_budgets, r, use_graph=False)
                else:
                    r.n_photons = trial_n
                    st = target.run(r)
                ms = st["kernel_ms"]
                best = ms if best is None else min(best, ms)
        except NotImplementedError:  # a shape this run cannot use (e.g. 1024 threads for FP64)
            continue
        table.append({"block_size": block, "blocks_per_cu": bpc, "claim": claim, "ms": round(best, 3)})
    target.zero()
    if not table:
        raise ValueError("no candidate launch shape is usable for this run")
    win = min(table, key=lambda t: t["ms"])
    res = dict(win)
    res["table"] = table
    _memo[memo_key] = res
    if cache_path:
        disk = _load(cache_path)
        disk[memo_key] = res
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(disk, f, indent=1)
    return res
