"""Batch-means statistics for Monte Carlo parity tests.

A run of N photons is split into B disjoint photon-index ranges; the per-bin
mean over batches and its standard error std(x_b, ddof=1)/sqrt(B) give the
per-bin Monte Carlo sigma.  Two independent estimators of the same quantity
are compared through z = (a - b) / sqrt(se_a^2 + se_b^2)."""
from __future__ import annotations

import numpy as np


def batch_mean_se(batches: np.ndarray):
    b = np.asarray(batches, dtype=np.float64)
    B = b.shape[0]
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(B)


def zscores(a_batches, b_batches, floor: float = 0.0):
    ma, sa = batch_mean_se(a_batches)
    mb, sb = batch_mean_se(b_batches)
    s = np.sqrt(sa ** 2 + sb ** 2)
    mask = s > floor
    z = np.zeros_like(ma)
    z[mask] = (ma[mask] - mb[mask]) / s[mask]
    return z, mask


def gpu_batches(scene, run_desc_fn, n_photons: int, B: int):
    """Run B chunks of n_photons // B photons with disjoint photon offsets."""
    per = n_photons // B
    out = []
    for b in range(B):
        scene.zero()
        scene.run(run_desc_fn(per, b * per))
        acc, _ = scene.read()
        out.append(acc)
    return np.array(out)


def oracle_batches(build, n_photons: int, B: int, seed: int, max_events: int = 1000, track: bool = True,
                   n_threads: int = 0):
    import oracle
    per = n_photons // B
    out = []
    for b in range(B):
        acc, _, _, _ = oracle.run(build, per, seed, max_events=max_events, track=track, n_threads=n_threads,
                                  photon_offset=b * per)
        out.append(acc)
    return np.array(out)
