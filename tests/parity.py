"""Trajectory-level parity of the FP64 xoshiro mode against the oracle.

In that mode the device follows the reference's per-photon streams and draw
order in FP64, so the two runs take the same branches: event, azimuth-trial
and hit counts must be equal, and every accumulator must agree to rounding.
Observed agreement is 1e-16 ... 2e-10 relative (the device's and glibc's
libm differ in the last bits of log/acos/sincos, and FP64 atomics add in
another order); FP64_RTOL is 100 times looser than the worst observed.  A
libm last-bit difference that flips a branch (a roulette draw or a plane test
landing on the other side) would change counts; none has been seen on these
scenes.
"""
import numpy as np

FP64_RTOL = 1e-8


def assert_trajectory_parity(st, gpu, hits, cpu_run, offsets, names=None, rtol=FP64_RTOL, label=""):
    """st, gpu, hits: the device run's stats, accumulators and hit counts;
    cpu_run: the oracle's (accumulators, hits, events, trials) of the same
    photons; offsets: (offset, length) per sensor."""
    cpu, chits, cev, ctr = cpu_run
    print(f"{label} events gpu/cpu {int(st['events'])} {cev} trials {int(st['phi_trials'])} {ctr} "
          f"hits {[int(h) for h in hits]} {[int(h) for h in chits]}")
    assert int(st["events"]) == int(cev), label
    assert int(st["phi_trials"]) == int(ctr), label
    assert [int(h) for h in hits] == [int(h) for h in chits], label
    names = names or [str(i) for i in range(len(offsets))]
    for (off, n), name in zip(offsets, names):
        g, c = gpu[off:off + n], cpu[off:off + n]
        scale = np.abs(c).max()
        if scale == 0:
            assert not np.any(g), (label, name)
            continue
        rel = np.abs(g - c).max() / scale
        print(f"{label} {name}: max rel diff {rel:.3e}")
        assert rel <= rtol, (label, name, rel)
