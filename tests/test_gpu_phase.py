"""FP32 path phases on the device: the CBS reverse-path factor
exp(i k0 (s_out + s_in).(r_n - r_1)) (detector.cpp:1231-1310) reaches 1e4-1e5
rad for wide exit angles and long paths, far beyond the +-256 revolutions the
hardware sine takes directly.  The device's sin and cos of FP32 phases up to
2e5 rad must equal the FP64 sine of the same FP32 value to within the
rounding of the revolution count (2 |x| 2^-24) plus the hardware sine's own
error, and stay on the unit circle."""
import ctypes

import numpy as np
import pytest

from luminis_mc import _lib

pytestmark = pytest.mark.gpu


def _phase(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(2 * x.size, dtype=np.float32)
    P = ctypes.POINTER(ctypes.c_float)
    _lib.check(_lib.lib().lmc_probe_phase(x.size, x.ctypes.data_as(P), out.ctypes.data_as(P)), "lmc_probe_phase")
    return out[0::2].astype(np.float64), out[1::2].astype(np.float64)


def test_fp32_phase_of_any_size(require_gpu):
    rng = np.random.default_rng(7)
    x = np.concatenate([np.linspace(-2e5, 2e5, 400001), rng.uniform(-2e5, 2e5, 100000),
                        np.linspace(-10.0, 10.0, 2001)]).astype(np.float32)
    s, c = _phase(x)
    xd = x.astype(np.float64)
    tol = 2.5e-7 * np.abs(xd) + 2e-5
    assert np.all(np.abs(s - np.sin(xd)) <= tol)
    assert np.all(np.abs(c - np.cos(xd)) <= tol)
    assert np.all(np.abs(s * s + c * c - 1.0) < 1e-5)
    # beyond 256 revolutions the phase is still a phase (no zeros, no clamping)
    big = np.abs(xd) > 2000.0
    assert np.mean(np.abs(s[big])) > 0.6 and np.mean(np.abs(c[big])) > 0.6
