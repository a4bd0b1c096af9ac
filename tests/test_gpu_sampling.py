"""Device azimuth sampler of the production path against its exact distribution.

The reference draws phi with density F(phi) = a cos^2 phi + b sin^2 phi +
c sin phi cos phi by rejection (src/phase.cpp:27-49); the production path samples
the same density exactly from two words as a mixture of a uniform and a
cos^2 lobe about the principal axis.  Checked on device: unit (cos, sin),
F(phi) returned equal to the quadratic form, the half-period bit adding pi,
and the lattice image of the two uniforms reproducing F's CDF
(azimuth_check.py; its sensitivity to plausible slips is a CPU test).
"""
import ctypes

import numpy as np
import pytest

import azimuth_check as ac
from luminis_mc import _lib

pytestmark = pytest.mark.gpu


def _run(abc, words, precision):
    n = len(words)
    out = np.zeros(3 * n)
    P = ctypes.POINTER(ctypes.c_uint32)
    _lib.check(_lib.lib().lmc_probe_azimuth(n, precision, _lib.dptr(np.ascontiguousarray(abc.ravel())),
                                            np.ascontiguousarray(words, dtype=np.uint32).ctypes.data_as(P),
                                            _lib.dptr(out)), "lmc_probe_azimuth")
    return out.reshape(n, 3)


@pytest.mark.parametrize("precision", [_lib.PREC_FP32, _lib.PREC_FP64])
def test_azimuth_matches_its_density(require_gpu, precision):
    forms = ac.forms(np.random.default_rng(11), 16)
    words = ac.lattice_words()
    n = len(words)
    half = n // 2  # second half: same words with the half-period bit set
    worst = 0.0
    for f in forms:
        res = _run(np.tile(f, (n, 1)), words, precision)
        cph, sph, F = res[:, 0], res[:, 1], res[:, 2]
        a, b, c = f
        assert np.allclose(cph ** 2 + sph ** 2, 1.0, atol=1e-5)
        assert np.allclose(F, a * cph ** 2 + b * sph ** 2 + c * cph * sph, atol=1e-5 * (a + b + 1e-30))
        assert np.allclose(cph[half:], -cph[:half], atol=1e-6) and np.allclose(sph[half:], -sph[:half], atol=1e-6)
        d = ac.cdf_discrepancy(f, cph, sph)
        worst = max(worst, d)
    print("max |CDF - exact| over forms", worst)
    assert worst < ac.CDF_TOL
