"""Stage C of the CBS reverse field (reverse_field_suffix, detector.cpp:1123-1161)
on the device.  The FP64 build forms the two in-plane basis vectors as the
reference does (normalised cross products); the FP32 build takes their
projections from the frames' handedness (m x n = h s: P1 = (m1, s1 x m1, s1)
is right-handed, the exit frame has the launch frame's h).  On random frames
of both handednesses the FP32 field must equal the FP64 one to FP32 rounding;
a sign slip in either identity moves it by O(1).  At exact backscatter
(s_n = -s1) the in-plane axis is the rounding residue of a vanishing cross
product in both builds (ill-defined, as in the reference): there the field is
only required to be finite and of unit norm."""
import numpy as np
import pytest

from luminis_mc import _lib

pytestmark = pytest.mark.gpu


def _rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 2] = -q[:, 2]
    return q.T  # rows: a right-handed orthonormal frame


def _cases(rng, n, hand, backscatter=False):
    out = np.zeros((n, 20))
    for i in range(n):
        f1 = _rot(rng)
        fn = _rot(rng)
        m1, s1 = f1[0], f1[2]
        mn, nn, sn = fn[0], fn[1], fn[2]
        if backscatter:
            # exit direction exactly opposite s1, exit frame about it
            sn = -s1
            mn = np.cross(f1[1], sn)
            mn /= np.linalg.norm(mn)
            nn = np.cross(sn, mn)
        if hand < 0:
            nn = -nn
        x = rng.normal(size=4)
        x /= np.linalg.norm(x)
        out[i] = np.concatenate([m1, s1, mn, nn, sn, x, [hand]])
    return out


def _run(frames, precision):
    n = len(frames)
    out = np.zeros(4 * n)
    _lib.check(_lib.lib().lmc_probe_reverse(n, precision, _lib.dptr(np.ascontiguousarray(frames.ravel())),
                                            _lib.dptr(out)), "lmc_probe_reverse")
    return out.reshape(n, 4)


@pytest.mark.parametrize("hand", [1.0, -1.0])
def test_fp32_reverse_suffix_equals_fp64(require_gpu, hand):
    rng = np.random.default_rng(11 if hand > 0 else 12)
    frames = _cases(rng, 4000, hand)
    s1 = frames[:, 3:6]
    sn = frames[:, 12:15]
    keep = np.linalg.norm(np.cross(-s1, sn), axis=1) > 1e-3  # the axis is ill-conditioned below
    assert keep.sum() > 3900
    e32 = _run(frames, _lib.PREC_FP32)[keep]
    e64 = _run(frames, _lib.PREC_FP64)[keep]
    assert np.all(np.isfinite(e32))
    assert np.allclose(np.sum(e64 ** 2, axis=1), 1.0, atol=1e-12)
    assert np.max(np.abs(e32 - e64)) < 2e-4
    assert np.mean(np.abs(e32 - e64)) < 2e-6
    back = _cases(rng, 64, hand, backscatter=True)
    for prec in (_lib.PREC_FP32, _lib.PREC_FP64):
        e = _run(back, prec)
        assert np.all(np.isfinite(e)) and np.allclose(np.sum(e ** 2, axis=1), 1.0, atol=1e-5)
