"""Checks on PhotonRecordSensor rows (layout of LMC_RECORD_FIELDS).

cbs_from_records rebuilds a FarFieldCBSSensor grid (steady state) from the
records of the same plane: each record carries everything the CBS sensor's
process_hit reads (detector.cpp:1231-1310) -- exit frame, forward Jones
vector, weight, r_1, r_n and, for events >= 2, the reverse-path Jones vector
the sensor computed.  The CBS accumulators and the records come from two
different code paths (the sensor's deposit and the record copy), so the
rebuilt grid matching the sensor's grid checks that the records carry the
state the sensor used, field by field."""
from __future__ import annotations

import numpy as np

SI, EV, DEPTH, T0, ARR, OPL, W, K = range(8)
R1, RN, HIT, DIR, M, N = slice(8, 11), slice(11, 14), slice(14, 17), slice(17, 20), slice(20, 23), slice(23, 26)


def jones(rows, first):
    return rows[:, first] + 1j * rows[:, first + 1], rows[:, first + 2] + 1j * rows[:, first + 3]


def cbs_from_records(rows, n_theta, n_phi, dtheta, dphi, s_in=(0.0, 0.0, 1.0)):
    """(8, n_theta, n_phi) Stokes grid [S0c..S3c, S0i..S3i] from records."""
    g = np.zeros((8, n_theta, n_phi))
    s = rows[:, DIR]
    theta = np.arccos(np.clip(-s[:, 2], -1, 1))
    phi = np.arctan2(s[:, 1], s[:, 0])
    phi = np.where(phi < 0, phi + 2 * np.pi, phi)
    it = (theta / dtheta).astype(np.int64)
    jp = (phi / dphi).astype(np.int64)
    ok = (it >= 0) & (it < n_theta) & (jp >= 0) & (jp < n_phi)
    ws = np.sqrt(rows[:, W])
    fm, fn = jones(rows, 26)
    rm, rn = jones(rows, 30)
    m, n = rows[:, M], rows[:, N]
    q = (s + np.asarray(s_in)[None, :]) * rows[:, K][:, None]
    pp = np.exp(1j * np.sum(q * (rows[:, RN] - rows[:, R1]), axis=1))
    multi = rows[:, EV] >= 2
    rm = np.where(multi, rm * pp, 0) * ws
    rn = np.where(multi, rn * pp, 0) * ws
    fm, fn = fm * ws, fn * ws
    efx, efy = fm * m[:, 0] + fn * n[:, 0], fm * m[:, 1] + fn * n[:, 1]
    erx, ery = rm * m[:, 0] + rn * n[:, 0], rm * m[:, 1] + rn * n[:, 1]
    ex, ey = efx + erx, efy + ery
    st = [np.abs(ex) ** 2 + np.abs(ey) ** 2, np.abs(ex) ** 2 - np.abs(ey) ** 2,
          2 * np.real(ex * np.conj(ey)), 2 * np.imag(ex * np.conj(ey)),
          np.abs(efx) ** 2 + np.abs(efy) ** 2 + np.abs(erx) ** 2 + np.abs(ery) ** 2,
          np.abs(efx) ** 2 - np.abs(efy) ** 2 + np.abs(erx) ** 2 - np.abs(ery) ** 2,
          2 * (np.real(efx * np.conj(efy)) + np.real(erx * np.conj(ery))),
          2 * (np.imag(efx * np.conj(efy)) + np.imag(erx * np.conj(ery)))]
    for k in range(8):
        np.add.at(g[k], (it[ok], jp[ok]), st[k][ok])
    return g


def frame_errors(rows):
    """Largest deviations of the record frames from orthonormal and of the
    forward Jones vector from unit norm."""
    m, n, s = rows[:, M], rows[:, N], rows[:, DIR]
    fm, fn = jones(rows, 26)
    dots = [np.abs(np.sum(a * b, axis=1)) for a, b in ((m, n), (m, s), (n, s))]
    norms = [np.abs(np.sum(a * a, axis=1) - 1) for a in (m, n, s)]
    return max(float(np.max(x)) for x in dots + norms), float(np.max(np.abs(np.abs(fm) ** 2 + np.abs(fn) ** 2 - 1)))
