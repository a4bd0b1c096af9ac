"""Post-processing of sensor accumulators (src/results.cpp, include/luminis/core/results.hpp).

These run unchanged on the GPU-filled buffers: normalise per photon and per
solid angle (dOmega = dphi (cos th_i - cos th_{i+1}), src/utils.cpp:76-82) or
per pixel area.
"""
from __future__ import annotations

import numpy as np

from .linalg import CMatrix, Matrix


def solid_angle_bin(i_theta: int, dtheta: float, dphi: float) -> float:
    return dphi * (np.cos(i_theta * dtheta) - np.cos((i_theta + 1) * dtheta))


class StokesMatrixProcessed:
    def __init__(self, S0, S1, S2, S3):
        self.S0, self.S1, self.S2, self.S3 = S0, S1, S2, S3


class StokesRadialProcessed:
    def __init__(self, S0, S1, S2, S3):
        self.S0, self.S1, self.S2, self.S3 = S0, S1, S2, S3


class FarFieldCBSProcessed:
    def __init__(self, dOmega, coherent, incoherent):
        self.dOmega, self.coherent, self.incoherent = dOmega, coherent, incoherent


class FarFieldCBSRadialProcessed:
    def __init__(self, theta_center, coherent, incoherent):
        self.theta_center, self.coherent, self.incoherent = theta_center, coherent, incoherent


class PlanarFluenceProcessed:
    def __init__(self, S0, S1, S2, S3):
        self.S0, self.S1, self.S2, self.S3 = S0, S1, S2, S3


class PlanarFieldProcessed:
    def __init__(self, Ex, Ey):
        self.Ex, self.Ey = Ex, Ey


def postprocess_farfield_cbs(det, n_photons: int, normalize_per_solid_angle: bool = True,
                             normalize_per_photon: bool = True, eps: float = 1e-30) -> FarFieldCBSProcessed:
    """src/results.cpp:61-103."""
    inv_n = (1.0 / float(n_photons)) if (normalize_per_photon and n_photons > 0) else 1.0
    it = np.arange(det.N_theta)
    if normalize_per_solid_angle:
        dom = solid_angle_bin(it, det.dtheta, det.dphi)
    else:
        dom = np.ones(det.N_theta)
    inv_om = np.where(dom > eps, 1.0 / np.where(dom > eps, dom, 1.0), 0.0)
    norm = (inv_n * inv_om)[:, None]
    g = det.grids()
    coherent, incoherent = [], []
    for t in range(det.N_t):
        coherent.append(StokesMatrixProcessed(*[Matrix(data=g[q, t] * norm) for q in range(4)]))
        incoherent.append(StokesMatrixProcessed(*[Matrix(data=g[4 + q, t] * norm) for q in range(4)]))
    dOmega = Matrix(data=np.repeat(dom[:, None], det.N_phi, axis=1))
    return FarFieldCBSProcessed(dOmega, coherent, incoherent)


def postprocess_planar_fluence(det, n_photons: int, normalize_per_photon: bool = True,
                               normalize_per_area: bool = True, eps: float = 1e-30) -> PlanarFluenceProcessed:
    """src/results.cpp:5-36."""
    inv_n = (1.0 / float(n_photons)) if (normalize_per_photon and n_photons > 0) else 1.0
    area = det.dx * det.dy
    norm = inv_n * ((1.0 / area) if (normalize_per_area and area > eps) else 1.0)
    g = det.grids()
    return PlanarFluenceProcessed(*[[Matrix(data=g[q, t] * norm) for t in range(det.N_t)] for q in range(4)])


def postprocess_planar_field(det, n_photons: int, normalize_per_photon: bool = True,
                             normalize_per_area: bool = True, eps: float = 1e-30) -> PlanarFieldProcessed:
    """src/results.cpp:38-58."""
    inv_n = (1.0 / float(n_photons)) if (normalize_per_photon and n_photons > 0) else 1.0
    area = det.dx * det.dy
    norm = inv_n * ((1.0 / area) if (normalize_per_area and area > eps) else 1.0)
    return PlanarFieldProcessed(CMatrix(data=det.Ex.data * norm), CMatrix(data=det.Ey.data * norm))
