"""Absorption r-z(-t) grid (include/luminis/core/absortion.hpp, src/absortion.cpp:13-130).

The GPU transport fills it: at every scattering vertex the absorbed weight
w * mu_a / mu_t of the layer is added to the (r, z) bin of the vertex in time
slice 0 and, when d_t > 0, in the slice of the arrival time."""
from __future__ import annotations

import math

import numpy as np

from .linalg import Matrix


class Absorption:
    def __init__(self, radius: float, depth: float, d_r: float, d_z: float, d_t: float = 0.0, t_max: float = 0.0):
        self.radius, self.depth, self.d_r, self.d_z = float(radius), float(depth), float(d_r), float(d_z)
        self.d_t, self.t_max = float(d_t), float(t_max)
        self.n_t = int(math.ceil(self.t_max / self.d_t)) + 1 if self.d_t > 0.0 else 1
        self.n_r = int(self.radius / self.d_r) + 1
        self.n_z = int(self.depth / self.d_z) + 1
        self.buffer = np.zeros(self.n_t * self.n_r * self.n_z)

    def grids(self) -> np.ndarray:
        return self.buffer.reshape(self.n_t, self.n_r, self.n_z)

    @property
    def time_slices(self):
        return [Matrix(data=self.grids()[k]) for k in range(self.n_t)]

    def record_absorption(self, photon, d_weight: float) -> None:
        """Absorption::record_absorption (absortion.cpp:47-77) for a photon object
        with pos, launch_time, opticalpath and velocity."""
        x, y, z = photon.pos.x, photon.pos.y, photon.pos.z
        r = math.sqrt(x * x + y * y)
        if r > self.radius or z < 0.0 or z > self.depth:
            return
        i_r, i_z = int(math.floor(r / self.d_r)), int(math.floor(z / self.d_z))
        if i_r >= self.n_r or i_z >= self.n_z:
            return
        g = self.grids()
        g[0, i_r, i_z] += d_weight
        if self.d_t > 0.0:
            t = photon.launch_time + photon.opticalpath / photon.velocity
            k = int(math.floor(t / self.d_t)) + 1
            if 1 <= k < self.n_t:
                g[k, i_r, i_z] += d_weight

    def get_absorption_image(self, n_photons: int, time_index: int = 0) -> Matrix:
        """Absorbed weight per photon and per voxel volume, mirrored about r = 0
        (absortion.cpp:79-118)."""
        if time_index < 0 or time_index >= self.n_t:
            return Matrix(0, 0)
        grid = self.grids()[time_index]
        n_r, len_z = grid.shape
        image = np.zeros((2 * n_r, len_z))
        for i in range(n_r):
            r_in, r_out = i * self.d_r, (i + 1) * self.d_r
            voxel = math.pi * (r_out * r_out - r_in * r_in) * self.d_z
            per_vol = grid[i] / (float(n_photons) * voxel)
            image[n_r - i] = per_vol
            if i != 0:
                image[n_r + i] = per_vol
        return Matrix(data=image)

    def get_total_image(self, n_photons: int) -> Matrix:
        return self.get_absorption_image(n_photons, 0)

    def clone(self) -> "Absorption":
        return Absorption(self.radius, self.depth, self.d_r, self.d_z, self.d_t, self.t_max)

    def merge_from(self, other: "Absorption") -> None:
        self.buffer += other.buffer

    def _fill_desc(self, d, out_offset: int) -> None:
        d.enabled = 1
        d.n_r, d.n_z, d.n_t = self.n_r, self.n_z, self.n_t
        d.radius, d.depth, d.d_r, d.d_z, d.d_t, d.t_max = (self.radius, self.depth, self.d_r, self.d_z, self.d_t,
                                                           self.t_max)
        d.out_offset = out_offset
