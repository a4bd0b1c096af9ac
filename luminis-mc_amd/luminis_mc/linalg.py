"""Small value types of the reference's Python surface (include/luminis/math/vec.hpp,
bound in python/python_module.cpp:32-111), backed by numpy.

``Matrix``/``CMatrix`` expose the same row-major float64 / complex128 layout as
the reference's buffer protocol, so ``numpy.asarray(det.S0_coh[0])`` works the
same way; sensor accumulators are views into one contiguous array per sensor.
"""
from __future__ import annotations

import numpy as np


class Vec3:
    __slots__ = ("x", "y", "z")

    def __init__(self, x: float = 0.0, y: float = 0.0, z: float = 0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def __repr__(self):
        return f"Vec3({self.x}, {self.y}, {self.z})"

    def as_tuple(self):
        return (self.x, self.y, self.z)


class Vec2:
    __slots__ = ("x", "y")

    def __init__(self, x: float = 0.0, y: float = 0.0):
        self.x, self.y = float(x), float(y)


class CVec2:
    __slots__ = ("m", "n")

    def __init__(self, m: complex = 0j, n: complex = 0j):
        self.m, self.n = complex(m), complex(n)


class Matrix:
    """Row-major float64 matrix; ``data`` may be a view into a sensor buffer."""

    _dtype = np.float64

    def __init__(self, rows: int = 0, cols: int = 0, data=None):
        if data is None:
            data = np.zeros((int(rows), int(cols)), dtype=self._dtype)
        self.data = data

    @property
    def rows(self) -> int:
        return int(self.data.shape[0])

    @property
    def cols(self) -> int:
        return int(self.data.shape[1])

    def get(self, i: int, j: int):
        return self.data[i, j]

    def set(self, i: int, j: int, value) -> None:
        self.data[i, j] = value

    def get_numpy(self):
        return np.array(self.data, copy=True)

    def __array__(self, dtype=None, copy=None):
        return self.data if dtype is None else self.data.astype(dtype)

    def __repr__(self):
        return f"{type(self).__name__}({self.rows}x{self.cols})"


class CMatrix(Matrix):
    _dtype = np.complex128

    @staticmethod
    def identity(size: int) -> "CMatrix":
        return CMatrix(data=np.eye(int(size), dtype=np.complex128))


def dot(a: Vec3, b: Vec3) -> float:
    return a.x * b.x + a.y * b.y + a.z * b.z


def cross(a: Vec3, b: Vec3) -> Vec3:
    """Normalised cross product (src/vec.cpp:15-28)."""
    r = Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x)
    n = norm(r)
    if n > 0.0:
        r = Vec3(r.x / n, r.y / n, r.z / n)
    return r


def norm(v: Vec3) -> float:
    return float(np.sqrt(dot(v, v)))


def matcmul(A: CMatrix, B: CMatrix, C: CMatrix) -> None:
    if A.cols != B.rows or A.rows != C.rows or B.cols != C.cols:
        raise ValueError("CMatrix multiplication dimension mismatch")
    C.data[...] = A.data @ B.data


def matcmulscalar(scalar: float, A: CMatrix) -> None:
    A.data *= scalar


def calculate_rotation_angle(n_from: Vec3, n_to: Vec3) -> float:
    nu, nv = norm(n_from), norm(n_to)
    if nu < 1e-12 or nv < 1e-12:
        return 0.0
    c = min(1.0, max(-1.0, dot(n_from, n_to) / (nu * nv)))
    return float(np.arccos(c))


class Rng:
    """xoshiro256++ with SplitMix64 seeding (include/luminis/math/rng.hpp:53-89)."""

    _M = (1 << 64) - 1

    def __init__(self, seed: int | None = None):
        if seed is None:
            seed = int.from_bytes(np.random.bytes(8), "little")
        self.seed_state(int(seed))

    def seed_state(self, seed: int) -> None:
        z = seed & self._M
        s = []
        for _ in range(4):
            z = (z + 0x9E3779B97F4A7C15) & self._M
            x = z
            x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & self._M
            x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & self._M
            s.append(x ^ (x >> 31))
        self.s = s

    def next_u64(self) -> int:
        M = self._M
        s = self.s
        rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M
        result = (rotl((s[0] + s[3]) & M, 23) + s[0]) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        return result

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * 2.0 ** -53

    def normal(self, mean: float, stddev: float) -> float:
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        return float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)) * stddev + mean


def mix_seed(base: int, tid: int) -> int:
    """Per-photon stream key (src/rng.cpp:5-10)."""
    M = (1 << 64) - 1
    z = (base + 0x9E3779B97F4A7C15 * (tid + 1)) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)
