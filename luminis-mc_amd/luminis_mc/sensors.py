"""Photon source and sensors.

Laser: include/luminis/core/laser.hpp:45-90, src/laser.cpp:29-99.
Sensors: include/luminis/core/detector.hpp:95-699, src/detector.cpp.  Each
sensor owns one contiguous float64 accumulator in the device layout; the
reference's per-time-bin Matrix lists (``S0_coh[t]``, ``S0[t]``) and the
complex ``Ex``/``Ey`` grids are views into it, so results accumulate (+=)
across runs exactly as the reference's merge_from does.
"""
from __future__ import annotations

import enum
import math

import numpy as np

from . import _lib
from .linalg import CMatrix, CVec2, Matrix, Vec3


class LaserSource(enum.IntEnum):
    Point = 0
    Uniform = 1
    Gaussian = 2


class TemporalProfile(enum.IntEnum):
    Delta = 0
    Gaussian = 1
    TopHat = 2
    Exponential = 3
    PulseTrain = 4
    CW = 5


class CrossingDirection(enum.IntEnum):
    Forward = 0
    Backward = 1
    Both = 2


GEOMETRY_EPSILON = 1e-9


class Laser:
    """Launch state: position (0,0,1e-9), direction +z, frame m=+x n=+y (laser.cpp:29-40)."""

    def __init__(self, m_state: complex, n_state: complex, wavelength: float, sigma: float,
                 source_type: LaserSource):
        self.position = Vec3(0, 0, GEOMETRY_EPSILON)
        self.direction = Vec3(0, 0, 1)
        self.local_m = Vec3(1, 0, 0)
        self.local_n = Vec3(0, 1, 0)
        self.polarization = CVec2(complex(m_state), complex(n_state))
        self.wavelength = float(wavelength)
        self.sigma = float(sigma)
        self.source_type = LaserSource(source_type)
        self.temporal_profile = TemporalProfile.Delta
        self.pulse_duration = 0.0
        self.repetition_rate = 0.0
        self.time_offset = 0.0

    def set_temporal_profile(self, profile, pulse_duration: float = 0.0, repetition_rate: float = 0.0,
                             time_offset: float = 0.0) -> None:
        self.temporal_profile = TemporalProfile(profile)
        self.pulse_duration = float(pulse_duration)
        self.repetition_rate = float(repetition_rate)
        self.time_offset = float(time_offset)


class Sensor:
    """Sensor plane with stackable filters (detector.hpp:95-180, detector.cpp:314-417)."""

    kind = 0

    def __init__(self, z: float, absorb: bool = True, estimator: bool = False):
        self.id = 0
        self.origin = Vec3(0, 0, z)
        self.normal = Vec3(0, 0, 1)
        self.backward_normal = Vec3(0, 0, -1)
        self.m_polarization = Vec3(1, 0, 0)
        self.n_polarization = Vec3(0, 1, 0)
        self.hits = 0
        self.absorb_photons = bool(absorb)
        self.estimator_enabled = bool(estimator) if z == 0.0 else False
        self.filter_theta_enabled = False
        self.filter_theta_min = self.filter_theta_max = 0.0
        self._cache_cos_theta_min = self._cache_cos_theta_max = 0.0
        self.filter_phi_enabled = False
        self.filter_phi_min = self.filter_phi_max = 0.0
        self.filter_position_enabled = False
        self.filter_x_min = self.filter_x_max = self.filter_y_min = self.filter_y_max = 0.0
        self.filter_direction_enabled = False
        self.filter_direction = CrossingDirection.Both
        self.filter_events_enabled = False
        self.filter_events_min = self.filter_events_max = 0
        self.buffer = np.zeros(0)

    def set_theta_limit(self, min_theta: float, max_theta: float) -> None:
        self.filter_theta_enabled = True
        self.filter_theta_min, self.filter_theta_max = float(min_theta), float(max_theta)
        c1, c2 = math.cos(min_theta), math.cos(max_theta)
        self._cache_cos_theta_min, self._cache_cos_theta_max = min(c1, c2), max(c1, c2)

    def set_phi_limit(self, min_phi: float, max_phi: float) -> None:
        self.filter_phi_enabled = True
        self.filter_phi_min, self.filter_phi_max = float(min_phi), float(max_phi)

    def set_position_limit(self, x_min: float, x_max: float, y_min: float, y_max: float) -> None:
        self.filter_position_enabled = True
        self.filter_x_min, self.filter_x_max = float(x_min), float(x_max)
        self.filter_y_min, self.filter_y_max = float(y_min), float(y_max)

    def set_direction_limit(self, direction) -> None:
        self.filter_direction_enabled = True
        self.filter_direction = CrossingDirection(direction)

    def set_events_limit(self, min_events: int, max_events: int) -> None:
        self.filter_events_enabled = True
        self.filter_events_min, self.filter_events_max = int(min_events), int(max_events)

    # -- device descriptor
    def out_len(self) -> int:
        return int(self.buffer.size)

    def _fill_desc(self, d: _lib.SensorDesc) -> None:
        d.kind = self.kind
        d.absorb = int(self.absorb_photons)
        d.estimator = int(self.estimator_enabled)
        d.z = self.origin.z
        d.f_dir_en = int(self.filter_direction_enabled)
        d.f_dir = int(self.filter_direction)
        d.f_ev_en = int(self.filter_events_enabled)
        d.f_ev_min, d.f_ev_max = self.filter_events_min, self.filter_events_max
        d.f_th_en = int(self.filter_theta_enabled)
        d.cos_th_min, d.cos_th_max = self._cache_cos_theta_min, self._cache_cos_theta_max
        d.f_ph_en = int(self.filter_phi_enabled)
        d.ph_min, d.ph_max = self.filter_phi_min, self.filter_phi_max
        d.f_pos_en = int(self.filter_position_enabled)
        d.x_min, d.x_max = self.filter_x_min, self.filter_x_max
        d.y_min, d.y_max = self.filter_y_min, self.filter_y_max

    def clone(self):
        import copy
        c = copy.deepcopy(self)
        c.buffer = np.zeros_like(self.buffer)
        c.hits = 0
        return c


def _matrix_list(arr: np.ndarray, cls=Matrix):
    return [cls(data=arr[t]) for t in range(arr.shape[0])]


class FarFieldCBSSensor(Sensor):
    """Far-field coherent backscattering detector at z = 0, absorbing
    (detector.cpp:851-907).  Bin counts from widths: N = ceil(max / width)."""

    kind = _lib.SENSOR_FARFIELD_CBS

    def __init__(self, theta_max: float, phi_max: float, len_t: float, d_theta: float, d_phi: float,
                 d_t: float, estimator: bool = False):
        super().__init__(0.0, True, estimator)
        self.theta_max, self.phi_max, self.t_max = float(theta_max), float(phi_max), float(len_t)
        self.dtheta, self.dphi, self.dt = float(d_theta), float(d_phi), float(d_t)
        self.N_t = 1 if self.dt == 0 else int(math.ceil(self.t_max / self.dt)) + 1
        self.N_theta = int(math.ceil(self.theta_max / self.dtheta))
        self.N_phi = int(math.ceil(self.phi_max / self.dphi))
        self.theta_pp_max = 0.0
        self.phi_explicit = False
        self.phi_values = [(j + 0.5) * self.dphi for j in range(self.N_phi)]
        self.buffer = np.zeros(8 * self.N_t * self.N_theta * self.N_phi)

    def set_phi_slices(self, phi_angles) -> None:
        """Explicit phi columns, sorted and de-duplicated; resets the grids
        (detector.cpp:909-943)."""
        angles = [float(a) for a in phi_angles]
        if not angles:
            raise ValueError("set_phi_slices: phi_angles must be non-empty")
        for a in angles:
            if a < 0.0 or a > 2.0 * math.pi:
                raise ValueError("set_phi_slices: every angle must be in [0, 2*pi]")
        self.phi_values = sorted(set(angles))
        self.phi_explicit = True
        self.N_phi = len(self.phi_values)
        self.buffer = np.zeros(8 * self.N_t * self.N_theta * self.N_phi)

    def grids(self) -> np.ndarray:
        return self.buffer.reshape(8, self.N_t, self.N_theta, self.N_phi)

    S0_coh = property(lambda self: _matrix_list(self.grids()[0]))
    S1_coh = property(lambda self: _matrix_list(self.grids()[1]))
    S2_coh = property(lambda self: _matrix_list(self.grids()[2]))
    S3_coh = property(lambda self: _matrix_list(self.grids()[3]))
    S0_incoh = property(lambda self: _matrix_list(self.grids()[4]))
    S1_incoh = property(lambda self: _matrix_list(self.grids()[5]))
    S2_incoh = property(lambda self: _matrix_list(self.grids()[6]))
    S3_incoh = property(lambda self: _matrix_list(self.grids()[7]))

    def _fill_desc(self, d):
        super()._fill_desc(d)
        for i, v in enumerate([self.theta_max, self.phi_max, self.t_max, self.dtheta, self.dphi, self.dt]):
            d.g[i] = v
        d.n[0], d.n[1], d.n[2] = self.N_theta, self.N_phi, self.N_t
        d.theta_pp_max = self.theta_pp_max


class PlanarFluenceSensor(Sensor):
    """Planar Stokes fluence (x, y, t) grid (detector.cpp:704-841)."""

    kind = _lib.SENSOR_PLANAR_FLUENCE

    def __init__(self, z: float, len_x: float, len_y: float, len_t: float, dx: float, dy: float, dt: float,
                 absorb: bool = True, estimator: bool = False):
        super().__init__(z, absorb, estimator)
        self.len_x, self.len_y, self.len_t = float(len_x), float(len_y), float(len_t)
        self.dx, self.dy, self.dt = float(dx), float(dy), float(dt)
        self.N_t = 1 if self.dt == 0 else int(math.ceil(self.len_t / self.dt)) + 1
        self.N_x = int(math.ceil(self.len_x / self.dx))
        self.N_y = int(math.ceil(self.len_y / self.dy))
        self.buffer = np.zeros(4 * self.N_t * self.N_x * self.N_y)
        self.set_position_limit(-0.5 * self.len_x, 0.5 * self.len_x, -0.5 * self.len_y, 0.5 * self.len_y)

    def grids(self) -> np.ndarray:
        return self.buffer.reshape(4, self.N_t, self.N_x, self.N_y)

    S0 = property(lambda self: _matrix_list(self.grids()[0]))
    S1 = property(lambda self: _matrix_list(self.grids()[1]))
    S2 = property(lambda self: _matrix_list(self.grids()[2]))
    S3 = property(lambda self: _matrix_list(self.grids()[3]))

    def _fill_desc(self, d):
        super()._fill_desc(d)
        for i, v in enumerate([self.len_x, self.len_y, self.len_t, self.dx, self.dy, self.dt]):
            d.g[i] = v
        d.n[0], d.n[1], d.n[2] = self.N_x, self.N_y, self.N_t


class PlanarFieldSensor(Sensor):
    """Complex Jones field on an (x, y) grid (detector.cpp:490-568)."""

    kind = _lib.SENSOR_PLANAR_FIELD

    def __init__(self, z: float, len_x: float, len_y: float, dx: float, dy: float, absorb: bool = True,
                 estimator: bool = False):
        super().__init__(z, absorb, estimator)
        self.len_x, self.len_y = float(len_x), float(len_y)
        self.dx, self.dy = float(dx), float(dy)
        self.N_x = int(math.ceil(self.len_x / self.dx))
        self.N_y = int(math.ceil(self.len_y / self.dy))
        self.buffer = np.zeros(4 * self.N_x * self.N_y)
        self.set_position_limit(-0.5 * self.len_x, 0.5 * self.len_x, -0.5 * self.len_y, 0.5 * self.len_y)

    def _complex(self, which: int) -> np.ndarray:
        n = self.N_x * self.N_y
        return self.buffer[2 * n * which: 2 * n * (which + 1)].view(np.complex128).reshape(self.N_x, self.N_y)

    Ex = property(lambda self: CMatrix(data=self._complex(0)))
    Ey = property(lambda self: CMatrix(data=self._complex(1)))

    def _fill_desc(self, d):
        super()._fill_desc(d)
        for i, v in enumerate([self.len_x, self.len_y, self.dx, self.dy]):
            d.g[i] = v
        d.n[0], d.n[1] = self.N_x, self.N_y


class FarFieldFluenceSensor(Sensor):
    """Exit-direction Stokes histogram at z = 0 (no reference counterpart): the
    forward-field term of the CBS deposit (detector.cpp:1237-1265) for every
    exiting photon, on (theta, phi) bins of widths d_theta, d_phi."""

    kind = _lib.SENSOR_FARFIELD_FLUENCE

    def __init__(self, theta_max: float, phi_max: float, d_theta: float, d_phi: float, absorb: bool = True):
        super().__init__(0.0, absorb, False)
        self.theta_max, self.phi_max = float(theta_max), float(phi_max)
        self.dtheta, self.dphi = float(d_theta), float(d_phi)
        self.N_theta = int(math.ceil(self.theta_max / self.dtheta))
        self.N_phi = int(math.ceil(self.phi_max / self.dphi))
        self.buffer = np.zeros(4 * self.N_theta * self.N_phi)

    def grids(self) -> np.ndarray:
        return self.buffer.reshape(4, self.N_theta, self.N_phi)

    S0 = property(lambda self: Matrix(data=self.grids()[0]))
    S1 = property(lambda self: Matrix(data=self.grids()[1]))
    S2 = property(lambda self: Matrix(data=self.grids()[2]))
    S3 = property(lambda self: Matrix(data=self.grids()[3]))

    def _fill_desc(self, d):
        super()._fill_desc(d)
        for i, v in enumerate([self.theta_max, self.phi_max, self.dtheta, self.dphi]):
            d.g[i] = v
        d.n[0], d.n[1] = self.N_theta, self.N_phi


class ReflectanceSensor(Sensor):
    """Sum of the weights of photons crossing its plane (no reference
    counterpart; equals the weight sum of a PhotonRecordSensor at the same z).
    With the default backward-direction filter at z = 0 this is the total
    diffuse reflectance times N."""

    kind = _lib.SENSOR_REFLECTANCE

    def __init__(self, z: float = 0.0, absorb: bool = True):
        super().__init__(z, absorb, False)
        self.set_direction_limit(CrossingDirection.Backward)
        self.buffer = np.zeros(1)

    @property
    def total_weight(self) -> float:
        return float(self.buffer[0])


class StatisticsSensor(Sensor):
    """Histograms of the photons crossing its plane: scattering events, exit
    polar and azimuthal angles, penetration depth, arrival time and weight
    (detector.cpp:1571-1888); each histogram is off until its set_* call."""

    kind = _lib.SENSOR_STATISTICS
    _KEYS = ("events", "theta", "phi", "depth", "time", "weight")

    def __init__(self, z: float, absorb: bool = False):
        super().__init__(z, absorb, False)
        for k in self._KEYS:
            setattr(self, f"{k}_histogram_bins_set", False)
        self.max_events = 0
        self.min_theta = self.max_theta = self.dtheta = 0.0
        self.n_bins_theta = 0
        self.min_phi = self.max_phi = self.dphi = 0.0
        self.n_bins_phi = 0
        self.max_depth = self.ddepth = 0.0
        self.n_bins_depth = 0
        self.h_max_time = self.h_dtime = 0.0
        self.n_bins_time = 0
        self.max_weight = self.dweight = 0.0
        self.n_bins_weight = 0
        self.N_t, self.dt, self.t_max = 1, 0.0, 0.0
        self._resize()

    # -- histogram configuration (detector.cpp:1802-1888)
    def set_time_resolution(self, len_t: float, dt: float) -> None:
        self.t_max, self.dt = float(len_t), float(dt)
        if self.dt <= 0.0:
            self.N_t, self.dt = 1, 0.0
        else:
            self.N_t = int(math.ceil(len_t / dt)) + 1
        self._resize()

    def set_events_histogram_bins(self, max_events: int) -> None:
        self.max_events = int(max_events)
        self.events_histogram_bins_set = True
        self._resize()

    def set_theta_histogram_bins(self, min_theta: float, max_theta: float, n_bins: int) -> None:
        self.min_theta, self.max_theta, self.n_bins_theta = float(min_theta), float(max_theta), int(n_bins)
        self.dtheta = (self.max_theta - self.min_theta) / self.n_bins_theta
        self.theta_histogram_bins_set = True
        self._resize()

    def set_phi_histogram_bins(self, min_phi: float, max_phi: float, n_bins: int) -> None:
        self.min_phi, self.max_phi, self.n_bins_phi = float(min_phi), float(max_phi), int(n_bins)
        self.dphi = (self.max_phi - self.min_phi) / self.n_bins_phi
        self.phi_histogram_bins_set = True
        self._resize()

    def set_depth_histogram_bins(self, max_depth: float, n_bins: int) -> None:
        self.max_depth, self.n_bins_depth = float(max_depth), int(n_bins)
        self.ddepth = self.max_depth / self.n_bins_depth
        self.depth_histogram_bins_set = True
        self._resize()

    def set_time_histogram_bins(self, max_time: float, n_bins: int) -> None:
        self.h_max_time, self.n_bins_time = float(max_time), max(1, int(n_bins))
        self.h_dtime = self.h_max_time / int(n_bins)
        self.time_histogram_bins_set = True
        self._resize()

    def set_weight_histogram_bins(self, max_weight: float, n_bins: int) -> None:
        self.max_weight, self.n_bins_weight = float(max_weight), int(n_bins)
        self.dweight = self.max_weight / self.n_bins_weight
        self.weight_histogram_bins_set = True
        self._resize()

    # -- accumulator layout
    def _sizes(self):
        nt = self.N_t
        return [("events", nt, self.max_events), ("theta", nt, self.n_bins_theta), ("phi", nt, self.n_bins_phi),
                ("depth", nt, self.n_bins_depth), ("time", 1, self.n_bins_time), ("weight", 1, self.n_bins_weight)]

    def _resize(self) -> None:
        n = sum(a * b for k, a, b in self._sizes() if getattr(self, f"{k}_histogram_bins_set"))
        self.buffer = np.zeros(n)

    def _block(self, key):
        off = 0
        for k, a, b in self._sizes():
            if not getattr(self, f"{k}_histogram_bins_set"):
                continue
            if k == key:
                blk = np.rint(self.buffer[off:off + a * b]).astype(np.int64).reshape(a, b)
                return blk.tolist() if k in ("events", "theta", "phi", "depth") else blk[0].tolist()
            off += a * b
        return [] if key in ("time", "weight") else []

    events_histogram = property(lambda self: self._block("events"))
    theta_histogram = property(lambda self: self._block("theta"))
    phi_histogram = property(lambda self: self._block("phi"))
    depth_histogram = property(lambda self: self._block("depth"))
    time_histogram = property(lambda self: self._block("time"))
    weight_histogram = property(lambda self: self._block("weight"))

    def stat_params(self) -> np.ndarray:
        flags = [float(getattr(self, f"{k}_histogram_bins_set")) for k in self._KEYS]
        return np.array(flags + [self.max_events, self.min_theta, self.max_theta, self.n_bins_theta, self.dtheta,
                                 self.min_phi, self.max_phi, self.n_bins_phi, self.dphi,
                                 self.max_depth, self.n_bins_depth, self.ddepth,
                                 self.h_max_time, self.n_bins_time, self.h_dtime,
                                 self.max_weight, self.n_bins_weight, self.dweight,
                                 self.N_t, self.dt, self.t_max], dtype=np.float64)


class PhotonRecord:
    """One detected photon's state (photon.hpp:259-291)."""

    __slots__ = ("events", "penetration_depth", "launch_time", "arrival_time", "opticalpath", "weight", "k",
                 "position_first_scattering", "position_last_scattering", "position_detector", "direction", "m",
                 "n", "polarization_forward", "polarization_reverse")

    def __init__(self):
        self.events = 0
        self.penetration_depth = self.launch_time = self.arrival_time = self.opticalpath = 0.0
        self.weight = self.k = 0.0
        self.position_first_scattering = Vec3()
        self.position_last_scattering = Vec3()
        self.position_detector = Vec3()
        self.direction = Vec3(0, 0, 1)
        self.m = Vec3(1, 0, 0)
        self.n = Vec3(0, 1, 0)
        self.polarization_forward = CVec2(1, 0)
        self.polarization_reverse = CVec2(0, 0)

    @classmethod
    def from_row(cls, r) -> "PhotonRecord":
        """From one LMC_RECORD_FIELDS row of lmc_scene_records."""
        p = cls()
        p.events = int(r[1])
        p.penetration_depth, p.launch_time, p.arrival_time, p.opticalpath, p.weight, p.k = (float(x) for x in r[2:8])
        p.position_first_scattering = Vec3(*r[8:11])
        p.position_last_scattering = Vec3(*r[11:14])
        p.position_detector = Vec3(*r[14:17])
        p.direction = Vec3(*r[17:20])
        p.m = Vec3(*r[20:23])
        p.n = Vec3(*r[23:26])
        p.polarization_forward = CVec2(complex(r[26], r[27]), complex(r[28], r[29]))
        p.polarization_reverse = CVec2(complex(r[30], r[31]), complex(r[32], r[33]))
        return p


class PhotonRecordSensor(Sensor):
    """Keeps a PhotonRecord per accepted crossing (detector.cpp:427-485).
    Records arrive in crossing-queue order, not photon order; ``records`` is
    the same data as an (n, 34) array (columns as LMC_RECORD_FIELDS)."""

    kind = _lib.SENSOR_PHOTON_RECORD

    def __init__(self, z: float, absorb: bool = True):
        super().__init__(z, absorb, False)
        self.records = np.zeros((0, _lib.RECORD_FIELDS))

    @property
    def recorded_photons(self) -> list:
        return [PhotonRecord.from_row(r) for r in self.records]

    def add_records(self, rows: np.ndarray) -> None:
        """merge_from (detector.cpp:455-461): append another copy's records."""
        if len(rows):
            self.records = np.concatenate([self.records, np.asarray(rows, dtype=np.float64)])

    def clone(self):
        c = super().clone()
        c.records = np.zeros((0, _lib.RECORD_FIELDS))
        return c


class SensorsGroup:
    """Container; add_detector stores a clone and returns the internal copy
    (python_module.cpp:315-322, detector.cpp:166-176)."""

    def __init__(self):
        self.detectors: list[Sensor] = []

    def add_detector(self, detector: Sensor) -> Sensor:
        c = detector.clone()
        c.id = len(self.detectors)
        self.detectors.append(c)
        return c

    def clone(self) -> "SensorsGroup":
        g = SensorsGroup()
        for d in self.detectors:
            g.add_detector(d)
        return g
