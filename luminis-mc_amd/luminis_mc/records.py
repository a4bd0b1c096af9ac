"""Typed run parameters and saved results (python/luminis_mc/records.py:1-558).

Two groups of dataclasses, with the reference's names and fields so that code
reading results keeps working:

* parameters: `SimParams` (run controls, laser, host index, layers, free-form
  extras) with `RunParams`, `LaserParams`, `MediumParams`, `LayerParams`; a
  JSON form (`to_json` / `from_json`) stored in the results file's ``/params``
  attribute ``params_json``;
* results: `SensorMeta` and one ``*Result`` per sensor type, built from a
  sensor's saved meta attributes and data arrays by `from_h5`, the
  post-processed ``*ProcessedResult`` and `AbsorptionResult`.

`encode_photon_records` writes a PhotonRecordSensor as parallel arrays under
``records/`` and is registered on the schema table here.
"""
from __future__ import annotations

import ast
import json
from dataclasses import dataclass, field, replace
from typing import Any, Callable, Dict, List, Optional

import numpy as np

from . import schema


def _to_numpy(x) -> Optional[np.ndarray]:
    """A Matrix / CMatrix, a list of them (stacked on a new first axis) or an
    array-like as ndarray; None for what has no numeric array form."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        return None if x.dtype == object else x
    if hasattr(x, "get_numpy"):
        a = np.asarray(x.get_numpy())
        return None if a.dtype == object else a
    if isinstance(x, (list, tuple)):
        if not x:
            return None
        parts = [_to_numpy(y) for y in x]
        if all(p is not None for p in parts) and len({p.shape for p in parts}) == 1:
            return np.stack(parts, axis=0)
    a = np.asarray(x)
    return None if a.dtype == object else a


def _py(v) -> Any:
    """A value read from a file as plain Python: bytes decoded, 0-d arrays and
    numpy scalars unwrapped, arrays left as arrays."""
    if isinstance(v, bytes):
        return v.decode("utf-8", errors="replace")
    if isinstance(v, np.ndarray):
        return v.item() if v.ndim == 0 else v
    if isinstance(v, np.generic):
        return v.item()
    return v


# ------------------------------------------------------------------ params ----
@dataclass
class MediumParams:
    kind: str
    mu_s: float
    mu_a: float
    mu_t: float
    n_particle: float
    n_medium: float
    wavelength: float
    radius: Optional[float] = None
    mean_free_path: Optional[float] = None


@dataclass
class LayerParams:
    """A layer's z range and content: kind "homogeneous" (one `medium`) or
    "mixture" (`species` with their `number_densities` and the totals)."""
    z_min: float
    z_max: float
    kind: str = "homogeneous"
    medium: Optional[MediumParams] = None
    species: Optional[List[MediumParams]] = None
    number_densities: Optional[List[float]] = None
    mu_s_total: Optional[float] = None
    mu_a_total: Optional[float] = None
    mu_t_total: Optional[float] = None
    mfp_total: Optional[float] = None

    def to_dict(self) -> Dict[str, Any]:
        d: Dict[str, Any] = {"z_min": self.z_min, "z_max": self.z_max, "kind": self.kind}
        if self.kind == "mixture":
            d.update(species=[dict(vars(s)) for s in (self.species or [])],
                     number_densities=list(self.number_densities or []), mu_s_total=self.mu_s_total,
                     mu_a_total=self.mu_a_total, mu_t_total=self.mu_t_total, mfp_total=self.mfp_total)
        else:
            d["medium"] = dict(vars(self.medium)) if self.medium is not None else None
        return d

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "LayerParams":
        if d.get("kind", "homogeneous") == "mixture":
            return cls(z_min=d["z_min"], z_max=d["z_max"], kind="mixture",
                       species=[MediumParams(**s) for s in d.get("species", [])],
                       number_densities=list(d.get("number_densities", [])), mu_s_total=d.get("mu_s_total"),
                       mu_a_total=d.get("mu_a_total"), mu_t_total=d.get("mu_t_total"), mfp_total=d.get("mfp_total"))
        m = d.get("medium")
        return cls(z_min=d["z_min"], z_max=d["z_max"], kind="homogeneous",
                   medium=MediumParams(**m) if m is not None else None)


@dataclass
class LaserParams:
    source_type: str
    wavelength: float
    sigma: float
    m_state: complex
    n_state: complex


@dataclass
class RunParams:
    n_photons: int
    n_threads: int
    seed: int
    max_events: int
    track_reverse_paths: bool


@dataclass
class SimParams:
    """A run's configuration; `extra` holds derived quantities (g, l*, CBS
    cone width, ...) that live on no simulation object."""
    run: RunParams
    laser: LaserParams
    sample_n_medium: float
    layers: List[LayerParams]
    extra: Dict[str, Any] = field(default_factory=dict)

    def to_dict(self) -> Dict[str, Any]:
        las = self.laser
        return {"run": dict(vars(self.run)),
                "laser": {"source_type": las.source_type, "wavelength": las.wavelength, "sigma": las.sigma,
                          "m_state": [las.m_state.real, las.m_state.imag],
                          "n_state": [las.n_state.real, las.n_state.imag]},
                "sample_n_medium": self.sample_n_medium,
                "layers": [L.to_dict() for L in self.layers],
                "extra": self.extra}

    def to_json(self) -> str:
        return json.dumps(self.to_dict(), default=str)

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "SimParams":
        las = d["laser"]
        return cls(run=RunParams(**d["run"]),
                   laser=LaserParams(source_type=las["source_type"], wavelength=las["wavelength"],
                                     sigma=las["sigma"], m_state=complex(*las["m_state"]),
                                     n_state=complex(*las["n_state"])),
                   sample_n_medium=d["sample_n_medium"],
                   layers=[LayerParams.from_dict(L) for L in d["layers"]],
                   extra=d.get("extra", {}))

    @classmethod
    def from_json(cls, s: str) -> "SimParams":
        return cls.from_dict(json.loads(s))


# ----------------------------------------------------------------- results ----
@dataclass
class SensorMeta:
    """The base-sensor attributes every saved sensor has, and all of its meta
    attributes in `raw`."""
    type: str
    id: str
    hits: int
    absorb_photons: bool
    estimator_enabled: bool
    origin: np.ndarray
    normal: np.ndarray
    raw: Dict[str, Any] = field(default_factory=dict)

    @classmethod
    def from_raw(cls, m: Dict[str, Any]) -> "SensorMeta":
        g = lambda k, dflt: _py(m.get(k, dflt))  # noqa: E731
        return cls(type=str(g("type", "")), id=str(g("id", "")), hits=int(g("hits", 0)),
                   absorb_photons=bool(g("absorb_photons", False)),
                   estimator_enabled=bool(g("estimator_enabled", False)),
                   origin=np.asarray(m.get("origin", np.zeros(3))), normal=np.asarray(m.get("normal", np.zeros(3))),
                   raw={k: _py(v) for k, v in m.items()})


def _floats(v) -> Optional[np.ndarray]:
    """A stored list of floats: an array, or the text form a list takes when
    written as an attribute ("[0.1, 0.2]")."""
    if v is None:
        return None
    v = _py(v)
    if isinstance(v, str):
        return np.asarray(ast.literal_eval(v), dtype=float)
    return np.asarray(v, dtype=float)


@dataclass
class FarFieldCBSResult:
    """Raw far-field CBS Stokes grids, each (N_t, N_theta, N_phi)."""
    meta: SensorMeta
    S0_coh: np.ndarray
    S1_coh: np.ndarray
    S2_coh: np.ndarray
    S3_coh: np.ndarray
    S0_incoh: np.ndarray
    S1_incoh: np.ndarray
    S2_incoh: np.ndarray
    S3_incoh: np.ndarray
    N_theta: int
    N_phi: int
    N_t: int
    theta_max: float
    phi_max: float
    phi_explicit: bool = False
    phi_values: Optional[np.ndarray] = None

    @property
    def theta(self) -> np.ndarray:
        return np.linspace(0.0, self.theta_max, self.N_theta)

    @property
    def phi(self) -> np.ndarray:
        """The explicit phi slices when the sensor had them, else N_phi even angles over [0, phi_max]."""
        if self.phi_explicit and self.phi_values is not None:
            return np.asarray(self.phi_values, dtype=float)
        return np.linspace(0.0, self.phi_max, self.N_phi)

    @classmethod
    def from_h5(cls, meta: Dict[str, Any], data: Dict[str, np.ndarray]) -> "FarFieldCBSResult":
        g = lambda k, dflt: _py(meta.get(k, dflt))  # noqa: E731
        grids = {k: data[k] for k in schema.SENSOR_SCHEMAS["FarFieldCBSSensor"].data_attrs}
        return cls(meta=SensorMeta.from_raw(meta), **grids, N_theta=int(g("N_theta", 0)), N_phi=int(g("N_phi", 0)),
                   N_t=int(g("N_t", 0)), theta_max=float(g("theta_max", 0.0)), phi_max=float(g("phi_max", 0.0)),
                   phi_explicit=bool(g("phi_explicit", False)), phi_values=_floats(meta.get("phi_values")))


@dataclass
class PlanarFluenceResult:
    """Raw planar-fluence Stokes grids, each (N_t, N_x, N_y)."""
    meta: SensorMeta
    S0: np.ndarray
    S1: np.ndarray
    S2: np.ndarray
    S3: np.ndarray
    N_x: int
    N_y: int
    N_t: int
    dx: float
    dy: float
    dt: float

    @classmethod
    def from_h5(cls, meta: Dict[str, Any], data: Dict[str, np.ndarray]) -> "PlanarFluenceResult":
        g = lambda k, dflt: _py(meta.get(k, dflt))  # noqa: E731
        return cls(meta=SensorMeta.from_raw(meta), S0=data["S0"], S1=data["S1"], S2=data["S2"], S3=data["S3"],
                   N_x=int(g("N_x", 0)), N_y=int(g("N_y", 0)), N_t=int(g("N_t", 0)), dx=float(g("dx", 0.0)),
                   dy=float(g("dy", 0.0)), dt=float(g("dt", 0.0)))


@dataclass
class PlanarFieldResult:
    """Raw complex planar field, Ex and Ey each (N_x, N_y)."""
    meta: SensorMeta
    Ex: np.ndarray
    Ey: np.ndarray
    N_x: int
    N_y: int
    dx: float
    dy: float

    @classmethod
    def from_h5(cls, meta: Dict[str, Any], data: Dict[str, np.ndarray]) -> "PlanarFieldResult":
        g = lambda k, dflt: _py(meta.get(k, dflt))  # noqa: E731
        return cls(meta=SensorMeta.from_raw(meta), Ex=data["Ex"], Ey=data["Ey"], N_x=int(g("N_x", 0)),
                   N_y=int(g("N_y", 0)), dx=float(g("dx", 0.0)), dy=float(g("dy", 0.0)))


@dataclass
class StatisticsResult:
    """The histograms a StatisticsSensor kept (None where a histogram was not set)."""
    meta: SensorMeta
    events_histogram: Optional[np.ndarray]
    theta_histogram: Optional[np.ndarray]
    phi_histogram: Optional[np.ndarray]
    depth_histogram: Optional[np.ndarray]
    time_histogram: Optional[np.ndarray]
    weight_histogram: Optional[np.ndarray]

    @classmethod
    def from_h5(cls, meta: Dict[str, Any], data: Dict[str, np.ndarray]) -> "StatisticsResult":
        names = schema.SENSOR_SCHEMAS["StatisticsSensor"].data_attrs
        return cls(meta=SensorMeta.from_raw(meta), **{k: data.get(k) for k in names})


@dataclass
class PhotonRecordsResult:
    """A PhotonRecordSensor's records as parallel arrays of length n_records."""
    meta: SensorMeta
    n_records: int
    events: np.ndarray
    weight: np.ndarray
    penetration_depth: np.ndarray
    launch_time: np.ndarray
    arrival_time: np.ndarray
    r_first: np.ndarray
    r_last: np.ndarray
    E_forward: np.ndarray
    E_reverse: np.ndarray

    @classmethod
    def from_h5(cls, meta: Dict[str, Any], records: Dict[str, np.ndarray]) -> "PhotonRecordsResult":
        keys = ("events", "weight", "penetration_depth", "launch_time", "arrival_time", "r_first", "r_last",
                "E_forward", "E_reverse")
        return cls(meta=SensorMeta.from_raw(meta), n_records=int(_py(meta.get("n_records", 0))),
                   **{k: records.get(k, np.empty(0)) for k in keys})


@dataclass
class AbsorptionResult:
    """An Absorption grid: `slices` (n_slices, n_r, n_z) of absorbed weight and
    `images` of the same per photon and voxel volume, mirrored about r = 0;
    slice 0 is time-integrated, slice k >= 1 the time bin k."""
    radius: float
    depth: float
    d_r: float
    d_z: float
    d_t: float
    t_max: float
    n_t: int
    slices: np.ndarray
    images: np.ndarray

    @property
    def total(self) -> np.ndarray:
        return self.slices[0]

    @property
    def total_image(self) -> np.ndarray:
        return self.images[0]

    @classmethod
    def from_h5(cls, meta: Dict[str, Any], slices: List[np.ndarray], images: List[np.ndarray]) -> "AbsorptionResult":
        g = lambda k, dflt: _py(meta.get(k, dflt))  # noqa: E731
        return cls(radius=float(g("radius", 0.0)), depth=float(g("depth", 0.0)), d_r=float(g("d_r", 0.0)),
                   d_z=float(g("d_z", 0.0)), d_t=float(g("d_t", 0.0)), t_max=float(g("t_max", 0.0)),
                   n_t=int(g("n_t", 0)), slices=np.stack(slices) if slices else np.empty(0),
                   images=np.stack(images) if images else np.empty(0))


RESULT_BUILDERS: Dict[str, Callable[[Dict[str, Any], Dict[str, np.ndarray]], Any]] = {
    "FarFieldCBSSensor": FarFieldCBSResult.from_h5,
    "PlanarFluenceSensor": PlanarFluenceResult.from_h5,
    "PlanarFieldSensor": PlanarFieldResult.from_h5,
    "StatisticsSensor": StatisticsResult.from_h5,
    "PhotonRecordSensor": PhotonRecordsResult.from_h5,
}


@dataclass
class FarFieldCBSProcessedResult:
    """Post-processed CBS Stokes, coherent and incoherent, each (N_t, N_theta, N_phi)."""
    theta: np.ndarray
    phi: np.ndarray
    coh_s0: np.ndarray
    coh_s1: np.ndarray
    coh_s2: np.ndarray
    coh_s3: np.ndarray
    inc_s0: np.ndarray
    inc_s1: np.ndarray
    inc_s2: np.ndarray
    inc_s3: np.ndarray

    @property
    def N_t(self) -> int:
        return self.coh_s0.shape[0]

    @property
    def N_theta(self) -> int:
        return self.coh_s0.shape[1]

    @property
    def N_phi(self) -> int:
        return self.coh_s0.shape[2]


@dataclass
class PlanarFluenceProcessedResult:
    S0: np.ndarray
    S1: np.ndarray
    S2: np.ndarray
    S3: np.ndarray


@dataclass
class PlanarFieldProcessedResult:
    Ex: np.ndarray
    Ey: np.ndarray


def encode_photon_records(sensor, g, g_meta, g_data, write_dataset, write_attr) -> None:
    """A PhotonRecordSensor under ``records/``: n_records in meta/, and the
    events, weight, penetration depth, launch and arrival times, first and last
    scattering positions (n, 3) and forward and reverse Jones vectors (n, 2,
    complex) of its records as parallel datasets."""
    rows = getattr(sensor, "records", None)
    if rows is not None:
        rows = np.asarray(rows, dtype=np.float64).reshape(-1, 34)
        n = rows.shape[0]
        cols = {"events": rows[:, 1].astype(np.int32), "weight": rows[:, 6], "penetration_depth": rows[:, 2],
                "launch_time": rows[:, 3], "arrival_time": rows[:, 4], "r_first": rows[:, 8:11],
                "r_last": rows[:, 11:14],
                "E_forward": rows[:, [26, 28]] + 1j * rows[:, [27, 29]],
                "E_reverse": rows[:, [30, 32]] + 1j * rows[:, [31, 33]]}
    else:
        rec = list(sensor.recorded_photons)
        n = len(rec)
        v3 = lambda p: [p.x, p.y, p.z]  # noqa: E731
        cols = {"events": np.array([r.events for r in rec], dtype=np.int32),
                "weight": np.array([r.weight for r in rec], dtype=np.float64),
                "penetration_depth": np.array([r.penetration_depth for r in rec], dtype=np.float64),
                "launch_time": np.array([r.launch_time for r in rec], dtype=np.float64),
                "arrival_time": np.array([r.arrival_time for r in rec], dtype=np.float64),
                "r_first": np.array([v3(r.position_first_scattering) for r in rec], dtype=np.float64),
                "r_last": np.array([v3(r.position_last_scattering) for r in rec], dtype=np.float64),
                "E_forward": np.array([[r.polarization_forward.m, r.polarization_forward.n] for r in rec],
                                      dtype=np.complex128),
                "E_reverse": np.array([[r.polarization_reverse.m, r.polarization_reverse.n] for r in rec],
                                      dtype=np.complex128)}
    write_attr(g_meta, "n_records", int(n))
    if n == 0:
        return
    rg = g.require_group("records")
    for k, a in cols.items():
        write_dataset(rg, k, np.ascontiguousarray(a))


schema.SENSOR_SCHEMAS["PhotonRecordSensor"] = replace(schema.SENSOR_SCHEMAS["PhotonRecordSensor"],
                                                      encoder=encode_photon_records)
