"""Experiment persistence: results files (python/luminis_mc/manager.py:1-1288).

`Experiment` makes an output directory (``<base_dir>/<timestamp>_<name>/``),
writes run parameters, sensor results, the Absorption grid and
post-processed arrays into one HDF5 file (``results.h5``) and can keep a copy
of the calling script; `ResultsLoader` reads such a file back, as plain arrays
or as the typed results of `luminis_mc.records`.  The file layout::

    /params                 attrs: params_json + flat scalars; array datasets
    /sensors/<name>/meta    attrs: type, base-sensor attributes, type attributes
    /sensors/<name>/data    datasets: Stokes grids, fields, histograms
    /sensors/<name>/records PhotonRecordSensor records (parallel arrays)
    /absorption/<name>/meta, /absorption/<name>/data/time_slice_<k>[_image]
    /derived/...            post-processed arrays

follows the reference's, with the sensor table of `luminis_mc.schema`, so a
results file written by either is read by the other.  HDF5 goes through the
HDF5 C library (`luminis_mc.h5`).  Arrays above 256 elements are stored
chunked, shuffled and gzip level 4 compressed, smaller ones plain.
"""
from __future__ import annotations

import enum
import os
import shutil
from datetime import datetime
from pathlib import Path
from typing import Any, Dict, List, Optional, Sequence, Union

import numpy as np

from . import h5, schema
from .records import (AbsorptionResult, FarFieldCBSProcessedResult, FarFieldCBSResult, LaserParams, LayerParams,
                      MediumParams, PhotonRecordsResult, PlanarFieldProcessedResult, PlanarFieldResult,
                      PlanarFluenceProcessedResult, PlanarFluenceResult, RESULT_BUILDERS, RunParams, SimParams,
                      StatisticsResult, _py, _to_numpy)


def on_progress(done: int, total: int) -> None:
    """A progress callback printing the percentage done on one line."""
    print(f"\rSimulation progress: {done * 100 // max(1, total)}%", end="", flush=True)
    if done == total:
        print()


def _is_scalar(x) -> bool:
    return isinstance(x, (str, bool, int, float, np.integer, np.floating, np.bool_)) and not isinstance(x, enum.Enum)


def _clean_scalar(x):
    return x.item() if isinstance(x, (np.integer, np.floating, np.bool_)) else x


def _as_array(x) -> Optional[np.ndarray]:
    """An array form of x (arrays, Matrix objects, lists of numbers or of
    equal-shaped matrices), or None."""
    if x is None:
        return None
    if isinstance(x, (list, tuple)) and len(x) == 0:
        return None
    return _to_numpy(x)


def _write_dataset(g, name: str, arr, compression: str = "gzip", clevel: int = 4):
    """(Re)write dataset `name` in group g: chunked, shuffled and compressed when
    it holds more than 256 elements."""
    if name in g:
        del g[name]
    arr = np.asarray(arr)
    big = arr.size > 256
    return g.create_dataset(name, data=arr, compression=compression if big else None,
                            compression_opts=clevel if big else None, shuffle=big)


def _write_attr(g, key: str, val: Any) -> None:
    """One attribute: scalars as they are, enums by name, Vec3 / Vec2 as float
    arrays, CVec2 as a complex pair, anything else as its text; None skipped."""
    if val is None:
        return
    if isinstance(val, enum.Enum) or (not _is_scalar(val) and hasattr(val, "name") and
                                      isinstance(getattr(val, "name"), str)):
        g.attrs[key] = str(val.name)
    elif _is_scalar(val):
        g.attrs[key] = _clean_scalar(val)
    elif all(hasattr(val, a) for a in ("x", "y", "z")):
        g.attrs[key] = np.array([val.x, val.y, val.z], dtype=np.float64)
    elif hasattr(val, "x") and hasattr(val, "y"):
        g.attrs[key] = np.array([val.x, val.y], dtype=np.float64)
    elif hasattr(val, "m") and hasattr(val, "n"):
        g.attrs[key] = np.array([val.m, val.n], dtype=np.complex128)
    else:
        g.attrs[key] = str(val)


# ------------------------------------------------------------ parameters ----
def _medium_params(m) -> MediumParams:
    opt = lambda a: float(getattr(m, a)) if hasattr(m, a) else None  # noqa: E731
    return MediumParams(kind=type(m).__name__, mu_s=float(m.mu_s), mu_a=float(m.mu_a), mu_t=float(m.mu_t),
                        n_particle=float(m.n_particle), n_medium=float(m.n_medium), wavelength=float(m.wavelength),
                        radius=opt("radius"), mean_free_path=opt("mean_free_path"))


def _layer_params(L) -> LayerParams:
    if hasattr(L, "species") and not hasattr(L, "medium"):
        return LayerParams(z_min=float(L.z_min), z_max=float(L.z_max), kind="mixture",
                           species=[_medium_params(s) for s in L.species],
                           number_densities=[float(n) for n in L.number_densities],
                           mu_s_total=float(L.mu_s_total), mu_a_total=float(L.mu_a_total),
                           mu_t_total=float(L.mu_t_total), mfp_total=float(L.mfp_total))
    return LayerParams(z_min=float(L.z_min), z_max=float(L.z_max), kind="homogeneous",
                       medium=_medium_params(L.medium))


def capture_params(config, extra: Optional[Dict[str, Any]] = None) -> SimParams:
    """A SimParams snapshot of a SimConfig (run controls, laser, host index,
    every layer); `extra` is kept as given.  ValueError without a sample or a laser."""
    if getattr(config, "sample", None) is None:
        raise ValueError("capture_params: config.sample is not set.")
    if getattr(config, "laser", None) is None:
        raise ValueError("capture_params: config.laser is not set.")
    las, pol = config.laser, config.laser.polarization
    st = las.source_type
    return SimParams(
        run=RunParams(n_photons=int(getattr(config, "n_photons", 0)), n_threads=int(getattr(config, "n_threads", 1)),
                      seed=int(getattr(config, "seed", 0)), max_events=int(getattr(config, "MAX_EVENTS", 0)),
                      track_reverse_paths=bool(getattr(config, "track_reverse_paths", False))),
        laser=LaserParams(source_type=str(st.name if hasattr(st, "name") else st), wavelength=float(las.wavelength),
                          sigma=float(las.sigma), m_state=complex(pol.m), n_state=complex(pol.n)),
        sample_n_medium=float(config.sample.refractive_index),
        layers=[_layer_params(L) for L in config.sample.layers],
        extra=dict(extra) if extra else {})


def derived_quantities(medium, volume_fraction: float) -> Dict[str, Any]:
    """Scattering efficiency, anisotropy g, scattering and transport mean free
    paths for particles at `volume_fraction` (l_s = 4r / (3 f Q_sca), l* =
    l_s / (1 - g)), CBS cone width 1 / (k l*), size parameter and relative
    index of a single-particle medium, for `capture_params(extra=...)`."""
    ph = medium.phase_function
    q_sca = float(ph.scattering_efficiency())
    g = float(ph.get_anisotropy_factor()[0])
    r, n_m, n_p, lam = float(medium.radius), float(medium.n_medium), float(medium.n_particle), float(medium.wavelength)
    l_s = 4.0 * r / (3.0 * volume_fraction * q_sca)
    l_star = l_s / (1.0 - g)
    k = 2.0 * np.pi * n_m / lam
    return {"volume_fraction": float(volume_fraction), "scattering_efficiency": q_sca, "anisotropy_g": g,
            "mean_free_path": l_s, "transport_mean_free_path": l_star, "theta_coherent": 1.0 / (k * l_star),
            "size_parameter": 2.0 * np.pi * r * n_m / lam, "m_relative": n_p / n_m}


def derived_quantities_mixture(species: Sequence[Any], number_densities: Sequence[float]) -> Dict[str, Any]:
    """Mixture aggregates: per-species mu_s = n_i sigma_s,i and g_i, their sum,
    g_eff = sum(mu_s,i g_i) / sum(mu_s,i), l_s = 1 / sum(mu_s,i), l* = 1 /
    sum(mu_s,i (1 - g_i)) and the CBS cone width 1 / (k l*) with k the host
    medium's wavenumber."""
    species, nd = list(species), [float(n) for n in number_densities]
    if not species:
        raise ValueError("derived_quantities_mixture: species list is empty.")
    if len(species) != len(nd):
        raise ValueError("derived_quantities_mixture: species and number_densities differ in length.")
    mu_s = [n * float(s.phase_function.scattering_cross_section()) for s, n in zip(species, nd)]
    gs = [float(s.phase_function.get_anisotropy_factor()[0]) for s in species]
    tot = float(sum(mu_s))
    if tot <= 0.0:
        raise ValueError("derived_quantities_mixture: total mu_s is not positive "
                         "(check number densities and phase-function cross-sections).")
    red = float(sum(m * (1.0 - g) for m, g in zip(mu_s, gs)))
    l_star = 1.0 / red if red > 0.0 else float("inf")
    k = float(species[0].k)
    return {"mixture": True, "n_species": len(species), "number_densities": nd, "mu_s_per_species": mu_s,
            "anisotropy_g_per_species": gs, "mu_s_total": tot,
            "anisotropy_g_eff": float(sum(m * g for m, g in zip(mu_s, gs)) / tot),
            "scattering_mean_free_path": 1.0 / tot, "transport_mean_free_path": l_star,
            "theta_coherent": 1.0 / (k * l_star) if np.isfinite(l_star) else 0.0}


# ------------------------------------------------------------- Experiment ----
class Experiment:
    """One experiment's output directory and results file.

        with Experiment("run", base_dir="results") as exp:
            exp.save_params(config)
            run_simulation_parallel(config)
            exp.save_sensor(cbs, "farfield_cbs")

    The directory is ``<base_dir>/<YYYY-MM-DD_HH-MM-SS>_<name>`` (or
    ``<base_dir>/<name>`` with timestamped=False); an existing one is replaced.
    """

    def __init__(self, name: str, base_dir: str = "sim_results", h5_name: str = "results.h5",
                 timestamped: bool = True):
        stamp = datetime.now().strftime("%Y-%m-%d_%H-%M-%S")
        self.path = Path(base_dir) / (f"{stamp}_{name}" if timestamped else name)
        if self.path.exists():
            shutil.rmtree(self.path)
        self.path.mkdir(parents=True, exist_ok=True)
        self.h5_path = self.path / h5_name
        self.h5 = h5.File(self.h5_path, "w")
        self.g_params = self.h5.require_group("params")
        self.g_sensors = self.h5.require_group("sensors")
        self.g_derived = self.h5.require_group("derived")
        _write_attr(self.h5, "created_at", datetime.now().isoformat(timespec="seconds"))
        _write_attr(self.h5, "experiment_dir", str(self.path))
        self.readme = self.path / "README.md"
        print(f"Experiment: {self.path}")
        print(f"  HDF5: {self.h5_path}")

    def close(self) -> None:
        try:
            self.h5.flush()
        finally:
            self.h5.close()

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        self.close()

    def log_readme(self, text: str) -> None:
        self.readme.write_text(text, encoding="utf-8")

    def update_log_readme(self, status: str, runtime_s: float, finished_at: str,
                          error_message: Optional[str] = None) -> None:
        """Append a status / runtime / finish-time block (and the error) to README.md."""
        block = f"Status: {status}\nRuntime: {runtime_s:.2f} seconds\nFinished at: {finished_at}"
        if error_message:
            block += f"\nError: {error_message}"
        old = self.readme.read_text(encoding="utf-8") if self.readme.exists() else ""
        self.readme.write_text(old + "\n\n" + block if old else block, encoding="utf-8")

    def log_params(self, units: Optional[Dict[str, str]] = None, **kwargs) -> None:
        """Named parameters into /params: scalars, enums and vectors as
        attributes, arrays as datasets; units[name] as a ``unit:<name>``
        attribute (a ``unit`` attribute on a dataset)."""
        for k, v in kwargs.items():
            if _is_scalar(v) or isinstance(v, enum.Enum) or hasattr(v, "name") or hasattr(v, "x") or hasattr(v, "m"):
                _write_attr(self.g_params, k, v)
                if units and k in units:
                    self.g_params.attrs[f"unit:{k}"] = str(units[k])
                continue
            arr = _as_array(v)
            if arr is None:
                _write_attr(self.g_params, k, str(v))
                continue
            ds = _write_dataset(self.g_params, k, arr)
            if units and k in units:
                ds.attrs["unit"] = str(units[k])
        self.h5.flush()

    def save_params(self, config, *, extra: Optional[Dict[str, Any]] = None,
                    units: Optional[Dict[str, str]] = None) -> SimParams:
        """capture_params(config, extra) as JSON in /params ``params_json``, plus
        its most used scalars as flat /params attributes; returns the snapshot."""
        p = capture_params(config, extra)
        self.g_params.attrs["params_json"] = p.to_json()
        flat: Dict[str, Any] = {"n_photons": p.run.n_photons, "n_threads": p.run.n_threads, "seed": p.run.seed,
                                "max_events": p.run.max_events, "track_reverse_paths": p.run.track_reverse_paths,
                                "laser_source_type": p.laser.source_type, "wavelength": p.laser.wavelength,
                                "laser_sigma": p.laser.sigma, "sample_n_medium": p.sample_n_medium,
                                "n_layers": len(p.layers)}
        if p.layers:
            top = p.layers[0]
            flat["layer_kind"] = top.kind
            if top.kind == "mixture":
                flat.update(medium_kind="mixture", n_species=len(top.species or []), mu_s=top.mu_s_total,
                            mu_a=top.mu_a_total, mu_t=top.mu_t_total, mfp_total=top.mfp_total)
                if top.species:
                    flat["n_medium"] = top.species[0].n_medium
            else:
                m = top.medium
                flat.update(medium_kind=m.kind, mu_s=m.mu_s, mu_a=m.mu_a, mu_t=m.mu_t, n_particle=m.n_particle,
                            n_medium=m.n_medium)
                if m.radius is not None:
                    flat["radius"] = m.radius
                if m.mean_free_path is not None:
                    flat["mean_free_path"] = m.mean_free_path
        flat.update({k: v for k, v in p.extra.items() if _is_scalar(v)})
        self.log_params(units=units, **flat)
        return p

    def log_script(self, file_path: str, out_name: str = "script_snapshot.py") -> None:
        """Copy a script into the directory; its name and origin as root attributes."""
        file_path = str(file_path)
        if not os.path.exists(file_path):
            raise FileNotFoundError(f"Script file not found: {file_path}")
        shutil.copy(file_path, self.path / out_name)
        _write_attr(self.h5, "script_snapshot_file", out_name)
        _write_attr(self.h5, "script_original_path", file_path)
        self.h5.flush()

    def save_derived(self, key: str, arr, attrs: Optional[Dict[str, Any]] = None) -> None:
        """An array at /derived/<key> ("a/b/c": groups a, b, dataset c), with
        optional dataset attributes."""
        *groups, name = key.strip("/").split("/")
        g = self.g_derived
        for p in groups:
            g = g.require_group(p)
        ds = _write_dataset(g, name, np.asarray(arr))
        for ak, av in (attrs or {}).items():
            ds.attrs[ak] = _clean_scalar(av) if _is_scalar(av) else str(av)
        self.h5.flush()

    def save_absorption(self, absorption: Any, n_photons: int, name: str = "absorption") -> None:
        """An Absorption grid at /absorption/<name>: its geometry as meta
        attributes, each time slice and its image (per photon and voxel volume)
        as data/time_slice_<k> and data/time_slice_<k>_image."""
        g = self.h5.require_group("absorption").require_group(name)
        meta, data = g.require_group("meta"), g.require_group("data")
        for a in ("radius", "depth", "d_r", "d_z", "d_t", "t_max", "n_t"):
            _write_attr(meta, a, getattr(absorption, a))
        for k, ts in enumerate(absorption.time_slices):
            arr = _as_array(ts)
            if arr is not None:
                _write_dataset(data, f"time_slice_{k}", arr)
                _write_dataset(data, f"time_slice_{k}_image", _as_array(absorption.get_absorption_image(n_photons, k)))
        self.h5.flush()

    def save_sensors(self, sensors: Dict[str, Any]) -> None:
        for name, s in sensors.items():
            self.save_sensor(s, name)

    def save_sensor(self, sensor: Any, name: str) -> None:
        """A sensor at /sensors/<name> as its schema says (KeyError for a type
        the schema does not know)."""
        sch = schema.get_schema(type(sensor).__name__)
        g = self.g_sensors.require_group(name)
        meta, data = g.require_group("meta"), g.require_group("data")
        _write_attr(meta, "type", type(sensor).__name__)
        for a in schema.COMMON_META + sch.meta_attrs:
            if hasattr(sensor, a):
                _write_attr(meta, a, getattr(sensor, a))
        if sch.encoder is not None:
            sch.encoder(sensor, g, meta, data, _write_dataset, _write_attr)
        else:
            for a in sch.data_attrs:
                arr = _as_array(getattr(sensor, a))
                if arr is None or arr.size == 0:
                    _write_attr(meta, f"skipped:{a}", "empty (bins not set?)")
                    continue
                _write_dataset(data, a, arr)
        self.h5.flush()

    def save_processed(self, name: str, processed: Any, *, sensor: Any = None, theta: Optional[np.ndarray] = None,
                       phi: Optional[np.ndarray] = None) -> None:
        """A postprocess_* result under /derived/<name>: CBS as coherent/ and
        incoherent/ s0..s3 stacked over time bins (N_t, N_theta, N_phi) with
        axes/theta and axes/phi (from `sensor` unless given), planar fluence as
        s0..s3 (N_t, N_x, N_y), planar field as Ex, Ey."""
        cls = type(processed).__name__
        if cls == "FarFieldCBSProcessed":
            for comp, items in (("coherent", processed.coherent), ("incoherent", processed.incoherent)):
                for s in ("S0", "S1", "S2", "S3"):
                    self.save_derived(f"{name}/{comp}/{s.lower()}",
                                      np.stack([_to_numpy(getattr(it, s)) for it in items], axis=0))
            if theta is None and sensor is not None:
                theta = np.linspace(0.0, sensor.theta_max, sensor.N_theta)
            if phi is None and sensor is not None:
                phi = np.linspace(0.0, sensor.phi_max, sensor.N_phi)
            if theta is not None:
                self.save_derived(f"{name}/axes/theta", np.asarray(theta))
            if phi is not None:
                self.save_derived(f"{name}/axes/phi", np.asarray(phi))
        elif cls == "PlanarFluenceProcessed":
            for s in ("S0", "S1", "S2", "S3"):
                self.save_derived(f"{name}/{s.lower()}", np.stack([_to_numpy(m) for m in getattr(processed, s)]))
        elif cls == "PlanarFieldProcessed":
            self.save_derived(f"{name}/Ex", _to_numpy(processed.Ex))
            self.save_derived(f"{name}/Ey", _to_numpy(processed.Ey))
        else:
            raise TypeError(f"save_processed: unsupported processed type '{cls}'. "
                            "Supported: FarFieldCBSProcessed, PlanarFluenceProcessed, PlanarFieldProcessed.")


# ---------------------------------------------------------- ResultsLoader ----
class ResultsLoader:
    """Read access to an experiment directory's results file.

    `params_flat` holds /params' flat scalars and arrays at once; `params` is
    the typed SimParams from ``params_json``; sensors, derived arrays and
    Absorption grids are read when asked for.
    """

    def __init__(self, exp_dir: Union[str, Path], h5_name: str = "results.h5"):
        self.path = Path(exp_dir)
        self.h5_path = self.path / h5_name
        if not self.h5_path.exists():
            raise FileNotFoundError(f"HDF5 file not found: {self.h5_path}")
        self.h5 = h5.File(self.h5_path, "r")
        self.params_flat: Dict[str, Any] = {}
        self._params_json: Optional[str] = None
        self._params_cache: Optional[SimParams] = None
        g = self.h5.get("params")
        if g is not None:
            for k, v in g.attrs.items():
                if k == "params_json":
                    self._params_json = _py(v)
                else:
                    self.params_flat[k] = _py(v)
            for k in g.keys():
                if not g.is_group(k):
                    self.params_flat[k] = np.asarray(g[k])

    def close(self) -> None:
        self.h5.close()

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        self.close()

    def __getitem__(self, key: str) -> np.ndarray:
        """Any dataset by its path in the file, e.g. "sensors/cbs/data/S0_coh"."""
        return np.asarray(self.h5[key])

    def sensor_data(self, sensor_name: str, dataset: str) -> np.ndarray:
        return np.asarray(self.h5[f"sensors/{sensor_name}/data/{dataset}"])

    def sensor_meta(self, sensor_name: str) -> Dict[str, Any]:
        return dict(self.h5[f"sensors/{sensor_name}/meta"].attrs.items())

    def derived(self, key: str) -> np.ndarray:
        return np.asarray(self.h5[f"derived/{key}"])

    @property
    def params(self) -> SimParams:
        """The typed snapshot; KeyError when the file has no ``params_json``."""
        if self._params_cache is None:
            if self._params_json is None:
                raise KeyError("No typed params found (params_json missing). "
                               "Use `.params_flat` for the flat scalar dict.")
            self._params_cache = SimParams.from_json(self._params_json)
        return self._params_cache

    def sensor_names(self) -> List[str]:
        g = self.h5.get("sensors")
        return g.keys() if g is not None else []

    def _read_meta(self, name: str) -> Dict[str, Any]:
        return {k: _py(v) for k, v in self.h5[f"sensors/{name}/meta"].attrs.items()}

    def _read_group_arrays(self, path: str) -> Dict[str, np.ndarray]:
        g = self.h5.get(path)
        return {} if g is None else {k: np.asarray(g[k]) for k in g.keys()}

    def _read_data(self, name: str) -> Dict[str, np.ndarray]:
        return self._read_group_arrays(f"sensors/{name}/data")

    def _check_type(self, name: str, expected: str) -> Dict[str, Any]:
        meta = self._read_meta(name)
        if meta.get("type") != expected:
            raise TypeError(f"Sensor '{name}' has type '{meta.get('type')}', expected '{expected}'.")
        return meta

    def sensor(self, name: str) -> Any:
        """The typed result of a saved sensor, whatever its type."""
        meta = self._read_meta(name)
        t = meta.get("type")
        build = RESULT_BUILDERS.get(t)
        if build is None:
            raise TypeError(f"Sensor '{name}' has unsupported type '{t}'.")
        if t == "PhotonRecordSensor":
            return build(meta, self._read_group_arrays(f"sensors/{name}/records"))
        return build(meta, self._read_data(name))

    def far_field_cbs(self, name: str = "farfield_cbs") -> FarFieldCBSResult:
        return FarFieldCBSResult.from_h5(self._check_type(name, "FarFieldCBSSensor"), self._read_data(name))

    def planar_fluence(self, name: str = "planar_fluence") -> PlanarFluenceResult:
        return PlanarFluenceResult.from_h5(self._check_type(name, "PlanarFluenceSensor"), self._read_data(name))

    def planar_field(self, name: str = "planar_field") -> PlanarFieldResult:
        return PlanarFieldResult.from_h5(self._check_type(name, "PlanarFieldSensor"), self._read_data(name))

    def statistics(self, name: str = "statistics") -> StatisticsResult:
        return StatisticsResult.from_h5(self._check_type(name, "StatisticsSensor"), self._read_data(name))

    def photon_records(self, name: str = "photon_records") -> PhotonRecordsResult:
        meta = self._check_type(name, "PhotonRecordSensor")
        return PhotonRecordsResult.from_h5(meta, self._read_group_arrays(f"sensors/{name}/records"))

    def processed_cbs(self, name: str = "farfield_cbs") -> FarFieldCBSProcessedResult:
        base = f"derived/{name}"
        coh, inc = self._read_group_arrays(f"{base}/coherent"), self._read_group_arrays(f"{base}/incoherent")
        axes = self._read_group_arrays(f"{base}/axes")
        return FarFieldCBSProcessedResult(theta=axes.get("theta", np.empty(0)), phi=axes.get("phi", np.empty(0)),
                                          **{f"coh_s{q}": coh[f"s{q}"] for q in range(4)},
                                          **{f"inc_s{q}": inc[f"s{q}"] for q in range(4)})

    def processed_fluence(self, name: str) -> PlanarFluenceProcessedResult:
        d = self._read_group_arrays(f"derived/{name}")
        return PlanarFluenceProcessedResult(S0=d["s0"], S1=d["s1"], S2=d["s2"], S3=d["s3"])

    def processed_field(self, name: str) -> PlanarFieldProcessedResult:
        d = self._read_group_arrays(f"derived/{name}")
        return PlanarFieldProcessedResult(Ex=d["Ex"], Ey=d["Ey"])

    def absorption(self, name: str = "absorption") -> AbsorptionResult:
        """The Absorption grid: slices and images stacked over time bins (0 = integrated)."""
        meta = self.absorption_meta(name)
        g = self.h5.get(f"absorption/{name}/data")
        slices, images = [], []
        k = 0
        while g is not None and f"time_slice_{k}" in g:
            slices.append(np.asarray(g[f"time_slice_{k}"]))
            if f"time_slice_{k}_image" in g:
                images.append(np.asarray(g[f"time_slice_{k}_image"]))
            k += 1
        return AbsorptionResult.from_h5(meta, slices, images)

    def absorption_meta(self, name: str = "absorption") -> Dict[str, Any]:
        return dict(self.h5[f"absorption/{name}/meta"].attrs.items())

    def absorption_data(self, name: str = "absorption", time_index: int = 0) -> np.ndarray:
        return np.asarray(self.h5[f"absorption/{name}/data/time_slice_{time_index}"])

    def absorption_image(self, name: str = "absorption", time_index: int = 0) -> np.ndarray:
        return np.asarray(self.h5[f"absorption/{name}/data/time_slice_{time_index}_image"])

    def _first_of(self, paths: Sequence[str], what: str) -> np.ndarray:
        for p in paths:
            if p in self.h5:
                return np.asarray(self.h5[p])
        raise KeyError(f"No integrated absorption {what} found at {' or '.join(paths)}")

    def absorption_total_data(self, name: str = "absorption") -> np.ndarray:
        """Slice 0 (files that named it ``total`` are read too)."""
        base = f"absorption/{name}/data"
        return self._first_of([f"{base}/time_slice_0", f"{base}/total"], "dataset")

    def absorption_total_image(self, name: str = "absorption") -> np.ndarray:
        base = f"absorption/{name}/data"
        return self._first_of([f"{base}/time_slice_0_image", f"{base}/total_image"], "image")
