"""SimConfig and the GPU driver (replaces src/simulation.cpp:105-249).

``run_simulation_parallel(config)`` flattens the config's Sample / Laser /
SensorsGroup into the C ABI's scene descriptor, transports ``n_photons``
photons on the MI355X and adds the FP64 accumulators into the sensors'
buffers (the reference's clone/merge result semantics).  ``GpuScene`` keeps a
scene resident on the device for repeated runs (benchmarks, chunked runs).
"""
from __future__ import annotations

import ctypes
import math
import random

import numpy as np

from . import _lib
from .medium import MixtureLayer, Sample, SampleLayer
from .sensors import Laser, Sensor, SensorsGroup


class SimConfig:
    """Run controls (include/luminis/core/simulation.hpp:67-90) plus GPU launch knobs."""

    def __init__(self):
        self.MAX_EVENTS = 1000
        self.batch_size = 0
        self.show_progress = True
        self.progress_interval_pct = 5
        self.seed = random.getrandbits(32)
        self.n_threads = 1
        self.n_photons = 0
        self.pin_threads_to_cores = False
        self.track_reverse_paths = False
        self.sample: Sample | None = None
        self.laser: Laser | None = None
        self.detector: SensorsGroup | None = None
        self.absorption = None
        # GPU knobs (no counterpart in the reference)
        self.precision = "fp32"       # "fp32" (FP32 transport, FP64 accumulation) or "fp64"
        self.rng = "philox"           # "philox" or "xoshiro" (fp64 only: the reference's streams)
        self.photon_offset = 0
        self.block_size = 0
        self.blocks_per_cu = 0
        self.claim = 0
        self.last_stats: dict | None = None


def validate_sim_config(config: SimConfig) -> None:
    """simulation.cpp:95-103."""
    if config.sample is None:
        raise ValueError("SimConfig::sample must not be null")
    if config.laser is None:
        raise ValueError("SimConfig::laser must not be null")
    if config.detector is None:
        raise ValueError("SimConfig::detector must not be null")


class SceneBuild:
    """Flat descriptor of one scene plus the arrays it points into."""

    def __init__(self, sample: Sample, laser: Laser, sensors: list[Sensor], absorption=None):
        if not sample.layers:
            raise ValueError("Sample has no layers")
        tables: list[np.ndarray] = []
        n_tab = 0

        def put(arr) -> int:
            nonlocal n_tab
            a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
            off = n_tab
            tables.append(a)
            n_tab += a.size
            return off

        species_objs = []
        species_index = {}

        def species_of(medium) -> int:
            key = id(medium)
            if key not in species_index:
                species_index[key] = len(species_objs)
                species_objs.append(medium)
            return species_index[key]

        layer_specs = []
        for L in sample.layers:
            if isinstance(L, SampleLayer):
                idx = [species_of(L.medium)]
                layer_specs.append((L, idx, None))
            elif isinstance(L, MixtureLayer):
                # a mixture's species occupy one consecutive block
                first = len(species_objs)
                for m in L.species:
                    species_objs.append(m)
                idx = list(range(first, first + len(L.species)))
                layer_specs.append((L, idx, np.asarray(L.selection_cdf, dtype=np.float64)))
            else:
                raise TypeError(f"unsupported layer type {type(L).__name__}")

        sp_arr = (_lib.SpeciesDesc * len(species_objs))()
        for i, med in enumerate(species_objs):
            d = sp_arr[i]
            ph = med.phase_function
            d.phase_kind = ph.kind
            if ph.kind in (_lib.PHASE_TABLE_THETA, _lib.PHASE_TABLE_COS):
                t = ph.table
                d.cdf_n = t.cdf.size
                d.cdf_offset = put(t.cdf)
                d.cdf_min = t.min_val
                d.cdf_step = t.step
            if ph.kind == _lib.PHASE_HG:
                if ph.g == 0.0:
                    raise ValueError("Henyey-Greenstein sampling divides by 2g: g must be non-zero")
                d.hg_g = ph.g
            s1, s2, dth = med.s_tables()
            d.s_n = s1.size
            d.s1_offset = put(np.stack([s1.real, s1.imag], axis=1))
            d.s2_offset = put(np.stack([s2.real, s2.imag], axis=1))
            d.s_dtheta = dth
            d.mu_s, d.mu_a = med.mu_s, med.mu_a
            d.mfp = med.mean_free_path
            try:
                d.sigma_s = med.scattering_cross_section()
            except NotImplementedError:
                d.sigma_s = 0.0

        ly_arr = (_lib.LayerDesc * len(layer_specs))()
        for i, (L, idx, sel) in enumerate(layer_specs):
            d = ly_arr[i]
            d.z_min, d.z_max = L.z_min, L.z_max
            d.mu_s, d.mu_a, d.mu_t = L.mu_scattering(), L.mu_absorption(), L.mu_attenuation()
            d.mfp = L.free_path_mean
            if not (d.mfp > 0.0) or math.isinf(d.mfp):
                raise ValueError(f"layer {i}: mean free path is not set (call set_mean_free_path on its medium)")
            d.n = sample.refractive_index
            d.species_first = idx[0]
            d.species_count = len(idx)
            d.select_offset = put(sel) if sel is not None else 0

        se_arr = (_lib.SensorDesc * max(1, len(sensors)))()
        n_out = 0
        self.offsets = []
        for i, s in enumerate(sensors):
            d = se_arr[i]
            s._fill_desc(d)
            if s.kind == _lib.SENSOR_STATISTICS:  # histogram parameter block
                d.stat_offset = put(s.stat_params())
            if getattr(s, "phi_explicit", False):  # set_phi_slices columns
                d.n_phi_explicit = len(s.phi_values)
                d.phi_offset = put(np.asarray(s.phi_values, dtype=np.float64))
            d.out_offset = n_out
            d.out_len = s.out_len()
            self.offsets.append((n_out, s.out_len()))
            n_out += s.out_len()

        self.absorption = absorption
        self.abs_offset = n_out
        if absorption is not None:
            n_out += absorption.buffer.size
        self.tables = np.concatenate(tables) if tables else np.zeros(1)
        self.species = sp_arr
        self.layers = ly_arr
        self.sensor_descs = se_arr
        self.sensors = list(sensors)
        self.n_out = n_out

        lz = laser.position.z
        if sample.get_layer_index_at(lz) >= sample.size():
            raise ValueError("laser position lies outside the sample")
        desc = _lib.SceneDesc()
        desc.n_layers = len(layer_specs)
        desc.n_species = len(species_objs)
        desc.n_sensors = len(sensors)
        desc.layers = ctypes.cast(ly_arr, ctypes.POINTER(_lib.LayerDesc))
        desc.species = ctypes.cast(sp_arr, ctypes.POINTER(_lib.SpeciesDesc))
        desc.sensors = ctypes.cast(se_arr, ctypes.POINTER(_lib.SensorDesc))
        desc.tables = _lib.dptr(self.tables)
        desc.n_tables = self.tables.size
        desc.n_out = n_out
        ld = desc.laser
        for j, v in enumerate(laser.position.as_tuple()):
            ld.pos[j] = v
        for j, v in enumerate(laser.direction.as_tuple()):
            ld.dir[j] = v
        for j, v in enumerate(laser.local_m.as_tuple()):
            ld.m[j] = v
        for j, v in enumerate(laser.local_n.as_tuple()):
            ld.n[j] = v
        pol = laser.polarization
        ld.pol[0], ld.pol[1], ld.pol[2], ld.pol[3] = pol.m.real, pol.m.imag, pol.n.real, pol.n.imag
        ld.wavelength = laser.wavelength
        ld.sigma = laser.sigma
        ld.source = int(laser.source_type)
        ld.temporal = int(laser.temporal_profile)
        ld.pulse_duration = laser.pulse_duration
        ld.repetition_rate = laser.repetition_rate
        ld.time_offset = laser.time_offset
        desc.n_medium = sample.refractive_index
        if absorption is not None:
            absorption._fill_desc(desc.absorption, self.abs_offset)
        self.desc = desc

    def scatter_results(self, out: np.ndarray, hits: np.ndarray, records: np.ndarray | None = None) -> None:
        """Add a flat accumulator array into the sensors (merge_from), and
        PhotonRecordSensor records (rows of lmc_scene_records, column 0 the
        sensor's index) into their sensors."""
        for s, (off, n), h in zip(self.sensors, self.offsets, hits):
            s.buffer += out[off:off + n]
            s.hits += int(h)
        if records is not None and len(records):
            idx = records[:, 0].astype(np.int64)
            for i, s in enumerate(self.sensors):
                if s.kind == _lib.SENSOR_PHOTON_RECORD:
                    s.add_records(records[idx == i])
        if self.absorption is not None:
            self.absorption.buffer += out[self.abs_offset:self.abs_offset + self.absorption.buffer.size]


def run_desc_from(config: SimConfig, n_photons: int | None = None, photon_offset: int | None = None):
    r = _lib.RunDesc()
    r.n_photons = int(config.n_photons if n_photons is None else n_photons)
    r.photon_offset = int(config.photon_offset if photon_offset is None else photon_offset)
    r.seed = int(config.seed) & ((1 << 64) - 1)
    r.max_events = int(config.MAX_EVENTS)
    r.track_reverse_paths = int(bool(config.track_reverse_paths))
    prec = str(config.precision).lower()
    if prec not in ("fp32", "fp64"):
        raise ValueError("precision must be 'fp32' or 'fp64'")
    r.precision = _lib.PREC_FP32 if prec == "fp32" else _lib.PREC_FP64
    rng = str(config.rng).lower()
    if rng not in ("philox", "xoshiro"):
        raise ValueError("rng must be 'philox' or 'xoshiro'")
    r.rng = _lib.RNG_PHILOX if rng == "philox" else _lib.RNG_XOSHIRO
    r.block_size = int(config.block_size)
    r.blocks_per_cu = int(config.blocks_per_cu)
    r.claim = int(config.claim)
    return r


class GpuScene:
    """A scene resident on the device: create once, run many times."""

    def __init__(self, sample: Sample, laser: Laser, sensors, build: SceneBuild | None = None, absorption=None):
        self.build = build if build is not None else SceneBuild(sample, laser, list(sensors), absorption)
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().lmc_scene_create(ctypes.byref(self.build.desc), ctypes.byref(self._h)),
                   "lmc_scene_create")

    @classmethod
    def from_build(cls, build) -> "GpuScene":
        """A device scene from an already flattened descriptor (SceneBuild or alike)."""
        return cls(None, None, None, build=build)

    def run(self, run: _lib.RunDesc) -> dict:
        st = _lib.Stats()
        _lib.check(_lib.lib().lmc_scene_run(self._h, ctypes.byref(run), ctypes.byref(st)), "lmc_scene_run")
        return st.as_dict()

    def run_adjoint(self, run: _lib.RunDesc) -> tuple[np.ndarray, dict]:
        n_layers = self.build.desc.n_layers
        grad = np.zeros(3 * n_layers * self.build.n_out)
        st = _lib.Stats()
        _lib.check(_lib.lib().lmc_scene_run_adjoint(self._h, ctypes.byref(run), _lib.dptr(grad), grad.size,
                                                    ctypes.byref(st)), "lmc_scene_run_adjoint")
        return grad.reshape(3 * n_layers, self.build.n_out), st.as_dict()

    def read(self) -> tuple[np.ndarray, np.ndarray]:
        out = np.zeros(self.build.n_out)
        hits = np.zeros(max(1, len(self.build.sensors)), dtype=np.uint64)
        _lib.check(_lib.lib().lmc_scene_read(self._h, _lib.dptr(out), out.size,
                                             hits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))),
                   "lmc_scene_read")
        return out, hits

    def records(self) -> np.ndarray:
        """PhotonRecordSensor records kept since creation or zero(): an
        (n, RECORD_FIELDS) array, column 0 the sensor's index."""
        n = ctypes.c_int64(0)
        L = _lib.lib()
        _lib.check(L.lmc_scene_records(self._h, None, 0, ctypes.byref(n)), "lmc_scene_records")
        rows = np.zeros((n.value, _lib.RECORD_FIELDS))
        if n.value:
            _lib.check(L.lmc_scene_records(self._h, _lib.dptr(rows), n.value, ctypes.byref(n)), "lmc_scene_records")
        return rows

    def zero(self) -> None:
        _lib.check(_lib.lib().lmc_scene_zero(self._h), "lmc_scene_zero")

    def merge_into_sensors(self) -> None:
        out, hits = self.read()
        has_rec = any(s.kind == _lib.SENSOR_PHOTON_RECORD for s in self.build.sensors)
        self.build.scatter_results(out, hits, self.records() if has_rec else None)

    def close(self) -> None:
        if self._h:
            _lib.lib().lmc_scene_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBatch:
    """Many independent scenes resident on the device, run together in one
    persistent launch (lmc_batch_*).  Scene i's photon j follows the same
    random stream as photon j of a GpuScene run of that scene alone."""

    def __init__(self, scenes, builds=None):
        """scenes: iterable of (sample, laser, sensors) or (sample, laser,
        sensors, absorption) tuples; absorption is an Absorption grid or None."""
        self.builds = list(builds) if builds is not None else []
        for t in (scenes if builds is None else ()):
            if len(t) not in (3, 4):
                raise ValueError("a batch scene is (sample, laser, sensors[, absorption])")
            sa, la, se = t[:3]
            self.builds.append(SceneBuild(sa, la, list(se), t[3] if len(t) == 4 else None))
        if not self.builds:
            raise ValueError("a batch needs at least one scene")
        arr = (_lib.SceneDesc * len(self.builds))(*[b.desc for b in self.builds])
        self._descs = arr
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().lmc_batch_create(len(self.builds), arr, ctypes.byref(self._h)), "lmc_batch_create")

    @classmethod
    def from_builds(cls, builds) -> "GpuBatch":
        """A device batch from already flattened descriptors (SceneBuild or alike)."""
        return cls(None, builds=builds)

    def __len__(self):
        return len(self.builds)

    def run(self, budgets, run: _lib.RunDesc, use_graph: bool = False) -> dict:
        b = np.ascontiguousarray(budgets, dtype=np.uint64)
        if b.shape != (len(self.builds),):
            raise ValueError("one photon budget per scene")
        st = _lib.Stats()
        _lib.check(_lib.lib().lmc_batch_run(self._h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                            ctypes.byref(run), int(bool(use_graph)), ctypes.byref(st)),
                   "lmc_batch_run")
        return st.as_dict()

    def read(self, k: int) -> tuple[np.ndarray, np.ndarray]:
        bld = self.builds[k]
        out = np.zeros(bld.n_out)
        hits = np.zeros(max(1, len(bld.sensors)), dtype=np.uint64)
        _lib.check(_lib.lib().lmc_batch_read(self._h, int(k), _lib.dptr(out), out.size,
                                             hits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))),
                   "lmc_batch_read")
        return out, hits

    def records(self, k: int) -> np.ndarray:
        """Scene k's PhotonRecordSensor records: an (n, RECORD_FIELDS) array,
        column 0 the sensor's index in scene k."""
        n = ctypes.c_int64(0)
        L = _lib.lib()
        _lib.check(L.lmc_batch_records(self._h, int(k), None, 0, ctypes.byref(n)), "lmc_batch_records")
        rows = np.zeros((n.value, _lib.RECORD_FIELDS))
        if n.value:
            _lib.check(L.lmc_batch_records(self._h, int(k), _lib.dptr(rows), n.value, ctypes.byref(n)),
                       "lmc_batch_records")
        return rows

    def zero(self) -> None:
        _lib.check(_lib.lib().lmc_batch_zero(self._h), "lmc_batch_zero")

    def merge_into_sensors(self) -> None:
        for k, bld in enumerate(self.builds):
            out, hits = self.read(k)
            has_rec = any(s.kind == _lib.SENSOR_PHOTON_RECORD for s in bld.sensors)
            bld.scatter_results(out, hits, self.records(k) if has_rec else None)

    def close(self) -> None:
        if self._h:
            _lib.lib().lmc_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_simulation_batched(configs, use_graph: bool = True) -> dict:
    """Run several SimConfigs in one batched launch and add each one's results
    into its own sensors, as run_simulation_parallel would for each.  The
    configs share one random key and run controls (seed, MAX_EVENTS,
    track_reverse_paths, precision, rng, photon_offset), which must agree."""
    configs = list(configs)
    if not configs:
        raise ValueError("no configs")
    for c in configs:
        validate_sim_config(c)
    keys = ("seed", "MAX_EVENTS", "track_reverse_paths", "precision", "rng", "photon_offset",
            "block_size", "blocks_per_cu", "claim")
    c0 = configs[0]
    for c in configs[1:]:
        for k in keys:
            if getattr(c, k) != getattr(c0, k):
                raise ValueError(f"batched configs must share {k}")
    batch = GpuBatch([(c.sample, c.laser, c.detector.detectors, c.absorption) for c in configs])
    try:
        st = batch.run([int(c.n_photons) for c in configs], run_desc_from(c0, n_photons=0), use_graph)
        batch.merge_into_sensors()
    finally:
        batch.close()
    for c in configs:
        c.last_stats = st
    return st


def run_simulation_parallel(config: SimConfig) -> None:
    """Transport config.n_photons photons on the GPU and add the results into
    config.detector's sensors (python_module.cpp:664-673 -> simulation.cpp:105-249)."""
    validate_sim_config(config)
    scene = GpuScene(config.sample, config.laser, config.detector.detectors, absorption=config.absorption)
    try:
        config.last_stats = scene.run(run_desc_from(config))
        scene.merge_into_sensors()
    finally:
        scene.close()


def device_info() -> dict:
    info = _lib.DeviceInfo()
    _lib.check(_lib.lib().lmc_get_device_info(ctypes.byref(info)), "lmc_get_device_info")
    d = {k: getattr(info, k) for k, _ in info._fields_}
    d["name"] = info.name.decode()
    return d
