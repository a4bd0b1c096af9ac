"""ctypes binding of libluminis_hip.so (the C ABI declared in include/luminis_hip.h).

The structures below mirror the header field by field.  The library is loaded
from this package's ``_lib/`` directory only; if it is missing the import fails
with an explicit error (there is no CPU fallback for the transport path).
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libluminis_hip.so")

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_uint64 = ctypes.c_uint64
c_double = ctypes.c_double
c_dp = ctypes.POINTER(ctypes.c_double)

LMC_OK = 0
LMC_ERR_INVALID = 1
LMC_ERR_DEVICE = 2
LMC_ERR_NOMEM = 3
LMC_ERR_UNSUPPORTED = 4

PHASE_TABLE_THETA = 0
PHASE_TABLE_COS = 1
PHASE_HG = 2
PHASE_UNIFORM = 3

SENSOR_FARFIELD_CBS = 1
SENSOR_PLANAR_FLUENCE = 2
SENSOR_PLANAR_FIELD = 3
SENSOR_FARFIELD_FLUENCE = 4
SENSOR_REFLECTANCE = 5
SENSOR_STATISTICS = 6
SENSOR_PHOTON_RECORD = 7
RECORD_FIELDS = 34  # doubles per PhotonRecordSensor record (LMC_RECORD_FIELDS)
MAX_LAYERS = 8    # per scene (lmc_scene_create refuses more)
MAX_SPECIES = 64  # per scene
MAX_SENSORS = 8   # per scene

PREC_FP32 = 0
PREC_FP64 = 1
RNG_PHILOX = 0
RNG_XOSHIRO = 1


class SpeciesDesc(ctypes.Structure):
    _fields_ = [
        ("phase_kind", c_int32), ("cdf_n", c_int32), ("cdf_offset", c_int64),
        ("cdf_min", c_double), ("cdf_step", c_double), ("hg_g", c_double),
        ("s_n", c_int32), ("reserved0", c_int32),
        ("s1_offset", c_int64), ("s2_offset", c_int64), ("s_dtheta", c_double),
        ("mu_s", c_double), ("mu_a", c_double), ("mfp", c_double), ("sigma_s", c_double),
    ]


class LayerDesc(ctypes.Structure):
    _fields_ = [
        ("z_min", c_double), ("z_max", c_double), ("mu_s", c_double), ("mu_a", c_double),
        ("mu_t", c_double), ("mfp", c_double), ("n", c_double),
        ("species_first", c_int32), ("species_count", c_int32), ("select_offset", c_int64),
    ]


class SensorDesc(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32), ("absorb", c_int32), ("z", c_double),
        ("f_dir_en", c_int32), ("f_dir", c_int32),
        ("f_ev_en", c_int32), ("f_ev_min", c_int32), ("f_ev_max", c_int32),
        ("f_th_en", c_int32), ("cos_th_min", c_double), ("cos_th_max", c_double),
        ("f_ph_en", c_int32), ("reserved0", c_int32), ("ph_min", c_double), ("ph_max", c_double),
        ("f_pos_en", c_int32), ("reserved1", c_int32),
        ("x_min", c_double), ("x_max", c_double), ("y_min", c_double), ("y_max", c_double),
        ("g", c_double * 8), ("n", c_int32 * 4),
        ("out_offset", c_int64), ("out_len", c_int64),
        ("estimator", c_int32), ("n_phi_explicit", c_int32), ("phi_offset", c_int64), ("theta_pp_max", c_double),
        ("stat_offset", c_int64),
    ]


class LaserDesc(ctypes.Structure):
    _fields_ = [
        ("pos", c_double * 3), ("dir", c_double * 3), ("m", c_double * 3), ("n", c_double * 3),
        ("pol", c_double * 4), ("wavelength", c_double), ("sigma", c_double),
        ("source", c_int32), ("temporal", c_int32),
        ("pulse_duration", c_double), ("repetition_rate", c_double), ("time_offset", c_double),
    ]


class AbsorptionDesc(ctypes.Structure):
    _fields_ = [
        ("enabled", c_int32), ("n_r", c_int32), ("n_z", c_int32), ("n_t", c_int32),
        ("radius", c_double), ("depth", c_double), ("d_r", c_double), ("d_z", c_double),
        ("d_t", c_double), ("t_max", c_double), ("out_offset", c_int64),
    ]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_layers", c_int32), ("n_species", c_int32), ("n_sensors", c_int32), ("reserved0", c_int32),
        ("layers", ctypes.POINTER(LayerDesc)), ("species", ctypes.POINTER(SpeciesDesc)),
        ("sensors", ctypes.POINTER(SensorDesc)), ("tables", c_dp),
        ("n_tables", c_int64), ("n_out", c_int64),
        ("laser", LaserDesc), ("n_medium", c_double),
        ("absorption", AbsorptionDesc),
    ]


class RunDesc(ctypes.Structure):
    _fields_ = [
        ("n_photons", c_uint64), ("photon_offset", c_uint64), ("seed", c_uint64),
        ("max_events", c_int32), ("track_reverse_paths", c_int32),
        ("precision", c_int32), ("rng", c_int32),
        ("block_size", c_int32), ("blocks_per_cu", c_int32), ("claim", c_int32), ("reserved0", c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("photons", c_uint64), ("events", c_uint64), ("phi_trials", c_uint64),
        ("crossings", c_uint64), ("cbs_detections", c_uint64),
        ("kernel_ms", c_double), ("transport_ms", c_double), ("sensor_ms", c_double),
        ("transport_launches", c_uint64),
        ("block_size", c_int32), ("blocks_per_cu", c_int32), ("lds_tables", c_int32), ("aux_mode", c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("n_cu", c_int32), ("clock_khz", c_int32), ("lds_per_block", c_int64),
        ("lds_per_cu", c_int64), ("global_mem", c_int64), ("name", ctypes.c_char * 64),
    ]


_lib = None
HIP_RUNTIME = None  # the HIP runtime library loaded ahead of libluminis_hip.so, if any


def _preload_hip_runtime() -> None:
    """One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64
    (and HSA runtime) and its libraries ask for it by a name that does not
    match the system runtime's soname, so a process that loads this library
    (bound to the system runtime) and torch ends up with two runtimes on one
    device, and torch then finds no GPU.  When torch is installed, its runtime
    is loaded here first, without importing torch: the library's dependency on
    libamdhip64.so.7 then binds to it (same soname), and a later `import
    torch` reuses it.  LMC_HIP_RUNTIME=system keeps the system runtime."""
    global HIP_RUNTIME
    if os.environ.get("LMC_HIP_RUNTIME", "") == "system":
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        HIP_RUNTIME = path


def lib():
    """Load libluminis_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libluminis_hip.so not found at {LIB_PATH}; build it with "
            "`python luminis-mc_amd/build.py` (the transport path has no CPU fallback)")
    _preload_hip_runtime()
    L = ctypes.CDLL(LIB_PATH)
    L.lmc_abi_version.restype = ctypes.c_int
    L.lmc_last_error.restype = ctypes.c_char_p
    L.lmc_device_count.restype = ctypes.c_int
    L.lmc_get_device_info.argtypes = [ctypes.POINTER(DeviceInfo)]
    L.lmc_scene_create.argtypes = [ctypes.POINTER(SceneDesc), ctypes.POINTER(ctypes.c_void_p)]
    L.lmc_scene_destroy.argtypes = [ctypes.c_void_p]
    L.lmc_scene_destroy.restype = None
    L.lmc_scene_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(RunDesc), ctypes.POINTER(Stats)]
    L.lmc_scene_run_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(RunDesc), ctypes.POINTER(Stats)]
    L.lmc_scene_read.argtypes = [ctypes.c_void_p, c_dp, c_int64, ctypes.POINTER(c_uint64)]
    L.lmc_scene_zero.argtypes = [ctypes.c_void_p]
    L.lmc_scene_records.argtypes = [ctypes.c_void_p, c_dp, c_int64, ctypes.POINTER(c_int64)]
    L.lmc_scene_run_adjoint.argtypes = [ctypes.c_void_p, ctypes.POINTER(RunDesc), c_dp, c_int64,
                                        ctypes.POINTER(Stats)]
    L.lmc_batch_create.argtypes = [c_int32, ctypes.POINTER(SceneDesc), ctypes.POINTER(ctypes.c_void_p)]
    L.lmc_batch_destroy.argtypes = [ctypes.c_void_p]
    L.lmc_batch_destroy.restype = None
    L.lmc_batch_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(c_uint64), ctypes.POINTER(RunDesc), c_int32,
                                ctypes.POINTER(Stats)]
    L.lmc_batch_read.argtypes = [ctypes.c_void_p, c_int32, c_dp, c_int64, ctypes.POINTER(c_uint64)]
    L.lmc_batch_records.argtypes = [ctypes.c_void_p, c_int32, c_dp, c_int64, ctypes.POINTER(c_int64)]
    L.lmc_batch_zero.argtypes = [ctypes.c_void_p]
    L.lmc_mie_amplitudes.argtypes = [c_double, c_double, c_double, c_int32, c_dp, c_dp, c_dp]
    L.lmc_mie_info.argtypes = [c_double, c_double, c_double, c_dp, c_dp, c_dp]
    L.lmc_sampling_cdf.argtypes = [c_int32, c_double, c_double, c_dp, c_dp]
    L.lmc_probe_azimuth.argtypes = [c_int32, c_int32, c_dp, ctypes.POINTER(ctypes.c_uint32), c_dp]
    L.lmc_probe_philox.argtypes = [c_int32, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    L.lmc_probe_stream.argtypes = [ctypes.c_uint32, ctypes.c_uint32, c_uint64, c_int32,
                                   ctypes.POINTER(ctypes.c_uint32)]
    L.lmc_probe_phase.argtypes = [c_int32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.lmc_probe_reverse.argtypes = [c_int32, c_int32, c_dp, c_dp]
    if L.lmc_abi_version() != 2:
        raise ImportError("libluminis_hip.so ABI version mismatch")
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    """Map a status code to the reference's Python exceptions."""
    if rc == LMC_OK:
        return
    msg = lib().lmc_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == LMC_ERR_INVALID:
        raise ValueError(text)
    if rc == LMC_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise RuntimeError(text)


def dptr(a):
    return a.ctypes.data_as(c_dp)
