"""luminis_mc -- MI355X (gfx950) implementation of luminis-mc's photon-transport path.

Same public names as the reference package's transport surface
(python/luminis_mc/__init__.py:1-76): build a Sample, a Laser and a
SensorsGroup, set up a SimConfig, call run_simulation_parallel(config) and
post-process the sensors with postprocess_*.  The transport runs as HIP
kernels on the GPU through libluminis_hip.so.  Results persist to the
reference's HDF5 layout through Experiment / ResultsLoader (:mod:`manager`,
the HDF5 C library through ctypes).
"""
from ._lib import lib as _load_library
from .linalg import (CMatrix, CVec2, Matrix, Rng, Vec2, Vec3, calculate_rotation_angle, cross, dot, matcmul,
                     matcmulscalar, mix_seed, norm)
from .phase import (DrainePhaseFunction, HenyeyGreensteinPhaseFunction, MiePhaseFunction, PhaseFunction,
                    RayleighDebyeEMCPhaseFunction, RayleighDebyePhaseFunction, RayleighPhaseFunction,
                    SamplingTable, UniformPhaseFunction, form_factor, mie_amplitudes, mie_info)
from .medium import (Layer, MieMedium, MixtureLayer, RGDMedium, Sample, SampleLayer, ScatteringMedium,
                     lognormal_mie_mixture)
from .sensors import (CrossingDirection, FarFieldCBSSensor, FarFieldFluenceSensor, Laser, LaserSource,
                      PhotonRecord, PhotonRecordSensor, PlanarFieldSensor, PlanarFluenceSensor,
                      ReflectanceSensor, Sensor, SensorsGroup, StatisticsSensor,
                      TemporalProfile)
from .results import (FarFieldCBSProcessed, FarFieldCBSRadialProcessed, PlanarFieldProcessed,
                      PlanarFluenceProcessed, StokesMatrixProcessed, StokesRadialProcessed,
                      postprocess_farfield_cbs, postprocess_planar_field, postprocess_planar_fluence,
                      solid_angle_bin)
from .simulation import (GpuBatch, GpuScene, SimConfig, device_info, run_simulation_batched, run_simulation_parallel,
                         validate_sim_config)
from .absorption import Absorption
from .photon import LogLevel, Photon, gaussian_distribution, get_log_level, set_log_level, uniform_distribution
from .tuning import autotune
from .adjoint import DifferentiableTransport, crn_central_difference
from .manager import (Experiment, ResultsLoader, capture_params, derived_quantities, derived_quantities_mixture,
                      on_progress)
from .records import (AbsorptionResult, FarFieldCBSProcessedResult, FarFieldCBSResult, LaserParams, LayerParams,
                      MediumParams, PhotonRecordsResult, PlanarFieldProcessedResult, PlanarFieldResult,
                      PlanarFluenceProcessedResult, PlanarFluenceResult, RunParams, SensorMeta, SimParams,
                      StatisticsResult)
from .schema import SENSOR_SCHEMAS, SensorSchema

_load_library()

__all__ = [n for n in dir() if not n.startswith("_")]
