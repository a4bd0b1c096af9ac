"""How each sensor type is laid out in a results file (python/luminis_mc/schema.py:1-145).

A sensor saved under the name N lands in ``/sensors/N/`` as ``meta/`` (HDF5
attributes: the type name, the base-sensor attributes of COMMON_META and the
type's `meta_attrs`) and ``data/`` (one dataset per name in `data_attrs`; an
empty array is not written and is flagged by a ``skipped:<attr>`` attribute).
The writer (`manager.Experiment.save_sensor`) and the reader
(`manager.ResultsLoader`) both go by this table, and its names are the ones
the reference's files use, so files written here and there read the same.
PhotonRecordSensor is written by its own encoder (`records.encode_photon_records`)
into ``/sensors/N/records/``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional, Tuple

# base-sensor attributes every sensor's meta/ carries (detector.hpp Sensor fields)
COMMON_META: Tuple[str, ...] = (
    "id", "origin", "normal", "backward_normal", "n_polarization", "m_polarization", "hits",
    "absorb_photons", "estimator_enabled",
    "filter_theta_enabled", "filter_theta_min", "filter_theta_max",
    "filter_phi_enabled", "filter_phi_min", "filter_phi_max",
    "filter_position_enabled", "filter_x_min", "filter_x_max", "filter_y_min", "filter_y_max",
    "filter_direction_enabled", "filter_direction",
)


@dataclass(frozen=True)
class SensorSchema:
    """One sensor type: its class name, the scalar attributes for meta/, the
    array attributes for data/, and an optional encoder that writes the sensor
    itself, called as encoder(sensor, group, meta, data, write_dataset, write_attr)."""

    type_name: str
    meta_attrs: Tuple[str, ...] = ()
    data_attrs: Tuple[str, ...] = ()
    encoder: Optional[Callable] = field(default=None, compare=False)


SENSOR_SCHEMAS = {
    "FarFieldCBSSensor": SensorSchema(
        "FarFieldCBSSensor",
        meta_attrs=("N_theta", "N_phi", "N_t", "t_max", "theta_max", "phi_max", "dtheta", "dphi", "theta_pp_max",
                    "phi_values", "phi_explicit"),
        data_attrs=("S0_coh", "S1_coh", "S2_coh", "S3_coh", "S0_incoh", "S1_incoh", "S2_incoh", "S3_incoh"),
    ),
    "PlanarFluenceSensor": SensorSchema(
        "PlanarFluenceSensor",
        meta_attrs=("N_x", "N_y", "N_t", "dx", "dy", "dt", "len_x", "len_y", "len_t"),
        data_attrs=("S0", "S1", "S2", "S3"),
    ),
    "PlanarFieldSensor": SensorSchema(
        "PlanarFieldSensor",
        meta_attrs=("N_x", "N_y", "dx", "dy", "len_x", "len_y"),
        data_attrs=("Ex", "Ey"),
    ),
    "StatisticsSensor": SensorSchema(
        "StatisticsSensor",
        meta_attrs=("events_histogram_bins_set", "max_events",
                    "theta_histogram_bins_set", "min_theta", "max_theta", "n_bins_theta", "dtheta",
                    "phi_histogram_bins_set", "min_phi", "max_phi", "n_bins_phi", "dphi",
                    "depth_histogram_bins_set", "max_depth", "n_bins_depth", "ddepth",
                    "N_t", "dt", "t_max",
                    "time_histogram_bins_set", "h_max_time", "n_bins_time", "h_dtime",
                    "weight_histogram_bins_set", "max_weight", "n_bins_weight", "dweight"),
        data_attrs=("events_histogram", "theta_histogram", "phi_histogram", "depth_histogram", "time_histogram",
                    "weight_histogram"),
    ),
    # records written by records.encode_photon_records (registered there)
    "PhotonRecordSensor": SensorSchema("PhotonRecordSensor"),
}


def get_schema(type_name: str) -> SensorSchema:
    """The schema of a sensor class name; KeyError naming the known types otherwise."""
    if type_name not in SENSOR_SCHEMAS:
        raise KeyError(f"Unsupported sensor type: '{type_name}'. Known types: {sorted(SENSOR_SCHEMAS)}")
    return SENSOR_SCHEMAS[type_name]
