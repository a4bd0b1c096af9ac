"""HDF5 persistence (python/luminis_mc/manager.py:447-1288, schema.py, records.py):
Experiment writes, ResultsLoader reads, and the files hold the reference's
layout in h5py's encoding of each type.  The layout and the encodings are
checked with the HDF5 distribution's own h5dump / h5ls, a reader independent
of the ctypes binding that wrote the files (skipped where those tools are
absent); values are checked by reading back and against the sensors' arrays."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import luminis_mc as lmc
from luminis_mc import h5, schema

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
H5LS = shutil.which("h5ls") or ("/opt/conda/bin/h5ls" if os.path.exists("/opt/conda/bin/h5ls") else None)


def _have_hdf5() -> bool:
    try:
        h5.library()
        return True
    except RuntimeError:
        return False


pytestmark = pytest.mark.skipif(not _have_hdf5(), reason="no HDF5 C library on this host")


def _dump(path, *args) -> str:
    return subprocess.run([H5DUMP, *args, str(path)], check=True, capture_output=True, text=True).stdout


def _listing(path) -> dict:
    """{object path: "Group" | "Dataset"} from h5ls -r."""
    out = subprocess.run([H5LS, "-r", str(path)], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        name, kind = line.split()[0], line.split()[1]
        res[name] = kind
    return res


def test_binding_types_round_trip_and_match_h5py_encodings(tmp_path):
    p = tmp_path / "t.h5"
    rng = np.random.default_rng(1)
    arrays = {"f8": rng.normal(size=(3, 5, 40)), "i4": np.arange(7, dtype=np.int32),
              "i8": np.arange(300, dtype=np.int64), "c16": rng.normal(size=(300, 2)) + 1j * rng.normal(size=(300, 2)),
              "b": np.array([True, False, True])}
    with h5.File(p, "w") as f:
        g = f.require_group("a").require_group("b")
        for k, v in arrays.items():
            g.create_dataset(k, v, compression="gzip" if v.size > 256 else None, compression_opts=4,
                             shuffle=v.size > 256)
        g.attrs["n"] = 100
        g.attrs["x"] = 0.25
        g.attrs["flag"] = np.bool_(True)
        g.attrs["s"] = "FarFieldCBSSensor"
        g.attrs["v3"] = np.array([0.0, 0.0, 1.5])
        g.attrs["pol"] = np.array([1 + 2j, 3 - 4j])
        f.create_dataset("scalar", 3.25)
    with h5.File(p, "r") as f:
        g = f["a/b"]
        for k, v in arrays.items():
            r = np.asarray(g[k])
            assert r.dtype == v.dtype and r.shape == v.shape and np.array_equal(r, v), k
        assert sorted(g.keys()) == sorted(arrays)
        a = dict(g.attrs.items())
        assert a["n"] == 100 and isinstance(a["n"], np.integer)
        assert a["x"] == 0.25 and a["flag"] is np.True_ and a["s"] == "FarFieldCBSSensor"
        assert np.array_equal(a["v3"], [0, 0, 1.5]) and np.array_equal(a["pol"], [1 + 2j, 3 - 4j])
        assert f["scalar"][()] == 3.25
        assert "a/b/f8" in f and "a/c" not in f and "a/b/f8/x" not in f
    if H5DUMP is None:
        pytest.skip("h5dump not installed: encodings not cross-checked")
    d = " ".join(_dump(p, "-p", "-A").split())
    # complex128 as h5py's compound {r, i}; bool as its int8 enum; str as vlen UTF-8; big arrays filtered
    assert 'DATASET "c16" { DATATYPE H5T_COMPOUND { H5T_IEEE_F64LE "r"; H5T_IEEE_F64LE "i"; }' in d
    assert 'DATASET "b" { DATATYPE H5T_ENUM { H5T_STD_I8LE; "FALSE" 0; "TRUE" 1; }' in d
    assert 'ATTRIBUTE "s" { DATATYPE H5T_STRING { STRSIZE H5T_VARIABLE; STRPAD H5T_STR_NULLTERM; ' \
           'CSET H5T_CSET_UTF8; CTYPE H5T_C_S1; } DATASPACE SCALAR' in d
    assert 'ATTRIBUTE "n" { DATATYPE H5T_STD_I64LE DATASPACE SCALAR' in d
    assert 'DATASET "f8" { DATATYPE H5T_IEEE_F64LE DATASPACE SIMPLE { ( 3, 5, 40 ) / ( 3, 5, 40 ) }' in d
    assert "PREPROCESSING SHUFFLE COMPRESSION DEFLATE { LEVEL 4 }" in d
    full = _dump(p, "-d", "/a/b/i4")
    assert "(0): 0, 1, 2, 3, 4, 5, 6" in full
    # a planted mistake: a file whose complex pair is (i, r) is told apart
    with h5.File(tmp_path / "bad.h5", "w") as f:
        f.create_dataset("c", arrays["c16"].imag + 1j * arrays["c16"].real)
    with h5.File(tmp_path / "bad.h5", "r") as f:
        assert not np.array_equal(np.asarray(f["c"]), arrays["c16"])


def _sensors():
    rng = np.random.default_rng(7)
    cbs = lmc.FarFieldCBSSensor(0.1, 2 * math.pi, 1.0, 0.1 / 16, 2 * math.pi / 4, 0.25)
    flu = lmc.PlanarFluenceSensor(0.0, 2.0, 2.0, 1.0, 0.25, 0.25, 0.5, absorb=False)
    fld = lmc.PlanarFieldSensor(0.0, 2.0, 2.0, 0.25, 0.5)
    st = lmc.StatisticsSensor(0.0)
    st.set_events_histogram_bins(50)
    st.set_theta_histogram_bins(0.0, math.pi / 2, 9)
    st.set_weight_histogram_bins(1.0, 5)
    rec = lmc.PhotonRecordSensor(0.0)
    rows = rng.normal(size=(11, 34))
    rows[:, 1] = rng.integers(1, 40, 11)
    rec.add_records(rows)
    for s in (cbs, flu, fld):
        s.buffer[:] = rng.normal(size=s.buffer.size)
    st.buffer[:] = rng.integers(0, 9, st.buffer.size)
    cbs.hits, flu.hits = 1234, 77
    cbs.set_theta_limit(0.0, 0.2)
    return {"farfield_cbs": cbs, "planar_fluence": flu, "planar_field": fld, "statistics": st,
            "photon_records": rec}


def test_experiment_sensors_layout_and_typed_results(tmp_path):
    sens = _sensors()
    with lmc.Experiment("run", base_dir=str(tmp_path), timestamped=False) as exp:
        exp.save_sensors(sens)
        path = exp.h5_path
    assert path == tmp_path / "run" / "results.h5"
    with lmc.ResultsLoader(tmp_path / "run") as ld:
        assert sorted(ld.sensor_names()) == sorted(sens)
        r = ld.far_field_cbs()
        c = sens["farfield_cbs"]
        g = c.grids()
        for q, nm in enumerate(schema.SENSOR_SCHEMAS["FarFieldCBSSensor"].data_attrs):
            assert np.array_equal(getattr(r, nm), g[q]), nm  # (N_t, N_theta, N_phi) per Stokes grid
        assert (r.N_theta, r.N_phi, r.N_t) == (c.N_theta, c.N_phi, c.N_t) and r.meta.hits == 1234
        assert r.meta.type == "FarFieldCBSSensor" and np.array_equal(r.meta.origin, [0, 0, 0])
        assert np.allclose(r.phi_values, c.phi_values) and not r.phi_explicit
        m = ld.sensor_meta("farfield_cbs")
        assert m["filter_theta_enabled"] and m["filter_direction"] == "Both" and m["filter_theta_max"] == 0.2
        f = ld.planar_fluence()
        assert np.array_equal(np.stack([f.S0, f.S1, f.S2, f.S3]), sens["planar_fluence"].grids())
        e = ld.planar_field()
        assert np.array_equal(e.Ex, sens["planar_field"].Ex.data) and e.Ey.dtype == np.complex128
        s = ld.statistics()
        assert np.array_equal(s.events_histogram, sens["statistics"].events_histogram)
        assert np.array_equal(s.weight_histogram, sens["statistics"].weight_histogram)
        assert s.phi_histogram is None and "skipped:phi_histogram" in ld.sensor_meta("statistics")
        pr = ld.photon_records()
        rows = sens["photon_records"].records
        assert pr.n_records == 11 and np.array_equal(pr.events, rows[:, 1].astype(np.int32))
        assert np.array_equal(pr.weight, rows[:, 6]) and np.array_equal(pr.r_last, rows[:, 11:14])
        recs = sens["photon_records"].recorded_photons
        assert np.array_equal(pr.E_reverse[:, 1], [p.polarization_reverse.n for p in recs])
        assert type(ld.sensor("planar_field")).__name__ == "PlanarFieldResult"
        with pytest.raises(TypeError):
            ld.planar_fluence("farfield_cbs")
    if H5LS is None:
        pytest.skip("h5ls not installed: layout not cross-checked")
    lst = _listing(path)
    for name, s in sens.items():
        sch = schema.get_schema(type(s).__name__)
        assert lst[f"/sensors/{name}/meta"] == "Group"
        for a in sch.data_attrs:
            key = f"/sensors/{name}/data/{a}"
            assert key in lst or (name == "statistics" and a in ("phi_histogram", "depth_histogram",
                                                                  "time_histogram")), key
    assert lst["/sensors/photon_records/records/E_forward"] == "Dataset"
    attrs = _dump(path, "-A", "-g", "/sensors/planar_fluence/meta")
    for a in schema.COMMON_META + schema.SENSOR_SCHEMAS["PlanarFluenceSensor"].meta_attrs + ("type",):
        assert f'ATTRIBUTE "{a}"' in attrs, a


def test_params_absorption_derived_and_processed(tmp_path):
    import scenes
    cfg, det = scenes.config_multilayer(1000, seed=3, small=True)
    cfg.n_threads = 4
    ab = lmc.Absorption(2.0, 3.0, 0.1, 0.1, 0.5, 2.0)
    ab.buffer[:] = np.random.default_rng(2).random(ab.buffer.size)
    cbs = det["cbs"]
    cbs.buffer[:] = np.random.default_rng(3).random(cbs.buffer.size)
    proc = lmc.postprocess_farfield_cbs(cbs, 1000)
    with lmc.Experiment("p", base_dir=str(tmp_path), timestamped=False) as exp:
        sp = exp.save_params(cfg, extra={"anisotropy_g": 0.9, "note": "x"}, units={"wavelength": "mm"})
        exp.log_params(grid=np.arange(10.0), label="abc")
        exp.save_absorption(ab, 1000)
        exp.save_derived("cbs/enhancement", np.linspace(1, 2, 5), attrs={"units": "1"})
        exp.save_processed("farfield_cbs", proc, sensor=cbs)
        src = tmp_path / "script.py"
        src.write_text("print(1)\n")
        exp.log_script(str(src))
        exp.log_readme("notes")
        exp.update_log_readme("ok", 1.5, "2026-01-01T00:00:00")
    d = tmp_path / "p"
    assert (d / "script_snapshot.py").read_text() == "print(1)\n"
    assert "Status: ok" in (d / "README.md").read_text()
    with lmc.ResultsLoader(d) as ld:
        p = ld.params
        assert p == sp and p.run.n_threads == 4 and p.laser.source_type == "Gaussian"
        assert [L.kind for L in p.layers] == ["homogeneous"] * 3 and p.layers[1].medium.kind == "MieMedium"
        assert p.layers[2].z_max == math.inf and p.extra["note"] == "x"
        fl = ld.params_flat
        assert fl["n_layers"] == 3 and fl["medium_kind"] == "RGDMedium" and fl["unit:wavelength"] == "mm"
        assert fl["anisotropy_g"] == 0.9 and np.array_equal(fl["grid"], np.arange(10.0)) and fl["label"] == "abc"
        a = ld.absorption()
        assert a.slices.shape == (ab.n_t, ab.n_r, ab.n_z) and np.array_equal(a.slices, ab.grids())
        for k in range(ab.n_t):
            assert np.array_equal(a.images[k], ab.get_absorption_image(1000, k).data)
        assert np.array_equal(ld.absorption_total_data(), ab.grids()[0])
        assert ld.absorption_meta()["n_t"] == ab.n_t
        assert np.array_equal(ld.derived("cbs/enhancement"), np.linspace(1, 2, 5))
        pc = ld.processed_cbs()
        assert pc.N_t == cbs.N_t and np.array_equal(pc.coh_s0[0], proc.coherent[0].S0.data)
        assert np.array_equal(pc.inc_s3[-1], proc.incoherent[-1].S3.data)
        assert np.array_equal(pc.theta, np.linspace(0, cbs.theta_max, cbs.N_theta))


def test_mixture_params_and_errors(tmp_path):
    sp, dens, _ = lmc.lognormal_mie_mixture(0.633e-3, 0.4e-3, 0.3, 1.59, 1.33, 4, 30.0, 0.002, n_div=2000)
    sample = lmc.Sample(1.33)
    sample.add_mixture_layer(sp, dens, 0.0, math.inf)
    cfg = lmc.SimConfig()
    cfg.sample = sample
    cfg.laser = lmc.Laser(1.0, 0.0, 0.633e-3, 0.0, lmc.LaserSource.Point)
    p = lmc.capture_params(cfg)
    L = p.layers[0]
    assert L.kind == "mixture" and len(L.species) == 4 and L.mu_s_total == pytest.approx(30.0, rel=1e-12)
    assert lmc.SimParams.from_json(p.to_json()) == p
    q = lmc.derived_quantities_mixture(sp, dens)
    assert q["mu_s_total"] == pytest.approx(L.mu_s_total, rel=1e-9) and q["n_species"] == 4
    with pytest.raises(ValueError):
        lmc.capture_params(lmc.SimConfig())
    with lmc.Experiment("e", base_dir=str(tmp_path), timestamped=False) as exp:
        with pytest.raises(KeyError):
            exp.save_sensor(lmc.ReflectanceSensor(0.0), "refl")  # a type the file layout has no entry for
        with pytest.raises(TypeError):
            exp.save_processed("x", object())
    with lmc.ResultsLoader(tmp_path / "e") as ld:
        with pytest.raises(KeyError):
            ld.params
    with pytest.raises(FileNotFoundError):
        lmc.ResultsLoader(tmp_path / "missing")
