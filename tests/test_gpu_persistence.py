"""A GPU run saved and read back through the HDF5 layer (manager.py:802-942,
944-1288): the multilayer scene with every grid sensor filled by
run_simulation_parallel on the device, Experiment.save_sensor /
save_processed / save_params, and ResultsLoader's typed results equal to the
sensors' own arrays bit for bit."""
import numpy as np
import pytest

import luminis_mc as lmc
import scenes
from luminis_mc import h5

pytestmark = pytest.mark.gpu


def test_gpu_run_saved_and_loaded(require_gpu, tmp_path):
    try:
        h5.library()
    except RuntimeError as e:
        pytest.skip(str(e))
    cfg, det = scenes.config_multilayer(50000, seed=11, small=True, with_field=True)
    lmc.run_simulation_parallel(cfg)
    cbs, flu, fld = det["cbs"], det["fluence"], det["field"]
    assert cbs.hits > 0 and flu.hits > 0 and fld.hits > 0
    proc = lmc.postprocess_farfield_cbs(cbs, cfg.n_photons)
    with lmc.Experiment("gpu", base_dir=str(tmp_path), timestamped=False) as exp:
        exp.save_params(cfg)
        exp.save_sensors({"farfield_cbs": cbs, "planar_fluence": flu, "planar_field": fld})
        exp.save_processed("farfield_cbs", proc, sensor=cbs)
    with lmc.ResultsLoader(tmp_path / "gpu") as ld:
        assert ld.params.run.n_photons == 50000 and ld.params.run.seed == 11
        r = ld.far_field_cbs()
        assert np.array_equal(np.stack([r.S0_coh, r.S1_coh, r.S2_coh, r.S3_coh, r.S0_incoh, r.S1_incoh, r.S2_incoh,
                                        r.S3_incoh]), cbs.grids())
        assert r.meta.hits == cbs.hits
        f = ld.planar_fluence()
        assert np.array_equal(np.stack([f.S0, f.S1, f.S2, f.S3]), flu.grids()) and f.meta.hits == flu.hits
        e = ld.planar_field()
        assert np.array_equal(e.Ex, fld.Ex.data) and np.array_equal(e.Ey, fld.Ey.data)
        pc = ld.processed_cbs()
        assert all(np.array_equal(pc.coh_s0[t], proc.coherent[t].S0.data) for t in range(pc.N_t))
