"""GPU runs driven by the reference's own run inputs.

* The reference's test configurations of the golden cases (specfem_config.yaml,
  sources.yaml, STATIONS, unchanged) run through ``io.config.run_config``:
  the traces written in the reference's ASCII format match SPECFEM2D's within
  the reference's own bar (relative L1 per station <= 1e-3,
  newmark_tests.cpp:325-346), and the displacement matches the FP32 oracle
  within the forward bar (relative L2 <= 1e-5).
* A user start time (time-scheme t0) gives the FP32 oracle's traces within
  the same bar.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_case, rel_l1_per_station, rel_l2
from specfempp_amd.io import config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["elastic_free_80x60", "elastic_stacey_80x60"])
def test_reference_configuration_runs_unmodified(case, tmp_path):
    mesh, cfg = load_case(case)
    seis, t0, run = config.run_config(os.path.join(GOLDEN, case, "specfem_config.yaml"), out_dir=str(tmp_path))
    traces = np.load(os.path.join(GOLDEN, case, "traces.npz"))
    types = run.seismogram_types
    for k in cfg["trace_kinds"]:
        t = {"d": "displacement", "v": "velocity", "a": "acceleration"}[k]
        ours = seis[:, types.index(t)].astype(np.float64)
        err = rel_l1_per_station(ours.transpose(1, 2, 0), traces[k]).max()
        print(f"[config] {case}/{k}: vs_specfem2d_l1={err:.3e}", flush=True)
        assert err <= cfg["tolerance"]
    s32, _ = oracle.Oracle2D(mesh, 4, 1, run.dt).forward(run.sources, run.receivers, run.nstep, run.dt, 1,
                                                          ("displacement",))
    e32 = rel_l2(seis[:, types.index("displacement")], s32[:, 0])
    print(f"[config] {case}: vs_oracle_fp32={e32:.3e}", flush=True)
    assert e32 <= 1e-5
    r0 = run.receivers[0]
    sem = np.loadtxt(os.path.join(tmp_path, f"{r0['network']}.{r0['name']}.S2.BXZ.semd"))
    assert sem.shape == (run.nstep, 2)
    np.testing.assert_allclose(sem[:, 1], seis[:, types.index("displacement"), 0, 1], rtol=1e-6, atol=1e-30)


def test_user_start_time_matches_oracle():
    from specfempp_amd import Engine2D

    mesh, cfg = load_case("elastic_stacey_80x60")
    start = -0.2
    e = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], cfg["nstep"], 1, ("displacement",),
                 start_time=start)
    e.initialize()
    e.run()
    ours = e.seismograms().astype(np.float64)
    assert abs(e.t0 - start) < 1e-7
    e.close()
    o = oracle.Oracle2D(mesh, 4, 1, cfg["dt"])
    o.set_start_time(start)
    s32, t0 = o.forward(cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"], 1, ("displacement",))
    e32 = rel_l2(ours, s32)
    print(f"[config] start_time {start}: vs_oracle_fp32={e32:.3e}", flush=True)
    assert e32 <= 1e-5
