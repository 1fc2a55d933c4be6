"""Reading the reference's run inputs (CPU): sources.yaml, STATIONS and
specfem_config.yaml, and the user start time.

* The source reader gives the values the reference's own reader test expects
  for its data files (tests/unit-tests/io/sources/test_read_sources.cpp:25-46):
  force at (2500, 2500), angle 0, Ricker f0 10, tshift 5, factor 1e10; moment
  tensor at (2000, 3000), Mxx = Mzz = 1, Mxz = 0, Ricker f0 1, tshift 30,
  factor 1e10.  The Cosserat force is outside this engine's path and is
  refused with a clear error.
* The reference's test configurations of the two golden cases parse to the
  dt, step count, seismogram types, stations and sources those cases hold.
* User start time (time-scheme t0): in the FP64 oracle, of two starts well
  before the wavelet's onset, the one k steps earlier delays every trace by
  exactly k samples (the system is time invariant), to 1e-10 of the peak; a
  start later than the earliest source onset allows is refused by the oracle
  and by the engine's host assembly.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_case
from specfempp_amd.io import config


def test_source_reader_matches_reference_expectations():
    d = os.path.join(GOLDEN, "io_sources")
    (f,) = config.read_sources(os.path.join(d, "single_force.yaml"))
    assert f == dict(type="force", x=2500.0, z=2500.0, angle=0.0, f0=10.0, factor=1e10, tshift=5.0)
    (m,) = config.read_sources(os.path.join(d, "single_moment_tensor.yaml"))
    assert m == dict(type="moment_tensor", x=2000.0, z=3000.0, mxx=1.0, mzz=1.0, mxz=0.0, f0=1.0, factor=1e10,
                     tshift=30.0)
    with pytest.raises(NotImplementedError, match="cosserat-force"):
        config.read_sources(os.path.join(d, "single_cosserat_force.yaml"))


def test_source_reader_validation():
    ricker = {"f0": 5.0, "factor": 1.0}
    with pytest.raises(ValueError, match="number-of-sources"):
        config.read_sources({"number-of-sources": 2, "sources": [{"force": {"x": 0, "z": 0, "Ricker": ricker}}]})
    with pytest.raises(ValueError, match="adjoint"):
        config.read_sources({"number-of-sources": 2, "sources": [
            {"force": {"x": 0, "z": 0, "Ricker": ricker}},
            {"adjoint-source": {"station_name": "S", "network_name": "N", "x": 1, "z": 1}}]})
    with pytest.raises(NotImplementedError, match="Dirac"):
        config.read_sources({"number-of-sources": 1, "sources": [{"force": {"x": 0, "z": 0, "Dirac": ricker}}]})
    rec = config.read_receivers([{"network": "AA", "station": "S1", "x": 1.5, "z": 2.5}], angle=30.0)
    assert rec == [dict(name="S1", network="AA", x=1.5, z=2.5, angle=30.0)]


@pytest.mark.parametrize("case", ["elastic_free_80x60", "elastic_stacey_80x60"])
def test_reference_configurations_parse(case):
    _, cfg = load_case(case)
    run = config.read_config(os.path.join(GOLDEN, case, "specfem_config.yaml"))
    assert run.dt == cfg["dt"] and run.nstep == cfg["nstep"] and run.start_time == 0.0
    assert run.simulation == "forward" and run.nstep_between_samples == 1
    assert run.database.endswith("database.bin.gz")
    assert [(r["name"], r["network"], r["x"], r["z"]) for r in run.receivers] == \
        [(s["name"], s["network"], s["x"], s["z"]) for s in cfg["stations"]]
    assert run.sources == [dict(type="force", **s) for s in cfg["sources"]]


def test_user_start_time_delays_traces_by_whole_steps():
    mesh, cfg = load_case("elastic_free_80x60")
    rec = cfg["stations"][:6]
    dt = cfg["dt"]
    o = oracle.Oracle2D(mesh, 8, 1, dt)
    auto, t0 = o.forward(cfg["sources"], rec, 300, dt)
    # both starts lie 0.1 s (and more) before the automatic one, where the
    # Ricker wavelet is below 1e-40 of its peak, so the only difference of
    # the two runs is the k-step delay
    k, ta = 10, t0 - 0.1
    o.set_start_time(ta)
    a, t1 = o.forward(cfg["sources"], rec, 300, dt)
    o.set_start_time(ta - k * dt)
    b, t2 = o.forward(cfg["sources"], rec, 300, dt)
    assert abs(t1 - ta) < 1e-15 and abs(t2 - (ta - k * dt)) < 1e-15
    peak = np.abs(auto).max()
    assert np.abs(b[:k]).max() < 1e-12 * peak and np.abs(a).max() > 0.1 * peak
    np.testing.assert_allclose(b[k:], a[:-k], rtol=0, atol=1e-10 * peak)
    o.set_start_time(t0 + 0.01)
    with pytest.raises(Exception, match="start time"):
        o.forward(cfg["sources"], rec, 10, dt)


def test_engine_host_assembly_takes_the_user_start_time():
    from specfempp_amd import Engine2D, SfxError

    mesh, cfg = load_case("elastic_free_80x60")
    e = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], 20, host_only=True, start_time=-0.25)
    assert e.t0 == np.float32(-0.25)
    e.close()
    with pytest.raises(SfxError, match="start_time"):
        Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], 20, host_only=True, start_time=-0.05)
