"""The reference's ASCII I/O layout (include/io/ASCII/impl/*) and the
checkpoint / adjoint-workflow inputs, on the CPU:

- dataset files: ``.meta`` text ("rank R" / "dims d0 d1 ... "), one value per
  line, floats as ``std::scientific`` with 10 digits (native_type.tpp:118-136),
  memory order of LayoutLeft (column-major) views; float32 values survive the
  text round trip bit for bit; a reader checks the dims (datasetbase.hpp);
- the combined configuration's adjoint sources with External ASCII time
  functions and their start-time / step checks (external.cpp:82-100).
"""
import os

import numpy as np
import pytest
import yaml

from specfempp_amd.io import config, wavefield


def test_dataset_text_layout(tmp_path):
    a = np.array([[1.0, 2.5, -3.0], [4.0, 5.0, 6.0e-12]], dtype=np.float32)  # (2, 3) LayoutLeft
    wavefield.write_dataset(str(tmp_path), "field", a)
    assert open(tmp_path / "field.meta").read() == "rank 2\ndims 2 3 \n"
    lines = open(tmp_path / "field.txt").read().splitlines()
    assert lines[:3] == ["1.0000000000e+00", "4.0000000000e+00", "2.5000000000e+00"]  # column-major
    assert lines[-1] == "5.9999999760e-12"  # the float32 value as std::scientific, 10 digits, prints it
    i = np.array([-1, 0, 7], dtype=np.int32)
    wavefield.write_dataset(str(tmp_path), "idx", i)
    assert open(tmp_path / "idx.txt").read() == "-1\n0\n7\n"


@pytest.mark.parametrize("order", ["F", "C"])
def test_dataset_round_trip_is_exact(tmp_path, order):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal((7, 5, 5, 3, 2)) * 10.0 ** rng.integers(-30, 30, size=(7, 5, 5, 3, 2))).astype(np.float32)
    wavefield.write_dataset(str(tmp_path), "v", a, order=order)
    b = wavefield.read_dataset(str(tmp_path), "v", np.float32, order=order, shape=a.shape)
    assert np.array_equal(a, b)
    empty = np.zeros((0, 5, 5, 3, 1), dtype=np.float32)
    wavefield.write_dataset(str(tmp_path), "e", empty)
    assert wavefield.read_dataset(str(tmp_path), "e", np.float32).shape == empty.shape
    with pytest.raises(ValueError, match="do not match"):
        wavefield.read_dataset(str(tmp_path), "v", np.float32, shape=(7, 5, 5, 4, 2))


def _combined_config(tmp_path, t0, dt, nstep, bad=None):
    """A combined configuration in the reference's form (the Tromp 2005
    benchmark's adjoint_config.yaml / adjoint_source.yaml structure)."""
    adj = tmp_path / "adjoint_sources"
    adj.mkdir()
    t = t0 + dt * np.arange(nstep)
    for comp, f in (("BXX", 1.0), ("BXZ", -2.0)):
        tt = t + (1e-3 if bad == "t0" and comp == "BXX" else 0.0)
        np.savetxt(adj / f"AA.S0001.S2.{comp}.adj", np.stack([tt, f * np.sin(t)], axis=1))
    sources = {"number-of-sources": 2, "sources": [
        {"force": {"x": 2500.0, "z": 2500.0, "angle": 0.0, "source_surf": False, "vx": 0.0, "vz": 0.0,
                   "Ricker": {"f0": 10.0, "factor": 1e10, "tshift": 0.0}}},
        {"adjoint-source": {"x": 1500.0, "z": 2000.0, "station_name": "AA", "network_name": "S0001", "angle": 0.0,
                            "source-surf": False, "vx": 0.0, "vz": 0.0,
                            "External": {"format": "ASCII", "stf": {
                                "X-component": "adjoint_sources/AA.S0001.S2.BXX.adj",
                                "Z-component": "adjoint_sources/AA.S0001.S2.BXZ.adj"}}}}]}
    with open(tmp_path / "adjoint_source.yaml", "w") as fh:
        yaml.safe_dump(sources, fh)
    (tmp_path / "database.bin").write_bytes(b"")
    (tmp_path / "STATIONS").write_text("S0001 AA 1500.0 2000.0 0.0 0.0\n")
    doc = {"parameters": {
        "header": {"title": "combined"},
        "simulation-setup": {"quadrature": {"quadrature-type": "GLL4"},
                             "simulation-mode": {"combined": {
                                 "reader": {"wavefield": {"directory": "OUTPUT_FILES", "format": "ASCII"}},
                                 "writer": {"kernels": {"directory": "OUTPUT_FILES", "format": "ASCII"}}}},
                             "solver": {"time-marching": {"type-of-simulation": "combined",
                                                          "time-scheme": {"type": "Newmark", "dt": dt,
                                                                          "nstep": nstep, "t0": -t0}}}},
        "receivers": {"stations": "STATIONS", "angle": 0.0, "nstep_between_samples": 1,
                      "seismogram-type": ["displacement"]},
        "databases": {"mesh-database": "database.bin"},
        "sources": "adjoint_source.yaml"}}
    path = tmp_path / "adjoint_config.yaml"
    with open(path, "w") as fh:
        yaml.safe_dump(doc, fh)
    return str(path)


def test_combined_configuration_and_external_time_functions(tmp_path):
    t0, dt, nstep = -0.12, 1e-3, 50
    cfg = config.read_config(_combined_config(tmp_path, t0, dt, nstep))
    assert cfg.simulation == "combined" and len(cfg.sources) == 1 and len(cfg.adjoint_sources) == 1
    a = cfg.adjoint_sources[0]
    assert (a["x"], a["z"], a["station"], a["network"]) == (1500.0, 2000.0, "AA", "S0001")
    assert cfg.wavefield_reader["directory"] == os.path.join(str(tmp_path), "OUTPUT_FILES")
    assert cfg.kernel_writer["format"] == "ASCII"
    stf = config.read_external_stf(cfg.adjoint_sources, nstep, dt, t0)
    t = t0 + dt * np.arange(nstep)
    np.testing.assert_allclose(stf[0, :, 0], np.sin(t), rtol=1e-12)
    np.testing.assert_allclose(stf[0, :, 1], -2.0 * np.sin(t), rtol=1e-12)
    with pytest.raises(ValueError, match="time step"):
        config.read_external_stf(cfg.adjoint_sources, nstep, 2 * dt, t0)


def test_external_time_function_start_time_is_checked(tmp_path):
    t0, dt, nstep = -0.12, 1e-3, 20
    cfg = config.read_config(_combined_config(tmp_path, t0, dt, nstep, bad="t0"))
    with pytest.raises(ValueError, match="start time"):
        config.read_external_stf(cfg.adjoint_sources, nstep, dt, t0)


def test_hdf5_format_is_refused_with_a_reason(tmp_path):
    path = _combined_config(tmp_path, -0.1, 1e-3, 10)
    doc = yaml.safe_load(open(path))
    doc["parameters"]["simulation-setup"]["simulation-mode"]["combined"]["reader"]["wavefield"]["format"] = "HDF5"
    with open(path, "w") as fh:
        yaml.safe_dump(doc, fh)
    with pytest.raises(NotImplementedError, match="ASCII"):
        config.read_config(path)
