"""Seismogram read-out (CPU): rotation identities and the ASCII format."""
import numpy as np

from specfempp_amd.io.fortran_db import read_trace
from specfempp_amd.io.seismograms import station_traces, write_ascii


def test_rotation_identities():
    rng = np.random.default_rng(0)
    seis = rng.standard_normal((7, 2, 3, 2)).astype(np.float32)
    same, t = station_traces(seis, -0.12, 1e-3, 2, [0.0, 0.0, 0.0])
    assert np.array_equal(same, seis)
    assert np.allclose(t, -0.12 + np.arange(7) * 2e-3, atol=1e-7)
    quarter, _ = station_traces(seis, 0.0, 1e-3, 1, [90.0] * 3)
    assert np.allclose(quarter[..., 0], -seis[..., 1], atol=1e-6)
    assert np.allclose(quarter[..., 1], seis[..., 0], atol=1e-6)
    rot, _ = station_traces(seis, 0.0, 1e-3, 1, [30.0, -45.0, 200.0])
    assert np.allclose(np.linalg.norm(rot, axis=-1), np.linalg.norm(seis, axis=-1), rtol=1e-6)


def test_ascii_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    seis = rng.standard_normal((5, 2, 2, 2)).astype(np.float32)
    stations = [("S0001", "AA", 0.0), ("S0002", "AA", 0.0)]
    paths = write_ascii(str(tmp_path), stations, seis, ("displacement", "pressure"), -0.1, 1e-3)
    names = sorted(p.split("/")[-1] for p in paths)
    assert names == ["AA.S0001.S2.BXX.semd", "AA.S0001.S2.BXZ.semd", "AA.S0001.S2.PRE.semp",
                     "AA.S0002.S2.BXX.semd", "AA.S0002.S2.BXZ.semd", "AA.S0002.S2.PRE.semp"]
    tr = read_trace(str(tmp_path / "AA.S0002.S2.BXZ.semd"))
    assert tr.shape == (5, 2)
    assert np.allclose(tr[:, 1], seis[:, 0, 1, 1], rtol=1e-6)
    assert np.allclose(tr[:, 0], -0.1 + np.arange(5) * 1e-3, atol=1e-7)
