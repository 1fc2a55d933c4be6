import gzip
import json
import os
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


def load_case(name):
    """(Mesh2D, case.json dict) of a golden fixture directory."""
    from specfempp_amd.io.fortran_db import read_database_2d

    d = os.path.join(GOLDEN, name)
    with gzip.open(os.path.join(d, "database.bin.gz")) as fh:
        data = fh.read()
    fd, path = tempfile.mkstemp(suffix=".bin")
    try:
        with os.fdopen(fd, "wb") as out:
            out.write(data)
        mesh = read_database_2d(path)
    finally:
        os.remove(path)
    cfg = {}
    if os.path.exists(os.path.join(d, "case.json")):
        with open(os.path.join(d, "case.json")) as fh:
            cfg = json.load(fh)
    return mesh, cfg


def rel_l1_per_station(ours, ref):
    """Reference metric of newmark_tests.cpp:325-346: sum|x - x_ref| / sum|x_ref|
    per station over components and time.  Arrays are (nrec, 2, nsamples)."""
    return np.abs(ours - ref).sum(axis=(1, 2)) / np.abs(ref).sum(axis=(1, 2))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
