"""The reference's mesh-reader fixtures (tests/unit-tests/data/mesh: 80 x 60
elements of 9 nodes under flat and curved topography) through the reader,
the oracle and the host assembly of the engine (no GPU needed).

- The reader gives the fixture's element count, control-node count and one
  elastic material; the oracle's per-point properties equal the reader test's
  known answer (tests/unit-tests/mesh/materials/properties.cpp:15-35:
  kappa, mu, rho).
- The engine's numbering, per-medium numbering, compute order and the
  source/receiver locations (three stations sit on the curved free surface,
  so their element search ends in the Newton iterations of locate_point.cpp)
  equal the oracle's bit for bit.
"""
import numpy as np
import pytest

import oracle
from conftest import load_case

CASES = ["flat_topography_80x60", "curved_topography_80x60"]


@pytest.mark.parametrize("case", CASES)
def test_reader_and_material_known_answer(case):
    mesh, cfg = load_case(case)
    assert mesh.nspec == cfg["nspec"] and mesh.ngnod == cfg["ngnod"]
    assert len(mesh.materials) == 1
    rho, kappa, mu, wj = oracle.Oracle2D(mesh, 4, 1, cfg["dt"]).properties()
    ka = cfg["properties_known_answer"]
    for got, want in ((kappa, ka["kappa"]), (mu, ka["mu"]), (rho, ka["rho"])):
        # FP32 properties (the reference's type_real) against the known answer
        assert np.abs(got / want - 1).max() <= 2e-7
    assert np.all(wj > 0)


@pytest.mark.parametrize("case", CASES)
def test_host_assembly_numbering_and_locations_bit_exact(case):
    from specfempp_amd import Engine2D

    mesh, cfg = load_case(case)
    eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], cfg["nstep"], 1, host_only=True)
    orc = oracle.Oracle2D(mesh, 4, 1, cfg["dt"])
    ibool, midx, c2m = orc.numbering()
    assert np.array_equal(eng.index_mapping(), ibool)
    assert np.array_equal(eng.field_index(), midx)
    assert np.array_equal(eng.compute_to_mesh(), c2m)
    for which, pts in (("sources", cfg["sources"]), ("receivers", cfg["stations"])):
        ispec, loc = eng.locations(which)
        for k, p in enumerate(pts):
            e, xi, ga = orc.locate(p["x"], p["z"])
            assert ispec[k] == e
            assert np.float32(loc[k, 0]) == np.float32(xi) and np.float32(loc[k, 1]) == np.float32(ga)
    eng.close()
