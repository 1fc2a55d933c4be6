"""Host assembly of meshes with acoustic elements (no GPU needed): the
reference's acoustic and coupled elastic-acoustic Newmark fixtures
(tests/unit-tests/displacement_tests/Newmark/serial/*) through the engine's
host-only build, against the oracle bit for bit.

- compute order: the six element groups of material_definitions.hpp:76-103
  (elastic none, elastic Stacey, acoustic none, free surface, Stacey,
  composite), each in mesh order;
- numbering and both per-medium numberings (field_impl.tpp:39-62, elements
  of the other medium skipped);
- source and receiver element and reference coordinates.

The oracle's acoustic pieces are pinned by the six SPECFEM2D trace fixtures
(tests/test_oracle_fixtures.py).
"""
import numpy as np
import pytest

import oracle
from conftest import load_case

CASES = ["acoustic_free_80x60", "acoustic_stacey_80x60", "acoustic_composite_80x60", "acoustic_dirichlet_80x60",
         "coupled_acoustic_top", "coupled_acoustic_bottom"]


@pytest.mark.parametrize("case", CASES)
def test_acoustic_numbering_media_and_locations_bit_exact(case):
    from specfempp_amd import Engine2D

    mesh, cfg = load_case(case)
    eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], cfg["nstep"], 1, host_only=True)
    orc = oracle.Oracle2D(mesh, 4, 1, cfg["dt"])
    ibool, midx, c2m = orc.numbering()
    idx_ac, medium, _ = orc.acoustic()
    assert np.array_equal(eng.compute_to_mesh(), c2m)
    assert np.array_equal(eng.index_mapping(), ibool)
    assert np.array_equal(eng.field_index(), midx)
    assert np.array_equal(eng.field_index_acoustic(), idx_ac)
    assert eng.nspec_acoustic == int(medium.sum()) and eng.nspec_elastic == mesh.nspec - int(medium.sum())
    assert (eng.nglob_elastic, eng.nglob_acoustic) == (orc.nglob_el, orc.nglob_ac)
    # elastic elements come first in compute order
    assert np.all(medium[:eng.nspec_elastic] == 0) and np.all(medium[eng.nspec_elastic:] == 1)
    for which, pts in (("sources", cfg["sources"]), ("receivers", cfg["stations"])):
        ispec, loc = eng.locations(which)
        for k, p in enumerate(pts):
            e, xi, ga = orc.locate(p["x"], p["z"])
            assert ispec[k] == e
            assert np.float32(loc[k, 0]) == np.float32(xi) and np.float32(loc[k, 1]) == np.float32(ga)
    eng.close()


def test_coupled_mesh_refuses_the_paths_it_does_not_serve():
    from specfempp_amd import Engine2D, SfxError

    mesh, cfg = load_case("coupled_acoustic_top")
    with pytest.raises(SfxError, match="SFX_SCATTER_COLORED or _ATOMIC"):
        Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], 10, 1, scatter="gather", host_only=True)
    # a free-surface entry on an elastic element
    el = int(np.flatnonzero([not mesh.materials[m].is_acoustic for m in mesh.kmato])[0])
    mesh.afs_ispec = np.array([el], dtype=np.int32)
    mesh.afs_type = np.array([2], dtype=np.int32)
    with pytest.raises(SfxError, match="not an acoustic element"):
        Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], 10, 1, host_only=True)
