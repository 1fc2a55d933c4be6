"""Host assembly of the C-ABI library against the oracle (CPU, no device).

``sfx_engine_create_host_2d`` runs the engine's host half only: numbering
(radix sort by (x, z), merge, first-visit renumbering), compute ordering and
the source/receiver locator (grid-pruned nearest GLL point, corner-index
candidates, Newton).  Their results must be bit-exact with the oracle's
restatement of compute_mesh.cpp:45-170 / :240-279 and locate_point.cpp,
which scans every point and every element:

* on the reference's 9-node 80 x 60 databases;
* on a structured mesh whose interior nodes are moved at random by up to
  a quarter element (non-rectangular elements: the grid index and the
  bounding-box pruning see skewed boxes);
* for targets exactly on GLL points, element corners and edges, near the
  boundary and outside the mesh (where Newton clamps at +-1.01).
"""
import numpy as np
import pytest

import oracle
from conftest import load_case
from specfempp_amd import meshes


def _host_engine(mesh, sources, receivers, dt=1e-3, nstep=4):
    from specfempp_amd import Engine2D

    return Engine2D(mesh, sources, receivers, dt, nstep, 1, host_only=True)


def _compare(mesh, sources, receivers, dt=1e-3):
    eng = _host_engine(mesh, sources, receivers, dt)
    orc = oracle.Oracle2D(mesh, 4, 1, dt)
    ibool, midx, c2m = orc.numbering()
    assert np.array_equal(eng.index_mapping(), ibool)
    assert np.array_equal(eng.field_index(), midx)
    assert np.array_equal(eng.compute_to_mesh(), c2m)
    for which, pts in (("sources", sources), ("receivers", receivers)):
        ispec, loc = eng.locations(which)
        for k, p in enumerate(pts):
            e, xi, ga = orc.locate(p["x"], p["z"])
            assert ispec[k] == e, (which, k, p, ispec[k], e)
            assert np.float32(loc[k, 0]) == np.float32(xi) and np.float32(loc[k, 1]) == np.float32(ga), (which, k, p)
    eng.close()


@pytest.mark.parametrize("case", ["elastic_free_80x60", "elastic_stacey_80x60"])
def test_reference_databases(case):
    mesh, cfg = load_case(case)
    _compare(mesh, cfg["sources"], cfg["stations"], cfg["dt"])


def _jittered(n=30, h=10.0, seed=3):
    mesh = meshes.quad_mesh(n, n, 0.0, n * h, 0.0, n * h, materials=[meshes.elastic(2000.0, 3000.0, 1700.0)],
                            absorbing=("bottom", "right", "left", "top"))
    rng = np.random.default_rng(seed)
    c = mesh.coorg.copy()
    iz, ix = np.divmod(np.arange(c.shape[1]), n + 1)
    inner = (ix > 0) & (ix < n) & (iz > 0) & (iz < n)
    c[:, inner] += rng.uniform(-0.25 * h, 0.25 * h, size=(2, inner.sum()))
    mesh.coorg = c
    return mesh


def test_jittered_mesh_and_edge_case_targets():
    n, h = 30, 10.0
    mesh = _jittered(n, h)
    rng = np.random.default_rng(8)
    pts = [dict(x=float(x), z=float(z)) for x, z in rng.uniform(0, n * h, size=(60, 2))]
    nodes = mesh.coorg
    for k in rng.choice(nodes.shape[1], 20, replace=False):        # element corners
        pts.append(dict(x=float(nodes[0, k]), z=float(nodes[1, k])))
    for k in rng.choice(nodes.shape[1] - n - 2, 10, replace=False):  # mid-edges
        pts.append(dict(x=float(0.5 * (nodes[0, k] + nodes[0, k + 1])), z=float(0.5 * (nodes[1, k] + nodes[1, k + 1]))))
    pts += [dict(x=-15.0, z=140.0), dict(x=n * h + 7.0, z=-3.0), dict(x=1e-4, z=n * h - 1e-4),
            dict(x=150.0, z=n * h + 40.0), dict(x=0.0, z=0.0)]
    src = [dict(x=p["x"], z=p["z"], angle=0.0, f0=10.0, factor=1.0, tshift=0.0) for p in pts[:8]]
    _compare(mesh, src, pts)


def test_targets_on_gll_points_of_a_regular_mesh():
    w = meshes.config2(n=20, L=50.0, nstep=4, nrec=4, seed=9, corr=10.0)
    orc = oracle.Oracle2D(w.mesh, 4, 1, w.dt)
    coord, _ = orc.geometry()
    rng = np.random.default_rng(10)
    flat = coord.reshape(-1, 2)
    pts = [dict(x=float(x), z=float(z)) for x, z in flat[rng.choice(len(flat), 40, replace=False)]]
    _compare(w.mesh, w.sources, pts, w.dt)


def test_3d_block_path_host_assembly_numbers_like_the_oracle():
    from specfempp_amd import Engine3D

    w = meshes.config3(n=4, nstep=4, h=100.0, nrec_side=2)
    eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, scatter="block", host_only=True)
    assert np.array_equal(eng.index_mapping(), oracle.Oracle3D(w.mesh, 8, w.dt).ibool())
    eng.close()


def test_host_only_engine_reports_the_run_sizes():
    """Host-only engines carry every host-side quantity a batch packs: the
    sample count and the element coloring (a batch of host-only engines
    depends on both)."""
    from specfempp_amd import Engine2D

    mesh, cfg = load_case("elastic_free_80x60")
    e = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], 37, 4, ("displacement", "velocity"),
                 host_only=True)
    assert e.nsig == 10 and e.nrec == len(cfg["stations"]) and e.nsrc == len(cfg["sources"])
    e.close()


@pytest.mark.parametrize("chunk", [2, 32, 64])
def test_storage_chunk_numbering_matches_oracle(chunk):
    """Host assembly numbering at storage_chunk_size > 1 (chunk_config.hpp:12-28)
    equals the oracle's on the reference's 80 x 60 Stacey mesh."""
    from specfempp_amd import Engine2D

    mesh, cfg = load_case("elastic_stacey_80x60")
    eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], 4, 1, ("displacement",), chunk_size=chunk,
                   host_only=True)
    ibool, midx, c2m = oracle.Oracle2D(mesh, 4, chunk, cfg["dt"]).numbering()
    assert np.array_equal(eng.index_mapping(), ibool)
    assert np.array_equal(eng.field_index(), midx)
    assert np.array_equal(eng.compute_to_mesh(), c2m)
    eng.close()
