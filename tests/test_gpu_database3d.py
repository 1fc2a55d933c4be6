"""GPU parity of the 3-D mesh-database path (general hexahedra with per-point
geometry on irregular elements; the mesh layout of
src/io/mesh/impl/fortran/dim3/mesh.cpp) through the C ABI.

- An irregular mesh (interior nodes of half the mesh moved, two materials,
  Stacey on five faces, a moment tensor and a force, receivers inside and on
  the surface), COLORED and ATOMIC, against the FP32 and FP64 oracle on the
  same database: displacement within 1e-5 relative L2 of the FP32 oracle,
  velocity and acceleration on the FP32 rounding-floor rule; final
  wavefields on the same bars.
- The database read back from its files runs as the one written.
- An undistorted cube written as a database gives the structured engine's
  seismograms (regular elements; Stacey from unit normals and jacobian2Dw
  against the structured J grad xi form; the moment tensor on an element edge,
  located in the same element by both location rules): displacement within
  1e-5, velocity and acceleration within twice the FP32 rounding floor (the
  structured FP32 oracle's distance to the FP64 one; two FP32 runs that
  round differently are each about that far from the exact result).
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from test_gpu_parity import _check, _report

pytestmark = pytest.mark.gpu

TYPES = ("displacement", "velocity", "acceleration")


def _workload():
    from specfempp_amd import meshes

    nx, ny, nz, dt = 8, 6, 6, 1e-3
    mat = np.zeros((nz, ny, nx), np.int32)
    mat[nz // 2:] = 1
    db = meshes.hex_database(nx, ny, nz, 100.0, mat, [2500.0, 2700.0], [3000.0, 3600.0], [1730.0, 2000.0],
                             perturb=0.12, seed=3, dt=dt)
    src = [dict(type="moment_tensor", x=240.0, y=260.0, z=230.0, mxx=0.3, myy=-0.5, mzz=0.8, mxy=0.2, mxz=-0.1,
                myz=0.4, f0=15.0, factor=1e12),
           dict(type="force", x=610.0, y=390.0, z=330.0, fx=0.3, fy=-0.5, fz=0.8, f0=15.0, factor=1e10)]
    rec = [dict(x=150.0, y=150.0, z=150.0), dict(x=40.0, y=60.0, z=600.0), dict(x=660.0, y=450.0, z=300.0),
           dict(x=400.0, y=300.0, z=600.0), dict(x=720.0, y=520.0, z=80.0)]
    return db, src, rec, dt


@pytest.mark.parametrize("scatter", ["colored", "atomic"])
def test_database_mesh_matches_oracle(scatter, tmp_path):
    from specfempp_amd import Engine3DDatabase
    from specfempp_amd.io import database3d as d3

    db, src, rec, dt = _workload()
    assert 0 < len(db.irregular_element_number) < db.nspec
    nstep = 400
    # through the files, as the reference reads them
    P, D = str(tmp_path / "mesh_parameters.bin"), str(tmp_path / "database.bin")
    d3.write_database_3d(P, D, db)
    db2 = d3.read_database_3d(P, D)
    print(f"[case] 3-D database {db.nspec} elements ({len(db.irregular_element_number)} irregular), "
          f"{nstep} steps, scatter={scatter}", flush=True)
    eng = Engine3DDatabase(db2, src, rec, dt, nstep, 1, TYPES, scatter=scatter)
    eng.initialize()
    eng.run(0, nstep - 1)
    eng.update_wavefields(nstep - 1)
    eng.apply_corrector_phase()
    eng.compute_seismograms(nstep - 1)
    ours = eng.seismograms().astype(np.float64)
    fields = [eng.field(k).astype(np.float64) for k in TYPES]
    eng.close()
    ref = {}
    for prec in (4, 8):
        s, _, f = oracle.Oracle3D.from_database(db, prec, dt).forward(src, rec, nstep, dt, 1, TYPES,
                                                                      return_fields=True)
        ref[prec] = (s, f)
    for it, t in enumerate(TYPES):
        a, b32, b64 = ours[:, it], ref[4][0][:, it], ref[8][0][:, it]
        e32, e64, floor = rel_l2(a, b32), rel_l2(a, b64), rel_l2(b32, b64)
        _report(f"db3d_{scatter}/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        if t == "displacement":
            assert e32 <= 1e-5, e32
        else:
            _check(e32, e64, floor)
        a, b32, b64 = fields[it], ref[4][1][it], ref[8][1][it]
        e32, e64, floor = rel_l2(a, b32), rel_l2(a, b64), rel_l2(b32, b64)
        _report(f"db3d_{scatter}/field_{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)


def test_database_cube_equals_structured_engine():
    from specfempp_amd import Engine3D, Engine3DDatabase, meshes

    n, h, dt, nstep = 6, 100.0, 1e-3, 300
    mat = np.zeros((n, n, n), np.int32)
    mat[n // 2:] = 1
    args = (n, n, n, h, mat, [2500.0, 2700.0], [3000.0, 3600.0], [1730.0, 2000.0])
    mesh = meshes.hex_mesh(*args)
    db = meshes.hex_database(*args, absorbing=mesh["absorbing"], perturb=0.0, dt=dt)
    w = meshes.config3(n=n, nstep=nstep, h=h, nrec_side=3)
    print(f"[case] {n}^3 cube as a database vs the structured engine, {nstep} steps", flush=True)
    runs = []
    for eng in (Engine3D(mesh, w.sources, w.receivers, dt, nstep, 1, TYPES, scatter="gather"),
                Engine3DDatabase(db, w.sources, w.receivers, dt, nstep, 1, TYPES, scatter="colored")):
        eng.initialize()
        eng.run()
        runs.append(eng.seismograms().astype(np.float64))
        eng.close()
    o32, _ = oracle.Oracle3D(mesh, 4, dt).forward(w.sources, w.receivers, nstep, dt, 1, TYPES)
    o64, _ = oracle.Oracle3D(mesh, 8, dt).forward(w.sources, w.receivers, nstep, dt, 1, TYPES)
    for it, t in enumerate(TYPES):
        e = rel_l2(runs[1][:, it], runs[0][:, it])
        floor = rel_l2(o32[:, it], o64[:, it])
        _report(f"db3d_cube_vs_structured/{t}", rel_l2=e, fp32_floor=floor)
        assert e <= (1e-5 if t == "displacement" else max(1e-5, 2.0 * floor))
