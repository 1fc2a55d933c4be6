"""GPU edge cases of the 3-D and 2-D forward paths, through the C ABI.

- A non-cubic 8 x 6 x 4 box with two seeded materials, Stacey on all six
  faces, a force and a moment-tensor source (with a time shift), receivers on
  element faces, edges and corners, seismograms every third step: each 3-D
  scatter mode against the FP32 and FP64 oracle.
- One sfx_step call per step gives the bits of one sfx_run call (3-D gather).
- No sources and no receivers: the run is valid, the wavefield stays exactly
  zero and the seismogram array is empty (2-D and 3-D).
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from test_gpu_parity import TOL_SEISMO, _box_workload, _check, _report

pytestmark = pytest.mark.gpu

TYPES = ("displacement", "velocity", "acceleration")


@pytest.mark.parametrize("scatter", ["gather", "colored", "atomic"])
def test_3d_box_all_absorbing_matches_oracle(scatter):
    from specfempp_amd import Engine3D

    mesh, src, rec, dt, nstep = _box_workload()
    nsb = 3
    o32 = oracle.Oracle3D(mesh, 4, dt)
    s32, _ = o32.forward(src, rec, nstep, dt, nsb, TYPES)
    s64, _ = oracle.Oracle3D(mesh, 8, dt).forward(src, rec, nstep, dt, nsb, TYPES)
    print(f"[case] 3d box 8x6x4 all-absorbing scatter={scatter} nstep={nstep} nsb={nsb}", flush=True)
    eng = Engine3D(mesh, src, rec, dt, nstep, nsb, TYPES, scatter=scatter)
    assert np.array_equal(eng.index_mapping(), o32.ibool())
    assert rel_l2(eng.field("mass_inverse"), o32.mass_inverse()) < 1e-6
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    eng.close()
    assert ours.shape == s32.shape == ((nstep + nsb - 1) // nsb, len(TYPES), len(rec), 3)
    for it, t in enumerate(TYPES):
        e32, e64 = rel_l2(ours[:, it], s32[:, it]), rel_l2(ours[:, it], s64[:, it])
        floor = rel_l2(s32[:, it], s64[:, it])
        _report(f"3d_box_{scatter}/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)
        if t == "displacement":
            assert e32 <= TOL_SEISMO


def test_3d_step_calls_equal_run_bitwise():
    from specfempp_amd import Engine3D

    mesh, src, rec, dt, _ = _box_workload()
    nstep = 150
    print("[case] 3d box gather step() vs run()", flush=True)
    eng = Engine3D(mesh, src, rec, dt, nstep, 1, TYPES, scatter="gather")
    eng.initialize()
    eng.run()
    run = eng.seismograms()
    fields_run = [eng.field(k) for k in TYPES]
    eng.initialize()
    for istep in range(nstep):
        eng.step(istep)
    stepped = eng.seismograms()
    fields_step = [eng.field(k) for k in TYPES]
    eng.close()
    assert np.abs(run).max() > 0
    assert np.array_equal(run, stepped)
    for a, b in zip(fields_run, fields_step):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dim,scatter", [(3, "gather"), (3, "colored"), (2, "gather"), (2, "colored")])
def test_no_sources_no_receivers(dim, scatter):
    from specfempp_amd import Engine2D, Engine3D, meshes

    print(f"[case] empty sources/receivers dim={dim} scatter={scatter}", flush=True)
    if dim == 3:
        mesh = meshes.hex_mesh(3, 2, 4, 100.0, np.zeros((4, 2, 3), dtype=np.int32), [2500.0], [3000.0], [1730.0])
        eng = Engine3D(mesh, [], [], 1e-3, 20, 1, TYPES, scatter=scatter)
    else:
        mesh = meshes.quad_mesh(5, 4, 0.0, 250.0, 0.0, 200.0, materials=[meshes.elastic(2200.0, 2500.0, 1443.0)])
        eng = Engine2D(mesh, [], [], 1e-3, 20, 1, TYPES, scatter=scatter)
    eng.initialize()
    eng.run()
    s = eng.seismograms()
    fields = [eng.field(k) for k in TYPES]
    eng.close()
    assert s.shape == (20, len(TYPES), 0, dim)
    for f in fields:
        assert f.shape[1] > 0 and not np.any(f)


def test_3d_source_in_absorbing_element_matches_oracle():
    """A source in an element with an absorbing face: the gather path then adds
    sources and Stacey terms in two launches (one launch otherwise)."""
    from specfempp_amd import Engine3D

    mesh, src, rec, dt, _ = _box_workload()
    nstep = 300
    src = [dict(src[0], x=40.0, y=270.0, z=190.0), src[1]]  # first source in an x- face element
    print("[case] 3d box gather, source in an absorbing element", flush=True)
    s32, _ = oracle.Oracle3D(mesh, 4, dt).forward(src, rec, nstep, dt, 1, TYPES)
    s64, _ = oracle.Oracle3D(mesh, 8, dt).forward(src, rec, nstep, dt, 1, TYPES)
    eng = Engine3D(mesh, src, rec, dt, nstep, 1, TYPES, scatter="gather")
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    eng.close()
    for it, t in enumerate(TYPES):
        e32, e64 = rel_l2(ours[:, it], s32[:, it]), rel_l2(ours[:, it], s64[:, it])
        floor = rel_l2(s32[:, it], s64[:, it])
        _report(f"3d_box_src_on_boundary/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)
        if t == "displacement":
            assert e32 <= TOL_SEISMO


def test_3d_displacement_only_traces_equal_those_of_a_full_run():
    """Displacement traces alone skip the force assembly at the receivers; they
    equal the displacement column of a run that also records v and a."""
    from specfempp_amd import Engine3D

    mesh, src, rec, dt, _ = _box_workload()
    nstep = 120
    print("[case] 3d box gather, displacement-only traces", flush=True)
    runs = []
    for types in (("displacement",), TYPES):
        eng = Engine3D(mesh, src, rec, dt, nstep, 1, types, scatter="gather")
        eng.initialize()
        eng.run()
        runs.append(eng.seismograms())
        eng.close()
    assert np.abs(runs[0]).max() > 0
    assert np.array_equal(runs[0][:, 0], runs[1][:, 0])


@pytest.mark.parametrize("nsrc,nrec", [(2, 0), (0, 8), (0, 0)])
def test_3d_displacement_sampling_launch_with_missing_sources_or_receivers(nsrc, nrec):
    """Displacement traces are taken in the Stacey/sources launch: without
    receivers the wavefield equals that of a run with velocity traces (which
    takes the separate launches), and without sources everything stays zero."""
    from specfempp_amd import Engine3D

    mesh, src, rec, dt, _ = _box_workload()
    src, rec = src[:nsrc], rec[:nrec]
    nstep = 60
    print(f"[case] 3d box gather, displacement sampling launch, nsrc={nsrc} nrec={nrec}", flush=True)
    out = []
    for types in (("displacement",), ("displacement", "velocity")):
        eng = Engine3D(mesh, src, rec, dt, nstep, 1, types, scatter="gather")
        eng.initialize()
        eng.run()
        out.append(([eng.field(k) for k in TYPES], eng.seismograms()))
        eng.close()
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert out[0][1].shape == (nstep, 1, nrec, 3)
    assert np.array_equal(out[0][1][:, 0], out[1][1][:, 0])
    if nsrc == 0:
        assert not np.any(out[0][0][0]) and not np.any(out[0][1])
    else:
        assert np.abs(out[0][0][0]).max() > 0
