"""GPU parity: the HIP engine (through the C ABI) against the oracle.

Bars: numbering, per-medium numbering, compute ordering and source/receiver
element location bit-exact; FP32 seismograms within relative L2 <= 1e-5 of
the FP32 oracle (each test also prints its error against the FP64 oracle);
SPECFEM2D traces within the reference's own rel. L1 <= 1e-3.

Acceleration seismograms sit on an FP32 rounding floor above 1e-5: the FP32
oracle itself moves by ~3e-5 (rel. L2) when only its summation order changes
and is ~5e-5 from the FP64 oracle.  For a trace type on that floor the bar is
that the engine's distance to the FP64 oracle stays within 1.5x the FP32
oracle's own distance to it (the order-only variation above is ~0.6x of that
distance, so 1.5x leaves room for it and nothing more) (``_check``).
"""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_case, rel_l1_per_station, rel_l2

pytestmark = pytest.mark.gpu

# The 3-D block path is switched off on the device.  The cause of its one GPU
# memory fault was found on the host build and fixed (an over-read of the
# kernel's quadrature-row argument by its idle threads; the ASan + bounds build
# reports it before the fix and is clean after it, at every shape these tests
# use), but the fixed kernel has not been run on a GPU, so these tests stay
# skipped.
BLOCK_OFF = pytest.mark.skip(reason="3-D block path switched off on the device: fault cause fixed on the host "
                                    "build, the fixed kernel not yet run on a GPU")

TOL_SEISMO = 1e-5


def _engine2d(mesh, cfg, types=("displacement",), nstep=None, scatter="colored"):
    from specfempp_amd import Engine2D

    return Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], nstep or cfg["nstep"], 1, types,
                    scatter=scatter)


def _check(e32, e64, floor):
    """e32/e64: engine vs FP32/FP64 oracle; floor: FP32 oracle vs FP64 oracle."""
    assert e32 <= TOL_SEISMO or e64 <= 1.5 * floor, (e32, e64, floor)


def _report(name, **vals):
    print(f"[parity] {name}: " + ", ".join(f"{k}={v:.3e}" for k, v in vals.items()), flush=True)


@pytest.mark.parametrize("case", ["elastic_free_80x60", "elastic_stacey_80x60"])
def test_numbering_and_locations_bit_exact(case):
    mesh, cfg = load_case(case)
    eng = _engine2d(mesh, cfg)
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


@pytest.mark.parametrize("scatter", ["colored", "gather"])
@pytest.mark.parametrize("case", ["elastic_free_80x60", "elastic_stacey_80x60"])
def test_forward_seismograms_match_oracle_and_specfem2d(case, scatter):
    mesh, cfg = load_case(case)
    kinds = cfg["trace_kinds"]
    types = [{"d": "displacement", "v": "velocity", "a": "acceleration"}[k] for k in kinds]
    eng = _engine2d(mesh, cfg, types, scatter=scatter)
    if scatter == "gather":
        # 9-node elements with curved-capable geometry: recomputed per point
        # from the control nodes; one material: kappa, mu once per element
        lay = eng.layout()
        assert lay["geometry"] == "control_nodes_9" and lay["element_properties"], lay
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)  # (nsig, ntypes, nrec, 2)
    mass = eng.field("mass_inverse")
    eng.close()
    o32 = oracle.Oracle2D(mesh, 4, 1, cfg["dt"])
    o64 = oracle.Oracle2D(mesh, 8, 1, cfg["dt"])
    s32, _ = o32.forward(cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"], 1, types)
    s64, _ = o64.forward(cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"], 1, types)
    assert rel_l2(mass, o32.mass_inverse()) < 1e-6
    traces = np.load(os.path.join(GOLDEN, case, "traces.npz"))
    for it, k in enumerate(kinds):
        e32 = rel_l2(ours[:, it], s32[:, it])
        e64 = rel_l2(ours[:, it], s64[:, it])
        ref = rel_l1_per_station(ours[:, it].transpose(1, 2, 0), traces[k]).max()
        floor = rel_l2(s32[:, it], s64[:, it])
        _report(f"{case}/{scatter}/{k}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor,
                vs_specfem2d_l1=ref)
        _check(e32, e64, floor)
        if k == "d":
            assert e32 <= TOL_SEISMO
        if scatter == "gather":
            # the gather path sums each point's contributions in the oracle's
            # element order and rounds every product and sum as the oracle
            # (built without contraction) does: the same bits
            assert e32 == 0.0, (k, e32)
        assert ref <= cfg["tolerance"]


@pytest.mark.parametrize("scatter", ["colored", "gather"])
def test_non_affine_mesh_matches_oracle(scatter):
    """Interior nodes moved by up to a quarter element: the geometry varies over
    each element, so the colored stiffness reads it per point (affine meshes use
    one record per element) and the gather path recomputes it at every point
    from the element's 4 control nodes, giving the oracle's bits."""
    from specfempp_amd import Engine2D, meshes

    n, h = 24, 10.0
    mesh = meshes.quad_mesh(n, n, 0.0, n * h, 0.0, n * h, materials=[meshes.elastic(2000.0, 3000.0, 1700.0)],
                            absorbing=("bottom", "right", "left"))
    rng = np.random.default_rng(3)
    c = mesh.coorg.copy()
    iz, ix = np.divmod(np.arange(c.shape[1]), n + 1)
    inner = (ix > 0) & (ix < n) & (iz > 0) & (iz < n)
    c[:, inner] += rng.uniform(-0.25 * h, 0.25 * h, size=(2, inner.sum()))
    mesh.coorg = c
    src = [dict(x=117.3, z=131.9, angle=20.0, f0=60.0, factor=1e10, tshift=0.0)]
    rec = [dict(x=x, z=205.0) for x in (30.5, 90.2, 150.7, 210.1)]
    dt, nstep = meshes.stable_dt(0.75 * h, 3000.0), 700
    e = Engine2D(mesh, src, rec, dt, nstep, 1, ("displacement",), scatter=scatter)
    lay = e.layout()
    e.initialize()
    e.run()
    ours = e.seismograms().astype(np.float64)
    e.close()
    s32, _ = oracle.Oracle2D(mesh, 4, 1, dt).forward(src, rec, nstep, dt, 1, ("displacement",))
    s64, _ = oracle.Oracle2D(mesh, 8, 1, dt).forward(src, rec, nstep, dt, 1, ("displacement",))
    e32, e64, floor = rel_l2(ours, s32), rel_l2(ours, s64), rel_l2(s32, s64)
    _report(f"non_affine_24x24/{scatter}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
    assert e32 <= TOL_SEISMO
    if scatter == "gather":
        # the one-launch batch step reads the stored per-point geometry and
        # kappa, mu; the engine recomputes the geometry from the control nodes
        # and reads one (kappa, mu) per element: the same bits
        from specfempp_amd import Batch

        assert lay["geometry"] == "control_nodes_4" and lay["element_properties"], lay
        host = Engine2D(mesh, src, rec, dt, nstep, 1, ("displacement",), host_only=True)
        b = Batch([host])
        host.close()
        b.set_config(5, True, "gather")
        b.run()
        assert np.array_equal(b.seismograms(0), ours.astype(np.float32))
        b.close()


def test_gather_properties_per_point_and_per_element():
    """sfx_set_properties on the 2-D gather path: values that vary within an
    element make the element forces read kappa and mu per point, values
    constant over each element one pair per element; both runs match the FP32
    oracle with the same properties."""
    from specfempp_amd import Engine2D, meshes

    n, h = 16, 20.0
    mesh = meshes.quad_mesh(n, n, 0.0, n * h, 0.0, n * h, materials=[meshes.elastic(2000.0, 3000.0, 1700.0)],
                            absorbing=("bottom", "right", "left"))
    src = [dict(x=150.3, z=170.9, angle=10.0, f0=40.0, factor=1e10, tshift=0.0)]
    rec = [dict(x=x, z=250.0) for x in (40.5, 120.2, 200.7)]
    dt, nstep = meshes.stable_dt(h, 3300.0), 400
    rho0, kap0, mu0, _ = oracle.Oracle2D(mesh, 4, 1, dt).properties()
    rng = np.random.default_rng(5)
    per_point = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, size=rho0.shape)
    per_elem = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, size=(mesh.nspec, 1, 1)) * np.ones(rho0.shape)
    for scale, element_props in ((per_point, False), (per_elem, True)):
        props = [(np.asarray(x, dtype=np.float64) * scale).astype(np.float32) for x in (rho0, kap0, mu0)]
        e = Engine2D(mesh, src, rec, dt, nstep, 1, ("displacement",), scatter="gather")
        e.set_properties(*props)
        lay = e.layout()
        e.initialize()
        e.run()
        ours = e.seismograms().astype(np.float64)
        e.close()
        o = oracle.Oracle2D(mesh, 4, 1, dt)
        o.set_properties(*props)
        s32, _ = o.forward(src, rec, nstep, dt, 1, ("displacement",))
        _report(f"gather_properties/{'element' if element_props else 'point'}", vs_oracle_fp32=rel_l2(ours, s32))
        assert lay["element_properties"] == element_props, lay
        assert np.abs(s32).max() > 0
        assert rel_l2(ours, s32) <= TOL_SEISMO


@pytest.mark.parametrize("scatter", ["colored", "gather"])
def test_seam_path_matches_fused_path(scatter):
    """predictor / update_wavefields / corrector / compute_seismograms, called
    the way time_marching.tpp:33-84 calls them, equals the fused step (the
    gather path sums in a fixed order, so there the two agree bit for bit)."""
    mesh, cfg = load_case("elastic_stacey_80x60")
    nstep = 120
    eng = _engine2d(mesh, cfg, ("displacement", "velocity", "acceleration"), nstep=nstep, scatter=scatter)
    eng.initialize()
    total_dofs = 2 * eng.nglob
    for istep in range(nstep):
        assert eng.apply_predictor_phase() == total_dofs
        assert eng.update_wavefields(istep) == eng.nspec
        assert eng.apply_corrector_phase() == total_dofs
        eng.compute_seismograms(istep)
    seam = eng.seismograms().astype(np.float64)
    eng.initialize()
    eng.run()
    fused = eng.seismograms().astype(np.float64)
    eng.close()
    # The two runs differ only by the (nondeterministic) order of the atomic
    # adds; the FP32 oracle moves by 2.4e-7 / 2.4e-6 / 2.8e-5 (d / v / a) when
    # its summation order changes, so the bars are set a few times above that.
    for it, bar in enumerate((2e-6, 2e-5, 2e-4)):
        err = rel_l2(seam[:, it], fused[:, it])
        _report(f"seam_vs_fused/{scatter}/{it}", rel_l2=err)
        assert err <= (0.0 if scatter == "gather" else bar)


@pytest.mark.parametrize("scatter", ["colored", "gather"])
def test_config1_homogeneous_80x80(scatter):
    """The 2-D CPU-runnable configuration (80x80, free top, Stacey on three
    sides), all 1600 steps (before ~step 700 the receivers see only the
    numerical precursor, so shorter runs measure rounding noise only)."""
    from specfempp_amd import Engine2D, meshes

    w = meshes.config1(nstep=1600)
    eng = Engine2D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, ("displacement",), scatter=scatter)
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    eng.close()
    s32, _ = oracle.Oracle2D(w.mesh, 4, 1, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1, ["displacement"])
    s64, _ = oracle.Oracle2D(w.mesh, 8, 1, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1, ["displacement"])
    e32, e64 = rel_l2(ours, s32), rel_l2(ours, s64)
    _report(f"config1_1600/{scatter}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=rel_l2(s32, s64))
    assert e32 <= TOL_SEISMO


# The Ricker wavelet (f0 = 5 Hz) peaks at step 240 (t0 = -1.2 / f0); windows
# reach well past it so the traces are signal, not rounding noise.
@pytest.mark.parametrize("n,nstep,scatter", [pytest.param(6, 600, "block", marks=BLOCK_OFF),
                                            pytest.param(10, 500, "block", marks=BLOCK_OFF), (6, 600, "gather"),
                                            (10, 500, "gather"), (6, 600, "colored"), (6, 600, "atomic")])
def test_3d_layered_matches_oracle(n, nstep, scatter):
    from specfempp_amd import Engine3D, meshes

    w = meshes.config3(n=n, nstep=nstep, h=100.0, nrec_side=4)
    types = ("displacement", "velocity", "acceleration")
    eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, types, scatter=scatter)
    o32 = oracle.Oracle3D(w.mesh, 4, w.dt)
    assert np.array_equal(eng.index_mapping(), o32.ibool())
    assert rel_l2(eng.field("mass_inverse"), o32.mass_inverse()) < 1e-6
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    eng.close()
    s32, _ = o32.forward(w.sources, w.receivers, w.nstep, w.dt, 1, types)
    s64, _ = oracle.Oracle3D(w.mesh, 8, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1, types)
    for it, t in enumerate(types):
        e32, e64 = rel_l2(ours[:, it], s32[:, it]), rel_l2(ours[:, it], s64[:, it])
        floor = rel_l2(s32[:, it], s64[:, it])
        _report(f"3d_n{n}_{scatter}/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)
        if t == "displacement":
            assert e32 <= TOL_SEISMO


@pytest.mark.parametrize("n,scatter", [(5, "gather"), pytest.param(6, "block", marks=BLOCK_OFF)])
def test_3d_seam_path_and_fields_match_fused_path(n, scatter):
    """Seam calls (predictor / update_wavefields / corrector / seismograms) on
    the 3-D gather and block paths equal the fused step, and the wavefield read
    back after fused steps (including the acceleration) equals the seam path's;
    both paths sum in a fixed order, so two fused runs agree bit for bit."""
    from specfempp_amd import Engine3D, meshes

    w = meshes.config3(n=n, nstep=300, h=100.0, nrec_side=3)
    types = ("displacement", "velocity", "acceleration")
    eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, types, scatter=scatter)
    eng.initialize()
    for istep in range(w.nstep):
        eng.apply_predictor_phase()
        assert eng.update_wavefields(istep) == eng.nspec
        eng.apply_corrector_phase()
        eng.compute_seismograms(istep)
    seam = eng.seismograms().astype(np.float64)
    seam_fields = [eng.field(k) for k in ("displacement", "velocity", "acceleration")]
    eng.initialize()
    eng.run()
    fused = eng.seismograms().astype(np.float64)
    # after the fused loop the stored u, v hold the predictor of step nstep;
    # undo it to compare with the seam state (a is the assembled acceleration)
    a = eng.field("acceleration").astype(np.float64)
    v = eng.field("velocity").astype(np.float64) - w.dt / 2 * a
    eng.initialize()
    eng.run()
    fused2 = eng.seismograms()
    eng.close()
    assert np.array_equal(fused2, fused.astype(np.float32))
    for it, bar in enumerate((2e-6, 2e-5, 2e-4)):
        err = rel_l2(seam[:, it], fused[:, it])
        _report(f"3d_seam_vs_fused/{it}", rel_l2=err)
        assert err <= bar
    assert rel_l2(a, seam_fields[2]) <= 2e-4
    assert rel_l2(v, seam_fields[1]) <= 2e-5


def _box_workload():
    """8 x 6 x 4 elements (2-element blocks along every axis), two seeded
    materials, Stacey on all six faces, a force and a moment-tensor source, and
    receivers spread through the volume: on block faces, edges and corners and
    strictly inside blocks."""
    from specfempp_amd import meshes

    rng = np.random.default_rng(11)
    nx, ny, nz, h = 8, 6, 4, 100.0
    material = rng.integers(0, 2, size=(nz, ny, nx)).astype(np.int32)
    mesh = meshes.hex_mesh(nx, ny, nz, h, material, [2000.0, 2500.0], [3000.0, 4200.0], [1700.0, 2400.0],
                           absorbing=(1, 1, 1, 1, 1, 1))
    M = meshes.moment_tensor(5)
    src = [dict(type="force", x=330.0, y=270.0, z=190.0, fx=0.3, fy=-0.5, fz=0.8, f0=12.0, factor=1e10),
           dict(type="moment_tensor", x=610.0, y=420.0, z=230.0, mxx=M[0, 0], myy=M[1, 1], mzz=M[2, 2],
                mxy=M[0, 1], mxz=M[0, 2], myz=M[1, 2], f0=12.0, factor=1e12, tshift=0.01)]
    pts = [(400.0, 300.0, 200.0), (200.0, 200.0, 200.0), (150.0, 250.0, 120.0), (575.0, 125.0, 330.0),
           (400.0, 440.0, 60.0), (720.0, 530.0, 390.0), (50.0, 50.0, 10.0), (800.0, 600.0, 400.0)]
    rec = [dict(x=x, y=y, z=z) for x, y, z in pts]
    return mesh, src, rec, 1e-3, 400


@BLOCK_OFF
def test_3d_block_box_all_absorbing_matches_oracle():
    from specfempp_amd import Engine3D

    mesh, src, rec, dt, nstep = _box_workload()
    types = ("displacement", "velocity", "acceleration")
    o32 = oracle.Oracle3D(mesh, 4, dt)
    s32, _ = o32.forward(src, rec, nstep, dt, 1, types)
    s64, _ = oracle.Oracle3D(mesh, 8, dt).forward(src, rec, nstep, dt, 1, types)
    runs = {}
    for scatter in ("block", "gather"):
        eng = Engine3D(mesh, src, rec, dt, nstep, 1, types, scatter=scatter)
        assert np.array_equal(eng.index_mapping(), o32.ibool())
        eng.initialize()
        eng.run()
        runs[scatter] = eng.seismograms().astype(np.float64)
        eng.close()
    for it, t in enumerate(types):
        ours = runs["block"][:, it]
        e32, e64 = rel_l2(ours, s32[:, it]), rel_l2(ours, s64[:, it])
        floor = rel_l2(s32[:, it], s64[:, it])
        _report(f"3d_box_block/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor,
                vs_gather=rel_l2(ours, runs["gather"][:, it]))
        _check(e32, e64, floor)
        if t == "displacement":
            assert e32 <= TOL_SEISMO


@BLOCK_OFF
def test_3d_block_step_calls_equal_run_bitwise():
    """Stepping one call per step stores the full acceleration after every
    step; the displacement and velocity updates are the same instructions, so
    the seismograms equal those of one run() call bit for bit."""
    from specfempp_amd import Engine3D

    mesh, src, rec, dt, _ = _box_workload()
    nstep = 150
    types = ("displacement", "velocity", "acceleration")
    eng = Engine3D(mesh, src, rec, dt, nstep, 1, types, scatter="block")
    eng.initialize()
    eng.run()
    run = eng.seismograms()
    a_run = eng.field("acceleration")
    eng.initialize()
    for istep in range(nstep):
        eng.step(istep)
    stepped = eng.seismograms()
    a_step = eng.field("acceleration")
    eng.close()
    assert np.array_equal(run, stepped)
    assert np.array_equal(a_run, a_step)


def test_pressure_seismograms_match_oracle():
    """Pressure (wavefield.hpp:49-85) beside displacement on the golden
    Stacey case, and on a prescribed linear field, where the exact value is
    -kappa (a + d) at every receiver."""
    mesh, cfg = load_case("elastic_stacey_80x60")
    types = ("displacement", "pressure")
    eng = _engine2d(mesh, cfg, types)
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    s32, _ = oracle.Oracle2D(mesh, 4, 1, cfg["dt"]).forward(cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"],
                                                            1, types)
    s64, _ = oracle.Oracle2D(mesh, 8, 1, cfg["dt"]).forward(cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"],
                                                            1, types)
    for it, t in enumerate(types):
        e32, e64 = rel_l2(ours[:, it], s32[:, it]), rel_l2(ours[:, it], s64[:, it])
        floor = rel_l2(s32[:, it], s64[:, it])
        _report(f"pressure_case/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)
    assert np.all(ours[:, 1, :, 1] == 0.0)
    # linear field through the seam sampling
    o = oracle.Oracle2D(mesh, 8, 1, cfg["dt"])
    coord, _ = o.geometry()
    ibool, midx, _ = o.numbering()
    xg, zg = np.zeros(o.nglob), np.zeros(o.nglob)
    xg[midx[ibool.ravel()]] = coord[..., 0].ravel()
    zg[midx[ibool.ravel()]] = coord[..., 1].ravel()
    a, b, c, d = 2e-4, -3e-4, 5e-4, 7e-4
    eng.set_field("displacement", np.stack([a * xg + b * zg, c * xg + d * zg]))
    eng.compute_seismograms(0)
    got = eng.seismograms()[0, 1, :, 0].astype(np.float64)
    m = mesh.materials[0]
    kappa = m.density * (m.cp ** 2 - 4.0 / 3.0 * m.cs ** 2)
    exact = -kappa * (a + d)
    # FP32 rounding of the derivatives of a field of magnitude ~3 over 50 m
    # elements: the FP32 oracle's own error on the same field sets the bar
    lin = np.zeros((3, 2, o.nglob))
    lin[0, 0], lin[0, 1] = a * xg + b * zg, c * xg + d * zg
    o32 = oracle.Oracle2D(mesh, 4, 1, cfg["dt"]).sample(cfg["stations"], lin, ("pressure",))[0, :, 0]
    err, floor = np.max(np.abs(got - exact)) / abs(exact), np.max(np.abs(o32 - exact)) / abs(exact)
    _report("pressure_linear_field", max_rel=err, fp32_oracle_max_rel=floor)
    assert err <= max(1e-4, 2.0 * floor)
    eng.close()


@pytest.mark.parametrize("scatter", ["colored", "gather"])
def test_2d_moment_tensor_source_matches_oracle(scatter):
    """A 2-D moment-tensor source (moment_tensor_source.cpp:18-140: M . grad
    of the Lagrange interpolants, element-averaged derivatives of the
    reference coordinates) next to a force source, with a time shift, on a
    non-square mesh with Stacey on three sides."""
    from specfempp_amd import Engine2D, meshes

    nx, nz, h = 30, 24, 20.0
    mesh = meshes.quad_mesh(nx, nz, 0.0, nx * h, 0.0, nz * h, materials=[meshes.elastic(2300.0, 2800.0, 1600.0)],
                            absorbing=("bottom", "right", "left"))
    src = [dict(type="moment_tensor", x=217.3, z=251.9, mxx=0.8, mzz=-0.45, mxz=0.35, f0=25.0, factor=1e12,
                tshift=0.004),
           dict(x=401.0, z=150.0, angle=35.0, f0=25.0, factor=1e10, tshift=0.0)]
    rec = [dict(x=x, z=z) for x, z in ((80.5, 400.2), (300.0, 420.0), (517.3, 90.1), (250.0, 250.0))]
    dt, nstep = meshes.stable_dt(h, 2800.0), 900
    print(f"[case] 2-D moment tensor + force, {nx}x{nz}, {nstep} steps, scatter={scatter}", flush=True)
    e = Engine2D(mesh, src, rec, dt, nstep, 1, ("displacement", "velocity"), scatter=scatter)
    e.initialize()
    e.run()
    ours = e.seismograms().astype(np.float64)
    e.close()
    s32, _ = oracle.Oracle2D(mesh, 4, 1, dt).forward(src, rec, nstep, dt, 1, ("displacement", "velocity"))
    s64, _ = oracle.Oracle2D(mesh, 8, 1, dt).forward(src, rec, nstep, dt, 1, ("displacement", "velocity"))
    for it, t in enumerate(("displacement", "velocity")):
        e32, e64, floor = rel_l2(ours[:, it], s32[:, it]), rel_l2(ours[:, it], s64[:, it]), rel_l2(s32[:, it],
                                                                                                 s64[:, it])
        _report(f"2d_moment_tensor_{scatter}/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        assert e32 <= TOL_SEISMO


@pytest.mark.parametrize("chunk", [2, 64])
def test_storage_chunk_numbering_and_seismograms_match_oracle(chunk):
    """Numbering at storage_chunk_size > 1 (chunk_config.hpp:12-28: 64 on HIP
    builds, 32 on CUDA): index mapping, per-medium numbering and locations
    bit-exact against the oracle run with the same chunk, and the forward run
    (whose field layout follows that numbering) within the seismogram bar."""
    from specfempp_amd import Engine2D

    mesh, cfg = load_case("elastic_stacey_80x60")
    print(f"[case] elastic_stacey_80x60 chunk={chunk}", flush=True)
    eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], cfg["nstep"], 1, ("displacement",),
                   chunk_size=chunk)
    orc = oracle.Oracle2D(mesh, 4, chunk, cfg["dt"])
    ibool, midx, c2m = orc.numbering()
    assert np.array_equal(eng.index_mapping(), ibool)
    assert np.array_equal(eng.field_index(), midx)
    assert np.array_equal(eng.compute_to_mesh(), c2m)
    ispec, _ = eng.locations("receivers")
    assert [orc.locate(p["x"], p["z"])[0] for p in cfg["stations"]] == list(ispec)
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    eng.close()
    s32, _ = orc.forward(cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"], 1, ("displacement",))
    e32 = rel_l2(ours, s32)
    _report(f"chunk{chunk}", vs_oracle_fp32=e32)
    assert e32 <= TOL_SEISMO
