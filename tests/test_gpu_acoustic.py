"""GPU parity of the acoustic medium and the elastic-acoustic coupling (2-D),
through the C ABI, on the reference's own Newmark fixtures
(tests/unit-tests/displacement_tests/Newmark/serial/*):

- acoustic with a free surface on all sides (Dirac source; displacement,
  velocity, acceleration and pressure traces), with a free surface and
  Dirichlet sides, with Stacey and a free surface (composite elements), with
  Stacey only;
- both coupled meshes (an acoustic layer above or below an elastic one).

Bars:
- seismograms within 1e-5 relative L2 of the FP32 oracle, or, where FP32
  rounding alone moves the result further (every acoustic trace is a
  derivative of the potential: the FP32 oracle is up to 1.6e-5 from the FP64
  one on these fixtures), within 1.5x the FP32 oracle's own distance to the
  FP64 oracle (the rule of the elastic velocity and acceleration tests);
- every trace within the reference's own 1e-3 relative L1 per station of the
  SPECFEM2D traces (newmark_tests.cpp:325-346);
- the acoustic mass inverse within 1e-6 of the oracle;
- final wavefields of both media, on the same rule;
- the coupled loop driven seam by seam (acoustic predictor, update,
  corrector, then elastic) equals the fused run.
"""
import os

import numpy as np
import pytest

import _oracle_jobs as jobs
import oracle
from conftest import GOLDEN, load_case, rel_l1_per_station, rel_l2
from test_gpu_parity import _check, _report

pytestmark = pytest.mark.gpu

TYPES = {"d": "displacement", "v": "velocity", "a": "acceleration", "p": "pressure"}
CASES = ["acoustic_free_80x60", "acoustic_dirichlet_80x60", "acoustic_composite_80x60", "acoustic_stacey_80x60",
         "coupled_acoustic_top", "coupled_acoustic_bottom"]


def _engine(mesh, cfg, types, scatter, nstep=None):
    from specfempp_amd import Engine2D

    return Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], nstep or cfg["nstep"], 1, types,
                    scatter=scatter)


@pytest.mark.parametrize("scatter", ["colored", "atomic"])
@pytest.mark.parametrize("case", CASES)
def test_acoustic_seismograms_match_oracle_and_specfem2d(case, scatter):
    mesh, cfg = load_case(case)
    types = [TYPES[k] for k in cfg["trace_kinds"]]
    print(f"[case] {case} scatter={scatter} nstep={cfg['nstep']} types={types}", flush=True)
    with jobs.pool(2) as ex:
        futs = [ex.submit(jobs.forward_2d, mesh, prec, cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"],
                          tuple(types)) for prec in (4, 8)]
        eng = _engine(mesh, cfg, types, scatter)
        assert eng.nspec_acoustic > 0
        eng.initialize()
        eng.run()
        ours = eng.seismograms().astype(np.float64)
        # acoustic mass inverse against the oracle's
        minv = eng.acoustic_field("mass_inverse").astype(np.float64)
        _, _, ref_minv = oracle.Oracle2D(mesh, 4, 1, cfg["dt"]).acoustic()
        e_m = float(np.abs(minv - ref_minv).max() / np.abs(ref_minv).max())
        _report(f"{case}/mass_inverse_acoustic", rel_max=e_m)
        assert e_m <= 1e-6
        eng.close()
        f32, f64 = [f.result() for f in futs]
    traces = np.load(os.path.join(GOLDEN, case, "traces.npz"))
    for it, k in enumerate(cfg["trace_kinds"]):
        a, b32, b64 = ours[:, it], f32[:, it], f64[:, it]
        e32, e64, floor = rel_l2(a, b32), rel_l2(a, b64), rel_l2(b32, b64)
        _report(f"{case}/{scatter}/{k}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        # in an acoustic element every trace is a derivative of the potential
        # (grad chi / rho, or -chi_dot_dot): the FP32 oracle's own distance to
        # the FP64 one reaches 1.6e-5 on the acoustic fixtures, so the
        # rounding-floor rule applies to all of them
        _check(e32, e64, floor)
        mine = ours[:, it].transpose(1, 2, 0)
        if k == "p":
            mine = mine[:, :1]
        l1 = rel_l1_per_station(mine, traces[k])
        _report(f"{case}/{scatter}/{k}/specfem2d", rel_l1_max=float(l1.max()))
        assert l1.max() < cfg["tolerance"], (k, l1.max())


def _coupled_snapshot(eng, k):
    """Corrected state after k steps: k - 1 fused steps, then step k - 1
    through the seams in the reference's order (time_marching.tpp:33-84)."""
    eng.initialize()
    eng.run(0, k - 1)
    eng.update_wavefields_medium("acoustic", "forward", k - 1)
    eng.corrector_phase("acoustic")
    eng.update_wavefields_medium("elastic_psv", "forward", k - 1)
    eng.corrector_phase("elastic_psv")
    eng.compute_seismograms_wavefield("forward", k - 1)
    el = [eng.field(x).astype(np.float64) for x in ("displacement", "velocity", "acceleration")]
    ac = [eng.acoustic_field(x).astype(np.float64) for x in ("displacement", "velocity", "acceleration")]
    return el, ac


@pytest.mark.parametrize("case", ["coupled_acoustic_top", "acoustic_composite_80x60"])
def test_final_wavefields_of_both_media_match_oracle(case):
    mesh, cfg = load_case(case)
    nstep = 300
    print(f"[case] {case} wavefields after {nstep} steps", flush=True)
    eng = _engine(mesh, cfg, ["displacement"], "colored", nstep)
    el, ac = _coupled_snapshot(eng, nstep)
    eng.close()
    refs = {}
    for prec in (4, 8):
        _, _, fields = oracle.Oracle2D(mesh, prec, 1, cfg["dt"]).forward(cfg["sources"], cfg["stations"], nstep,
                                                                         cfg["dt"], 1, ("displacement",),
                                                                         return_fields=True)
        refs[prec] = fields
    coupled = eng.nspec_elastic > 0
    for i, kind in enumerate(("displacement", "velocity", "acceleration")):
        pairs = [("acoustic", ac[i], refs[4][1][i], refs[8][1][i])]
        if coupled:
            pairs.append(("elastic", el[i], refs[4][0][i], refs[8][0][i]))
        for medium, ours, r32, r64 in pairs:
            e32, e64, floor = rel_l2(ours, r32), rel_l2(ours, r64), rel_l2(r32, r64)
            _report(f"{case}/{medium}/{kind}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
            # the elastic layer of the coupled mesh holds a small field at step
            # 300 (the FP32 oracle is 1e-4 from the FP64 one there): the
            # rounding-floor rule, as for v and a
            _check(e32, e64, floor)


def test_coupled_loop_seam_by_seam_equals_fused_run():
    """time_marching<forward>::run driven one seam at a time through the
    solver mirror (specfempp_amd.solver) on the coupled mesh: the acoustic
    and elastic predictors, updates and correctors in the reference's order
    and the accounting of time_marching.tpp:90-108 give the fused run's
    seismograms (COLORED: the same kernels in the same order)."""
    from specfempp_amd.solver import Newmark, TimeMarching

    mesh, cfg = load_case("coupled_acoustic_bottom")
    nstep = 200
    a = _engine(mesh, cfg, ["displacement", "pressure"], "colored", nstep)
    a.initialize()
    a.run()
    fused = a.seismograms().astype(np.float64)
    a.close()
    b = _engine(mesh, cfg, ["displacement", "pressure"], "colored", nstep)
    ts = Newmark(nstep, 1, cfg["dt"], "forward")
    ts.link_assembly(b)
    TimeMarching(b, ts).run()
    seam = b.seismograms().astype(np.float64)
    b.close()
    err = rel_l2(seam, fused)
    _report("coupled/seam_vs_fused", rel_l2=err)
    assert err <= 1e-6
