"""The reference's loops driven seam by seam through the C ABI
(specfempp_amd.solver: Newmark, DomainKernels, FrechetKernels, TimeMarching,
the mirror of time_marching.tpp:9-247) against the engine's fused paths and
the oracle.

- forward: TimeMarching over Newmark<forward> equals run() (2-D colored and
  gather, 3-D gather), with the reference's DOF and element accounting;
- combined: forward_record, the forward state copied to the buffer
  wavefield, then TimeMarching over Newmark<combined> (adjoint field forward,
  backward field with -dt replaying the Stacey record, the buffer copied in
  after the first backward step, Fréchet kernels every step) equals
  sfx_adjoint_run within FP32 rounding of the two kernel orderings, and the
  FP32/FP64 oracle at the bars of test_gpu_adjoint.
"""
import numpy as np
import pytest

import oracle
from conftest import load_case, rel_l2
from test_gpu_adjoint import DT, NAMES, NSTEP, REC, SRC, TYPES, _check, _mesh, _report

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,scatter", [(2, "colored"), (2, "gather"), (3, "gather")])
def test_forward_loop_seam_by_seam_equals_fused_run(dim, scatter):
    from specfempp_amd import Engine2D, Engine3D, meshes
    from specfempp_amd.solver import Newmark, TimeMarching

    types = ("displacement", "velocity", "acceleration")
    if dim == 2:
        mesh, cfg = load_case("elastic_stacey_80x60")
        mk = lambda: Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], 200, 2, types,  # noqa: E731
                              scatter=scatter)
    else:
        w = meshes.config3(n=6, nstep=200, h=100.0, nrec_side=3)
        mk = lambda: Engine3D(w.mesh, w.sources, w.receivers, w.dt, 200, 2, types, scatter=scatter)  # noqa: E731
    print(f"[case] forward loop seam by seam, {dim}-D {scatter}", flush=True)
    fused = mk()
    fused.initialize()
    fused.run()
    ref = fused.seismograms().astype(np.float64)
    fused.close()
    eng = mk()
    ts = Newmark(200, 2, eng.dt, "forward")
    ts.link_assembly(eng)
    TimeMarching(eng, ts).run()
    got = eng.seismograms().astype(np.float64)
    assert ts.get_seismogram_step() == 100
    eng.close()
    for it, t in enumerate(types):
        err = rel_l2(got[:, it], ref[:, it])
        _report(f"seams_forward_{dim}d_{scatter}/{t}", vs_fused=err)
        assert err <= (1e-6 if t != "acceleration" else 1e-5)


def test_combined_loop_seam_by_seam_equals_adjoint_run_and_oracle():
    from specfempp_amd import Engine2D
    from specfempp_amd.solver import Newmark, TimeMarching

    mesh = _mesh()
    rng = np.random.default_rng(7)
    t = np.arange(NSTEP) * DT
    stf = np.zeros((len(REC), NSTEP, 2))
    for r in range(len(REC)):
        for c in range(2):
            a, f, ph = rng.uniform(0.5, 1.5), rng.uniform(8, 20), rng.uniform(0, 2 * np.pi)
            stf[r, :, c] = 1e-6 * a * np.sin(2 * np.pi * f * t + ph) * np.exp(-((t - 0.25) / 0.1) ** 2)
    print("[case] combined loop seam by seam vs sfx_adjoint_run", flush=True)
    fused = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, TYPES, scatter="colored")
    fused.initialize()
    fused.forward_record(0)
    fused.set_adjoint_sources(REC, stf)
    fused.adjoint_run()
    k_fused, b_fused = fused.kernels(), fused.seismograms().astype(np.float64)
    fused.close()

    eng = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, TYPES, scatter="colored")
    eng.initialize()
    eng.forward_record(0)
    eng.copy_wavefield("buffer", "forward")
    eng.set_adjoint_sources(REC, stf)
    ts = Newmark(NSTEP, 1, DT, "combined")
    ts.link_assembly(eng)
    TimeMarching(eng, ts).run()
    k_seam, b_seam = eng.kernels(), eng.seismograms().astype(np.float64)
    eng.close()

    results = {}
    for prec in (4, 8):
        o = oracle.Oracle2D(mesh, prec, 1, DT)
        _, _, fields, record = o.forward_record(SRC, REC, NSTEP, DT, 1, TYPES)
        results[prec] = o.combined(SRC, REC, stf, REC, NSTEP, DT, fields, record, 1, TYPES)
    (k32, b32), (k64, b64) = results[4], results[8]
    e = rel_l2(b_seam, b_fused)
    e32, e64, floor = rel_l2(b_seam, b32), rel_l2(b_seam, b64), rel_l2(b32, b64)
    _report("seams_combined/backward_seismograms", vs_fused=e, vs_oracle_fp32=e32, vs_oracle_fp64=e64,
            fp32_floor=floor)
    assert e <= 1e-5
    _check(e32, e64, floor)
    for k, name in enumerate(NAMES):
        ours = k_seam[name].astype(np.float64)
        e = rel_l2(ours, k_fused[name].astype(np.float64))
        e32, e64, floor = rel_l2(ours, k32[k]), rel_l2(ours, k64[k]), rel_l2(k32[k], k64[k])
        _report(f"seams_combined/K_{name}", vs_fused=e, vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        assert e <= 1e-4
        _check(e32, e64, floor)
