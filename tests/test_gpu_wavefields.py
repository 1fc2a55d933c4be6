"""GPU parity of wavefield snapshots: u, v and a over every global point,
engine against the FP32 and FP64 oracle, at several times of one run.

The fused step leaves u and v predicted for the next step, so each snapshot
is taken by running the fused loop up to the step before it and doing that
step through the seams (update_wavefields, then the corrector): the state is
then the corrected state after k steps, which is what the oracle holds after a
k-step run (time_marching.tpp:33-84 ends every step on the corrector).  The
predictor is then applied and the fused loop continues.

Bars: u within relative L2 <= 1e-5 of the FP32 oracle (the north-star bar
for wavefield snapshots).  v and a sit on the FP32 rounding floor: the FP32
oracle's own distance to the FP64 oracle reaches 1e-5 for v (1.04e-5 on the
6^3 cube at step 400) and 2e-5 for a, and moves with the summation order
alone; they must be within 1e-5 of the FP32 oracle or within 1.5x that
floor of the FP64 oracle, the bar of the seismogram tests.  Every case
prints its error against the FP64 oracle.
"""
import numpy as np
import pytest

import oracle
from conftest import load_case, rel_l2
from test_gpu_parity import BLOCK_OFF, _check, _report

pytestmark = pytest.mark.gpu

KINDS = ("displacement", "velocity", "acceleration")


def engine_snapshots(eng, steps):
    """Fields (u, v, a) after k steps for every k in ``steps`` (increasing).
    The seam-stepped steps also take their seismogram sample (one sample per
    step), so the seismograms cover every step."""
    eng.initialize()
    done, out = 0, []
    for k in steps:
        if k - 1 > done:
            eng.run(done, k - 1 - done)
        eng.update_wavefields(k - 1)
        eng.apply_corrector_phase()
        eng.compute_seismograms(k - 1)
        out.append([eng.field(x).astype(np.float64) for x in KINDS])
        eng.apply_predictor_phase()
        done = k
    return out


def compare(name, ours, f32, f64):
    for i, kind in enumerate(KINDS):
        e32, e64, floor = rel_l2(ours[i], f32[i]), rel_l2(ours[i], f64[i]), rel_l2(f32[i], f64[i])
        _report(f"{name}/{kind}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        if kind == "displacement":
            assert e32 <= 1e-5, (name, kind, e32)
        else:
            _check(e32, e64, floor)


@pytest.mark.parametrize("scatter", ["colored", "gather"])
def test_2d_wavefield_snapshots_match_oracle(scatter):
    from specfempp_amd import Engine2D

    mesh, cfg = load_case("elastic_stacey_80x60")
    steps = (100, 200, 300)
    print(f"[case] 2-D elastic_stacey_80x60 snapshots {steps} scatter={scatter}", flush=True)
    eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], steps[-1], 1, ("displacement",),
                   scatter=scatter)
    snaps = engine_snapshots(eng, steps)
    eng.close()
    for k, ours in zip(steps, snaps):
        ref = []
        for prec in (4, 8):
            _, _, f = oracle.Oracle2D(mesh, prec, 1, cfg["dt"]).forward(cfg["sources"], cfg["stations"], k,
                                                                        cfg["dt"], 1, ("displacement",),
                                                                        return_fields=True)
            ref.append(f)
        compare(f"2d_{scatter}_step{k}", ours, ref[0], ref[1])
        if scatter == "gather":  # the oracle's operations in the oracle's order: the same bits
            for it in range(3):
                assert np.array_equal(ours[it].astype(np.float32), ref[0][it].astype(np.float32)), (k, it)


@pytest.mark.parametrize("scatter", ["gather", "colored", "atomic", pytest.param("block", marks=BLOCK_OFF)])
def test_3d_wavefield_snapshots_match_oracle(scatter):
    from specfempp_amd import Engine3D, meshes

    w = meshes.config3(n=6, nstep=600, h=100.0, nrec_side=3)
    steps = (250, 400, 600)
    print(f"[case] 3-D layered 6^3 snapshots {steps} scatter={scatter}", flush=True)
    eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, steps[-1], 1, ("displacement",), scatter=scatter)
    snaps = engine_snapshots(eng, steps)
    eng.close()
    for k, ours in zip(steps, snaps):
        ref = []
        for prec in (4, 8):
            _, _, f = oracle.Oracle3D(w.mesh, prec, w.dt).forward(w.sources, w.receivers, k, w.dt,
                                                                  return_fields=True)
            ref.append(f)
        compare(f"3d_{scatter}_step{k}", ours, ref[0], ref[1])


@pytest.mark.timeout(600)
def test_3d_full_size_64cube_wavefield_matches_threaded_oracle():
    """The full 64^3 configuration (16,974,593 global points, 400 surface
    receivers), 300 steps: engine u, v, a over every point and the 400
    seismograms against the FP32 and FP64 oracle, each on half of the host's
    threads (the threaded oracle gives the serial loop's bits), the FP64 run
    in a worker process.  300 steps cover the source's rise and peak (step
    240); the surface receivers are 3 km up, beyond the wave's reach, so
    their traces are the scheme's leading-edge tails, small but not zero."""
    import _oracle_jobs as jobs
    from specfempp_amd import Engine3D, meshes

    nstep = 300
    w = meshes.config3(n=64, nstep=nstep)
    half = max(1, jobs.host_workers() // 2)
    print(f"[case] 3-D 64^3 full size, {nstep} steps, gather, oracle threads={half} x 2", flush=True)
    with jobs.pool(1) as ex:
        f64 = ex.submit(jobs.forward_3d, w.mesh, 8, w.sources, w.receivers, nstep, w.dt, half)
        eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, nstep, 1, ("displacement",), scatter="gather")
        assert eng.nglob == 16_974_593 and eng.nrec == 400
        ours = engine_snapshots(eng, (nstep,))[0]
        seis = eng.seismograms().astype(np.float64)
        eng.close()
        s32, _, g32 = oracle.Oracle3D(w.mesh, 4, w.dt).forward(w.sources, w.receivers, nstep, w.dt,
                                                               return_fields=True, nthreads=half)
        s64, g64 = f64.result()
    compare(f"3d_64cube_step{nstep}", ours, g32, g64)
    e32, e64, floor = rel_l2(seis, s32), rel_l2(seis, s64), rel_l2(s32, s64)
    _report("3d_64cube/seismograms", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor,
            max_abs_ref=float(np.abs(s64).max()))
    _check(e32, e64, floor)
