"""GPU parity of batched 2-D runs.

Each simulation of a batch must give the seismograms of its own engine run
alone (same kernels and coloring; only the order of float atomics on Stacey
and source points may differ, so the bar is relative L2 <= 1e-6) and those of
the FP32 oracle (<= 1e-5, the bar of every forward parity test), whatever
launch geometry the autotuner picks, with and without graph capture.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from specfempp_amd import meshes

pytestmark = pytest.mark.gpu

TYPES = ("displacement", "velocity")


def _workloads():
    return meshes.config4(nsim=6, seed=11, nmin=8, nmax=24, nstep_range=(150, 400), nrec=4)


def _engines(ws):
    from specfempp_amd import Engine2D

    return [Engine2D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, TYPES) for w in ws]


@pytest.fixture(scope="module")
def alone():
    ws = _workloads()
    out = []
    for e in _engines(ws):
        e.initialize()
        e.run()
        out.append(e.seismograms().astype(np.float64))
        e.close()
    return ws, out


@pytest.mark.parametrize("elems_per_block,graphs", [(5, True), (2, False), (10, True)])
def test_batch_matches_single_runs(alone, elems_per_block, graphs):
    from specfempp_amd import Batch

    ws, ref = alone
    engines = _engines(ws)
    b = Batch(engines)
    for e in engines:
        e.close()  # the batch holds its own copy
    b.set_config(elems_per_block, graphs)
    b.run()
    for i, w in enumerate(ws):
        got = b.seismograms(i).astype(np.float64)
        assert got.shape == ref[i].shape
        err = rel_l2(got, ref[i])
        print(f"[batch] E={elems_per_block} graphs={graphs} sim {i} ({w.name}, {w.nstep} steps): {err:.2e}",
              flush=True)
        assert err <= 1e-6
    info = b.info()
    assert info["point_steps"] == sum(((4 * w.notes["n"] + 1) ** 2) * w.nstep for w in ws)
    b.close()


def test_batch_matches_oracle_and_tuning_keeps_results(alone):
    from specfempp_amd import Batch

    ws, ref = alone
    b = Batch(_engines(ws))
    choice = b.tune(3)
    assert choice["elems_per_block"] in (2, 5, 10)
    b.run()
    for i in (0, len(ws) - 1):
        w = ws[i]
        s32, _ = oracle.Oracle2D(w.mesh, 4, 1, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1, TYPES)
        got = b.seismograms(i).astype(np.float64)
        err = rel_l2(got[:, 0], s32[:, 0])
        print(f"[batch] tuned {choice} sim {i} vs FP32 oracle: {err:.2e}", flush=True)
        assert err <= 1e-5
        assert rel_l2(got, ref[i]) <= 1e-6
    b.close()


def test_batch_of_host_only_engines(alone):
    """A batch built from host-only engines (no device memory of their own)
    gives the same seismograms."""
    from specfempp_amd import Batch, Engine2D

    ws, ref = alone
    engines = [Engine2D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, TYPES, host_only=True) for w in ws]
    b = Batch(engines)
    for e in engines:
        e.close()
    b.set_config(5, True)
    b.run()
    for i in range(len(ws)):
        assert rel_l2(b.seismograms(i).astype(np.float64), ref[i]) <= 1e-6
    b.close()


def test_batch_at_the_configuration_extremes_matches_oracle():
    """Members at the batched configuration's extremes (40^2 quads for 2400
    steps, 240^2 quads for 800 steps; h 50 m, vp ~ U[2000, 4000], one Ricker
    force of f0 ~ U[5, 20] Hz, 32 receivers, dt at Courant 0.4) in one batch
    with a mid-size member: each against the FP32 oracle (<= 1e-5) and its own
    engine run alone (<= 1e-6).  The oracle runs proceed in worker processes
    while the GPU runs."""
    import _oracle_jobs as jobs
    from specfempp_amd import Batch

    ws = (meshes.config4(nsim=1, seed=41, nmin=40, nmax=40, nstep_range=(2400, 2400))
          + meshes.config4(nsim=1, seed=42, nmin=240, nmax=240, nstep_range=(800, 800))
          + meshes.config4(nsim=1, seed=43, nmin=120, nmax=120, nstep_range=(1500, 1500)))
    print("[case] batch extremes: " + ", ".join(f"{w.name} x {w.nstep}" for w in ws), flush=True)
    with jobs.pool(len(ws)) as ex:
        fut = [ex.submit(jobs.forward_2d, w.mesh, 4, w.sources, w.receivers, w.nstep, w.dt, TYPES) for w in ws]
        alone = []
        for e in _engines(ws):
            e.initialize()
            e.run()
            alone.append(e.seismograms().astype(np.float64))
            e.close()
        b = Batch(_engines(ws))
        choice = b.tune(3)
        b.run()
        got = [b.seismograms(i).astype(np.float64) for i in range(len(ws))]
        b.close()
        ref = [f.result() for f in fut]
    for i, w in enumerate(ws):
        e32 = rel_l2(got[i][:, 0], ref[i][:, 0])
        e_alone = rel_l2(got[i], alone[i])
        print(f"[batch] extremes {choice} {w.name} x {w.nstep}: vs FP32 oracle {e32:.2e}, vs alone {e_alone:.2e}",
              flush=True)
        assert e32 <= 1e-5
        assert e_alone <= 1e-6
