"""GPU parity on the reference's 9-node mesh-reader fixtures (80 x 60
elements under flat and curved topography, free surfaces, a 10 Hz force at
(2500, 2500), six stations, three of them on the curved surface): the
engine's seismograms against the FP32 and FP64 oracle, for both 2-D
scatters, to the bars of test_gpu_parity.py."""
import functools

import numpy as np
import pytest

import oracle
from conftest import load_case, rel_l2
from test_gpu_parity import TOL_SEISMO, _check, _report

pytestmark = pytest.mark.gpu

TYPES = ("displacement", "velocity", "acceleration")


@functools.lru_cache(maxsize=None)
def _oracle_runs(case):
    mesh, cfg = load_case(case)
    args = (cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"], 1, TYPES)
    s32, _ = oracle.Oracle2D(mesh, 4, 1, cfg["dt"]).forward(*args)
    s64, _ = oracle.Oracle2D(mesh, 8, 1, cfg["dt"]).forward(*args)
    return s32, s64


@pytest.mark.parametrize("scatter", ["colored", "gather"])
@pytest.mark.parametrize("case", ["flat_topography_80x60", "curved_topography_80x60"])
def test_reference_mesh_seismograms_match_oracle(case, scatter):
    from specfempp_amd import Engine2D

    mesh, cfg = load_case(case)
    print(f"[case] {case} scatter={scatter} nstep={cfg['nstep']}", flush=True)
    eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], cfg["nstep"], 1, TYPES, scatter=scatter)
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    eng.close()
    s32, s64 = _oracle_runs(case)
    assert np.isfinite(ours).all() and np.abs(ours[:, 0]).max() > 0
    for it, t in enumerate(TYPES):
        e32, e64 = rel_l2(ours[:, it], s32[:, it]), rel_l2(ours[:, it], s64[:, it])
        floor = rel_l2(s32[:, it], s64[:, it])
        _report(f"{case}_{scatter}/{t}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)
        if t == "displacement":
            assert e32 <= TOL_SEISMO
