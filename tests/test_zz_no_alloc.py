"""sfx_run_noalloc (steps without device allocation, element forces in the
library's static per-GPU scratch) against sfx_run, through the C ABI.

Two calls in a row, the second on a smaller engine: both must give exactly
the seismograms and wavefields of sfx_run on a twin engine (same kernels in
the same order), and the scratch's contents left by the first call must not
leak into the second.  Engines above SFX_NOALLOC_MAX_ELEMENTS are refused,
and so are paths other than the 3-D gather path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TYPES = ("displacement", "velocity", "acceleration")
FIELDS = ("displacement", "velocity", "acceleration")


def _pair(n, nstep):
    from specfempp_amd import Engine3D, meshes

    w = meshes.config3(n=n, nstep=nstep, h=100.0, nrec_side=3)
    mk = lambda: Engine3D(w.mesh, w.sources, w.receivers, w.dt, nstep, 1, TYPES, scatter="gather")  # noqa: E731
    return mk(), mk()


@pytest.mark.parametrize("n,nstep", [(10, 300), (6, 300)])
def test_noalloc_run_equals_run(n, nstep):
    """n = 10 (1000 elements) first, then n = 6 (216 elements) reusing the
    scratch the larger run filled."""
    print(f"[case] sfx_run_noalloc vs sfx_run, {n}^3, {nstep} steps", flush=True)
    a, b = _pair(n, nstep)
    a.initialize()
    a.run()
    b.initialize()
    # split into two calls to cover a restart mid-run
    b.run_noalloc(0, nstep // 3)
    b.run_noalloc(nstep // 3, nstep - nstep // 3)
    ref, got = a.seismograms(), b.seismograms()
    assert np.array_equal(got, ref)
    for k in FIELDS:
        assert np.array_equal(b.field(k), a.field(k)), k
    a.close()
    b.close()


def test_noalloc_run_limits():
    from specfempp_amd import Engine2D, Engine3D, SfxError, meshes
    from specfempp_amd._native import load

    lib = load()
    n = 36  # 46,656 elements > SFX_NOALLOC_MAX_ELEMENTS = 43,690
    w = meshes.config3(n=n, nstep=4, h=100.0, nrec_side=2)
    e = Engine3D(w.mesh, w.sources, w.receivers, w.dt, 4, 1, ("displacement",), scatter="gather")
    e.initialize()
    with pytest.raises(SfxError, match="scratch limit"):
        e.run_noalloc()
    e.close()
    w = meshes.config3(n=4, nstep=4, h=100.0, nrec_side=2)
    e = Engine3D(w.mesh, w.sources, w.receivers, w.dt, 4, 1, ("displacement",), scatter="colored")
    e.initialize()
    with pytest.raises(SfxError, match="SFX_SCATTER_GATHER"):
        e.run_noalloc()
    e.close()
    w2 = meshes.config1(nstep=4, n=8)
    e = Engine2D(w2.mesh, w2.sources, w2.receivers, w2.dt, 4, 1, ("displacement",))
    e.initialize()
    assert lib.sfx_run_noalloc(e.h, 0, 4) < 0
    e.close()
