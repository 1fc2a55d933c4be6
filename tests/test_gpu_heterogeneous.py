"""GPU parity on the heterogeneous 2-D workload (Gaussian random field
medium, one material per element, Stacey on every side).

* Small instance (32 x 32 elements of 2.5 m): seismograms within the forward
  bar of the FP32 oracle (relative L2 <= 1e-5), error to the FP64 oracle
  printed.
* Instance at the full workload's spacing and field statistics over a
  1 km x 1 km square (400 x 400 elements, 2.56 M points): discrete
  reciprocity G_zz(b <- a) = G_zz(a <- b) and G_zx(b <- a) = G_xz(a <- b)
  (a property of the symmetric mass, Stacey and stiffness matrices that
  holds at any size; the oracle pins it in FP64 to 1e-10).  The FP32 bar is
  1e-4 relative L2, of the order of the FP32 rounding floor of a
  6000-step run.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from specfempp_amd import meshes

pytestmark = pytest.mark.gpu


def _engine(w, sources, receivers):
    from specfempp_amd import Engine2D

    return Engine2D(w.mesh, sources, receivers, w.dt, w.nstep, 1, ("displacement",))


def _run(w, sources, receivers):
    e = _engine(w, sources, receivers)
    e.initialize()
    e.run()
    s = e.seismograms().astype(np.float64)
    e.close()
    return s


def test_grf_small_matches_oracle():
    w = meshes.config2(n=32, L=80.0, nstep=3000, nrec=12, seed=5, corr=20.0)
    got = _run(w, w.sources, w.receivers)
    s32, _ = oracle.Oracle2D(w.mesh, 4, 1, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1,
                                                          ("displacement",))
    s64, _ = oracle.Oracle2D(w.mesh, 8, 1, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1,
                                                          ("displacement",))
    e32, e64, floor = rel_l2(got, s32), rel_l2(got, s64), rel_l2(s32, s64)
    print(f"[grf] 32x32: e32={e32:.2e} e64={e64:.2e} fp32 floor={floor:.2e}", flush=True)
    assert e32 <= 1e-5


def test_grf_reciprocity_at_full_spacing():
    w = meshes.config2(n=400, L=1000.0, nstep=6000, nrec=1, seed=2)
    a, b = dict(x=431.7, z=388.2), dict(x=612.9, z=547.3)   # 250 m apart

    def src(p, angle):
        return [dict(x=p["x"], z=p["z"], angle=angle, f0=15.0, factor=1e10, tshift=0.0)]

    z_at_b = _run(w, src(a, 0.0), [b])[:, 0, 0]   # (nsig, 2)
    z_at_a = _run(w, src(b, 0.0), [a])[:, 0, 0]
    x_at_b = _run(w, src(a, 90.0), [b])[:, 0, 0]
    ezz = rel_l2(z_at_a[:, 1], z_at_b[:, 1])
    exz = rel_l2(-z_at_a[:, 0], x_at_b[:, 1])
    print(f"[grf] 400x400 reciprocity: zz={ezz:.2e} xz={exz:.2e}, "
          f"peak |u_z| {np.abs(z_at_b[:, 1]).max():.3e}", flush=True)
    assert np.abs(z_at_b[:, 1]).max() > 0
    assert ezz <= 1e-4 and exz <= 1e-4
