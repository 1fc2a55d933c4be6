"""Run the 3-D mesh-database path of a host (ASan) build of the library on a
small irregular mesh: COLORED and ATOMIC fused runs, the seam calls, the
field getters; compare with the FP32 oracle on the same database.
Usage: check_database3d.py LIB.so"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import specfempp_amd._native as nat  # noqa: E402

nat._lib = None
nat.LIB_PATH = sys.argv[1]
nat.load()
from specfempp_amd import Engine3DDatabase, meshes  # noqa: E402

types = ("displacement", "velocity", "acceleration")
nstep, dt = 10, 1e-3
mat = np.zeros((3, 3, 4), np.int32)
mat[1:] = 1
db = meshes.hex_database(4, 3, 3, 100.0, mat, [2500.0, 2700.0], [3000.0, 3600.0], [1730.0, 2000.0],
                         perturb=0.12, seed=3, dt=dt)
src = [dict(type="moment_tensor", x=140.0, y=160.0, z=130.0, mxx=0.3, myy=-0.5, mzz=0.8, mxy=0.2, mxz=-0.1,
            myz=0.4, f0=50.0, factor=1e12),
       dict(type="force", x=310.0, y=90.0, z=230.0, fx=0.3, fy=-0.5, fz=0.8, f0=50.0, factor=1e10)]
rec = [dict(x=150.0, y=150.0, z=150.0), dict(x=40.0, y=60.0, z=290.0), dict(x=360.0, y=250.0, z=100.0)]
ref, _ = oracle.Oracle3D.from_database(db, 4, dt).forward(src, rec, nstep, dt, 1, types)
worst = 0.0
for scatter in ("colored", "atomic"):
    eng = Engine3DDatabase(db, src, rec, dt, nstep, 1, types, scatter=scatter)
    eng.initialize()
    eng.run()
    ours = eng.seismograms().astype(np.float64)
    _ = [eng.field(k) for k in ("displacement", "velocity", "acceleration", "mass_inverse")]
    eng.initialize()
    for istep in range(nstep):
        eng.apply_predictor_phase()
        eng.update_wavefields(istep)
        eng.apply_corrector_phase()
        eng.compute_seismograms(istep)
    seam = eng.seismograms().astype(np.float64)
    eng.close()
    err = float(np.linalg.norm(ours[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0]))
    seam_err = float(np.abs(seam - ours).max() / np.abs(ours).max())
    worst = max(worst, err, seam_err)
    print(f"database {scatter}: displacement rel L2 vs oracle {err:.3e}, seam vs fused {seam_err:.3e}", flush=True)
sys.exit(0 if worst <= 1e-5 else 1)
