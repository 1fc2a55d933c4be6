"""Run the 3-D GATHER path of a host (ASan) build of the library on a small
cube and compare it with the FP32 oracle.  Usage: check_gather3d.py LIB.so"""
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
from specfempp_amd import Engine3D, meshes  # noqa: E402

types = ("displacement", "velocity", "acceleration")
w = meshes.config3(n=3, nstep=12, h=100.0, nrec_side=2)
eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, types, scatter="gather")
# a layered medium: one (kappa, mu) per element
assert eng.layout()["element_properties"], eng.layout()
eng.initialize()
eng.run()
ours = eng.seismograms().astype(np.float64)
eng.initialize()
for istep in range(w.nstep):
    eng.apply_predictor_phase()
    eng.update_wavefields(istep)
    eng.apply_corrector_phase()
    eng.compute_seismograms(istep)
seam = eng.seismograms().astype(np.float64)
# the same run through sfx_run_noalloc (element forces in the static scratch)
eng.initialize()
eng.run_noalloc(0, 5)
eng.run_noalloc(5, w.nstep - 5)
noalloc = eng.seismograms().astype(np.float64)
eng.close()
assert np.array_equal(noalloc, ours), "sfx_run_noalloc differs from sfx_run"
ref, _ = oracle.Oracle3D(w.mesh, 4, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1, types)
err = float(np.linalg.norm(ours[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0]))
seam_err = float(np.abs(seam - ours).max() / np.abs(ours).max())
print(f"displacement rel L2 vs oracle {err:.3e}, seam vs fused {seam_err:.3e}", flush=True)
sys.exit(0 if err <= 1e-5 and seam_err <= 1e-5 else 1)
