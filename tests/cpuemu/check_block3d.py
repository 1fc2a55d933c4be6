"""Run the 3-D BLOCK path of a host (ASan + bounds) build of the library on a
small cube: fused steps and the seam calls, compared with each other and with
the FP32 oracle.  Usage: check_block3d.py LIB.so"""
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
# 4^3 elements = 8 blocks; a 50 Hz source rises within the 10 steps
w = meshes.config3(n=4, nstep=10, h=100.0, nrec_side=2)
w.sources[0].update(f0=50.0)
eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, types, scatter="block")
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
eng.close()
ref, _ = oracle.Oracle3D(w.mesh, 4, w.dt).forward(w.sources, w.receivers, w.nstep, w.dt, 1, types)
err = float(np.linalg.norm(ours[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0]))
seam_err = float(np.abs(seam - ours).max() / np.abs(ours).max())
print(f"block: displacement rel L2 vs oracle {err:.3e}, seam vs fused {seam_err:.3e}", flush=True)
sys.exit(0 if err <= 1e-5 and seam_err <= 1e-5 else 1)
