"""Run the 2-D acoustic and coupled elastic-acoustic paths of a host (ASan)
build of the library on the reference's acoustic and coupled Newmark
fixtures for a few steps: fused steps (COLORED and ATOMIC, every seismogram
type), the seam-by-seam loop, the field getters; compare with the FP32
oracle.  Usage: check_acoustic2d.py LIB.so [NSTEP]"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import specfempp_amd._native as nat  # noqa: E402

nat._lib = None
nat.LIB_PATH = sys.argv[1]
nat.load()
from conftest import load_case, rel_l2  # noqa: E402
from specfempp_amd import Engine2D  # noqa: E402
from specfempp_amd.solver import Newmark, TimeMarching  # noqa: E402

NSTEP = int(sys.argv[2]) if len(sys.argv) > 2 else 4
TYPES = ("displacement", "velocity", "acceleration", "pressure")
CASES = ["acoustic_free_80x60", "acoustic_dirichlet_80x60", "acoustic_composite_80x60", "acoustic_stacey_80x60",
         "coupled_acoustic_top", "coupled_acoustic_bottom"]
worst = 0.0
for case in CASES:
    mesh, cfg = load_case(case)
    # move the sources and stations so that a few steps see signal: the
    # first source at a station's place, the time shift zero
    ref, _ = oracle.Oracle2D(mesh, 4, 1, cfg["dt"]).forward(cfg["sources"], cfg["stations"], NSTEP, cfg["dt"], 1,
                                                            TYPES)
    for scatter in ("colored", "atomic"):
        eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], NSTEP, 1, TYPES, scatter=scatter)
        eng.initialize()
        eng.run()
        fused = eng.seismograms().astype(np.float64)
        _ = eng.field("displacement"), eng.acoustic_field("velocity"), eng.acoustic_field("mass_inverse")
        eng.close()
        eng = Engine2D(mesh, cfg["sources"], cfg["stations"], cfg["dt"], NSTEP, 1, TYPES, scatter=scatter)
        ts = Newmark(NSTEP, 1, cfg["dt"], "forward")
        ts.link_assembly(eng)
        TimeMarching(eng, ts).run()
        seam = eng.seismograms().astype(np.float64)
        eng.close()
        e = rel_l2(fused, ref) if np.linalg.norm(ref) > 0 else float(np.abs(fused).max())
        s = rel_l2(seam, fused) if np.linalg.norm(fused) > 0 else float(np.abs(seam).max())
        worst = max(worst, e)
        print(f"{case} {scatter}: rel L2 vs oracle {e:.3e}, seam vs fused {s:.3e}, |ref| {np.abs(ref).max():.3e}",
              flush=True)
sys.exit(0 if worst <= 1e-4 else 1)
