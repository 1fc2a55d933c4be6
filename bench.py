"""Benchmark: GLL point-steps per second of the 3-D elastic NGLL=5 FP32
forward Newmark loop (64^3 layered cube) on one MI355X, with the roofline of
the dominant kernel (stiffness) and the CPU oracle timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 64] [--no-cpu]

Prints one JSON line.  A "step" is one Newmark step of the whole hot path:
sources, stiffness (gather, gradient, stress, divergence, atomic scatter),
Stacey, seismograms, fused mass division + corrector + predictor.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes_3d(nspec: int, nglob: int) -> dict:
    """Bytes each kernel of the 3-D gather path must move per launch (FP32):
      elemforce_3d:       kappa + mu per local point (8 B x 125 / element), the
                          displacement once per global point (12 B), the element
                          forces stored once (12 B x 125 / element)
      assemble_update_3d: the element forces loaded once (12 B x 125 / element),
                          u, v, M^-1 read (36 B) and u, v written (24 B) per point
    and the survey's data-model figure for a whole step (143 B per point-step:
    gather index + kappa + mu per local point, displacement gather, acceleration
    read-modify-write, fused update read u v a M^-1 / write u v a)."""
    elem = nspec * 125 * 8 + nglob * 12 + nspec * 125 * 12
    upd = nspec * 125 * 12 + nglob * 60
    survey = nspec * 125 * 12 + nglob * 36 + nglob * 84
    return {"elemforce": elem, "update": upd, "per_pt_step": (elem + upd) / nglob,
            "survey_per_pt_step": survey / nglob}


def cpu_baseline(w, nsteps: int) -> dict:
    """The FP32 oracle (single thread) on the same 64^3 workload for a few steps."""
    import oracle

    t = time.perf_counter()
    o = oracle.Oracle3D(w.mesh, 4, w.dt)
    build_s = time.perf_counter() - t
    t = time.perf_counter()
    o.forward(w.sources, w.receivers, nsteps, w.dt)
    el = time.perf_counter() - t
    return {"value": o.nglob * nsteps / el, "unit": "GLL pt-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP32 oracle (C++ restatement, 1 thread), {nsteps} steps of the same "
                      f"{w.mesh['nx']}^3 workload, {el:.1f} s timed (+{build_s:.1f} s setup untimed)"}


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5) -> dict:
    """A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counters
    give the scale factor of the byte counters, its time the copy bandwidth."""
    import torch

    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t) / reps
    return {"bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes, "copies": reps + 1,
            "copy_GBs": 2 * nbytes / el / 1e9}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (scale reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine is single-GPU).",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from specfempp_amd import Engine3D, meshes

    w = meshes.config3(n=args.n, nstep=args.warmup + args.steps)
    t = time.perf_counter()
    eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, ("displacement",))
    setup_s = time.perf_counter() - t
    eng.initialize()
    if args.warmup:
        eng.run(0, args.warmup)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(args.warmup, args.steps)
    eng.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tm = eng.timings()
    seis = eng.seismograms()
    finite = bool(np.isfinite(seis).all())
    nglob, nspec = eng.nglob, eng.nspec
    eng.close()

    ms_per_step = 1e3 * elapsed / args.steps
    value = nglob * args.steps / elapsed
    ab = algorithmic_bytes_3d(nspec, nglob)
    kern = {"elemforce_3d_kernel": (ab["elemforce"], tm["stiffness_ms"]),
            "assemble_update_3d_kernel": (ab["update"], tm["update_ms"])}
    dominant = max(kern, key=lambda k: kern[k][1])
    nbytes, ms = kern[dominant]
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    per_kernel = {k: {"avg_ms": v[1], "algorithmic_bytes": v[0],
                      "GBs": (v[0] / (v[1] * 1e-3) / 1e9 if v[1] > 0 else 0.0)} for k, v in kern.items()}
    step_gbs = ab["per_pt_step"] * value / 1e9
    survey_gbs = ab["survey_per_pt_step"] * value / 1e9
    out = {
        "metric": "GLL pt-steps/s, 3D elastic NGLL=5 FP32 forward (64^3 layered, Stacey + free surface)",
        "value": value,
        "unit": "GLL pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3d_layered_cube_{args.n}^3_ngll5", "nspec": nspec, "nglob": nglob,
                   "dt": w.dt, "dt_stated": w.dt_stated, "nrec": len(w.receivers), "source": "moment tensor",
                   "scatter": "gather", "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant,
                     "algorithmic_bytes_per_launch": nbytes, "avg_ms": ms},
        "kernels": per_kernel,
        "kernels_ms": tm,
        "step_hbm_gbs": step_gbs,
        "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "algorithmic_bytes_per_pt_step": ab["per_pt_step"],
        "survey_bytes_per_pt_step": ab["survey_per_pt_step"],
        "survey_step_hbm_frac": survey_gbs / HBM_PEAK_GBS,
        "seismograms_finite": finite,
    }
    if args.calibrate_copy:
        out["copy_calibration"] = copy_calibration()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
