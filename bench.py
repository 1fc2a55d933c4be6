"""Benchmark: GLL point-steps per second of the 3-D elastic NGLL=5 FP32
forward Newmark loop (64^3 layered cube) on one MI355X, with the roofline of
the dominant kernel and the CPU oracle timed beside it; then the batched
multi-event leg (64 2-D simulations in one batch vs the same 64 run one after
another), the adjoint leg (200^2 2-D gradient vs a plain forward run) and the
config-2 leg (2-D Gaussian-random-field medium, 4000^2 quads of 2.5 m, 256 M
points, with the roofline of its dominant kernel).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 64] [--no-cpu]
                    [--no-batch] [--no-adjoint] [--no-config2] [--c2-steps K]
                    [--scatter auto|gather|colored|atomic]
    python bench.py --config2 [--c2-n 4000] [--steps K] [--warmup W]

Prints one JSON line.  --config2 measures only the config-2 workload.
A "step" is one Newmark step of the whole hot path:
sources, stiffness (gradient, stress, divergence, assembly), Stacey,
seismograms, fused mass division + corrector + predictor.
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


def algorithmic_bytes_block(n: int) -> dict:
    """Bytes each kernel of the 3-D block path must move per launch (FP32),
    n^3 elements in blocks of 2^3:
      block_step_3d:     the displacement once per point (12 B), kappa + mu per
                         local point (8 B x 125 / element), for the points
                         inside blocks v and M^-1 read and u, v written (48 B),
                         and the block-surface sums written (12 B per block and
                         surface point, 386 per block)
      surface_update_3d: those sums read, and u, v, M^-1 read and u, v written
                         (60 B) at the block-surface points."""
    nspec, nglob = n ** 3, (4 * n + 1) ** 3
    nb = nspec // 8
    nint = (7 * (n // 2)) ** 3
    nsurf = nglob - nint
    blk = nglob * 12 + nspec * 125 * 8 + nint * 48 + nb * 386 * 12
    srf = nb * 386 * 12 + nsurf * 60
    survey = nspec * 125 * 12 + nglob * 36 + nglob * 84
    return {"block_step_3d_kernel": blk, "surface_update_3d_kernel": srf, "per_pt_step": (blk + srf) / nglob,
            "survey_per_pt_step": survey / nglob}


def minv_bytes_3d(nx: int, ny: int, nz: int, absorbing) -> int:
    """Bytes of M^-1 that assemble_update_3d reads per launch: one component
    (16 B per row of 4 lattice points) in the tiles without absorbing-face
    points, where the three components are equal, and all three (48 B per row)
    in the tiles on an absorbing face (x-, x+, y-, y+, z-, z+ flags); the
    padding rows of the last tile layer in y and z are not read."""
    import numpy as np

    TX, TY, TZ = nx + 1, ny + 1, nz + 1
    ty, tz = np.meshgrid(np.arange(TY), np.arange(TZ), indexing="ij")
    rows_j = np.where(ty == ny, 1, 4)
    rows_k = np.where(tz == nz, 1, 4)
    rows = (rows_j * rows_k).astype(np.int64)  # rows per tile, the same for every tx
    on_face_yz = (((absorbing[2] and 1) and (ty == 0)) | ((absorbing[3] and 1) and (ty == ny)) |
                  ((absorbing[4] and 1) and (tz == 0)) | ((absorbing[5] and 1) and (tz == nz)))
    total = 0
    for tx in range(TX):
        on_face = on_face_yz | bool((absorbing[0] and tx == 0) or (absorbing[1] and tx == nx))
        total += int(np.sum(np.where(on_face, 48, 16) * rows))
    return total


def algorithmic_bytes_3d(nspec: int, nglob: int, minv_bytes: int = None, element_properties: bool = False) -> dict:
    """Bytes each kernel of the 3-D gather path must move per launch (FP32):
      elemforce_3d:       kappa + mu per local point (8 B x 125 / element), or
                          one pair per element (8 B) when the engine reads them
                          so (every element holds one value of each), the
                          displacement once per global point (12 B), the element
                          forces stored once (12 B x 125 / element)
      assemble_update_3d: the element forces loaded once (12 B x 125 / element),
                          u, v read (24 B) and u, v written (24 B) per point, and
                          M^-1 (minv_bytes, see minv_bytes_3d; 12 B per point
                          when not given)
    and the survey's data-model figure for a whole step (143 B per point-step:
    gather index + kappa + mu per local point, displacement gather, acceleration
    read-modify-write, fused update read u v a M^-1 / write u v a)."""
    if minv_bytes is None:
        minv_bytes = nglob * 12
    props = nspec * 8 if element_properties else nspec * 125 * 8
    elem = props + nglob * 12 + nspec * 125 * 12
    upd = nspec * 125 * 12 + nglob * 48 + minv_bytes
    survey = nspec * 125 * 12 + nglob * 36 + nglob * 84
    return {"elemforce": elem, "update": upd, "per_pt_step": (elem + upd) / nglob,
            "survey_per_pt_step": survey / nglob}


FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (packed FP32 instructions)
FP32_VALU_UNPACKED_TFLOPS = 78.6  # one FP32 operation per lane per cycle (packed FP32 is not built)


def elemforce_3d_flops(nspec: int) -> float:
    """FP32 operations elemforce_3d must perform per launch (FMA = 2): per
    local point, the three 5-term derivative contractions of three components
    (45 FMA), the isotropic stress (22), the three 5-term divergence
    contractions (45 FMA) and their weighting (18): 220 per point, 125 points
    per element."""
    return nspec * 125 * (90 + 22 + 90 + 18)


def measured_valu_flops(workload: str, scatter: str, kernel: str):
    """FP32 VALU operations per launch of a kernel from the committed
    rocprofv3 instruction-count summary, or None."""
    path = os.path.join(ROOT, "profiles", "gather3d_64cube_valu.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        prof = json.load(fh)
    if prof.get("workload") != workload or prof.get("scatter") != scatter or kernel not in prof["per_dispatch"]:
        return None
    return prof["per_dispatch"][kernel]["fp32_flops"]


def measured_traffic(workload: str, scatter: str, kernel: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each scaled by
    the copy-kernel factor recorded with it), or None."""
    path = os.path.join(ROOT, "profiles", "gather3d_64cube_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        pmc = json.load(fh)
    if pmc.get("workload") != workload or pmc.get("scatter") != scatter or kernel not in pmc["kib_per_dispatch"]:
        return None
    c = pmc["kib_per_dispatch"][kernel]
    return 1024.0 * (c["FETCH_SIZE"] * pmc["scale"]["FETCH_SIZE"] + c["WRITE_SIZE"] * pmc["scale"]["WRITE_SIZE"])


def cpu_threads() -> int:
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set, else the
    CPUs this process may run on, at most 16 (the host-thread share one GPU
    is given on these machines)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
    return max(1, min(16, n))


def host_info() -> dict:
    """nproc, the CPUs this process may use, and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS", "")}


def cpu_baseline(w, nsteps: int) -> dict:
    """The FP32 oracle on the same 64^3 workload for a few steps: on all host
    threads (the element loop in parallel, the serial loop's bits), and on one
    thread.  The threaded rate is taken from the difference of a short and a
    longer run, so the per-run setup (point-to-element index) is not counted."""
    import oracle

    t = time.perf_counter()
    o = oracle.Oracle3D(w.mesh, 4, w.dt)
    build_s = time.perf_counter() - t
    t = time.perf_counter()
    o.forward(w.sources, w.receivers, nsteps, w.dt, nthreads=1)
    el1 = time.perf_counter() - t
    one = o.nglob * nsteps / el1
    nt = cpu_threads()
    out = {"value": one, "unit": "GLL pt-steps/s", "cores": 1, "kind": "port", "host": host_info(),
           "sample": f"FP32 oracle (C++ restatement, 1 thread), {nsteps} steps of the same "
                     f"{w.mesh['nx']}^3 workload, {el1:.1f} s timed (+{build_s:.1f} s setup untimed)"}
    if nt > 1:
        extra = 4 * nsteps
        t = time.perf_counter()
        o.forward(w.sources, w.receivers, 2, w.dt, nthreads=nt)
        ta = time.perf_counter() - t
        t = time.perf_counter()
        o.forward(w.sources, w.receivers, 2 + extra, w.dt, nthreads=nt)
        tb = time.perf_counter() - t
        out.update(value=o.nglob * extra / (tb - ta), cores=nt, single_thread_value=one,
                   cores_unused=max(0, (os.cpu_count() or nt) - nt),
                   sample=f"FP32 oracle (C++ restatement, OpenMP element loop, {nt} threads): "
                          f"{2 + extra} minus 2 steps of the same {w.mesh['nx']}^3 workload, "
                          f"{tb - ta:.1f} s; 1 thread: {nsteps} steps, {el1:.1f} s")
    return out


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


def batched_leg(nsim: int, tune_steps: int) -> dict:
    """Config 4: nsim 2-D simulations (40^2..240^2 quads, own vp, f0, source,
    32 receivers, 800..2400 steps) advanced as one batch vs run one after
    another with the plain engine, on its colored and on its gather path (the
    faster of the two is the sequential baseline).  All from a zero state; the
    batch's seismograms are checked against the sequential run of the same
    arithmetic (gather for the one-launch-per-step scatter, else colored)."""
    import numpy as np

    from specfempp_amd import Batch, Engine2D, meshes

    ws = meshes.config4(nsim=nsim)
    seq, seq_s, setup_s = {}, {}, 0.0
    for scatter in ("colored", "gather"):
        t = time.perf_counter()
        engines = [Engine2D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, ("displacement",), scatter=scatter)
                   for w in ws]
        setup_s += time.perf_counter() - t
        t = time.perf_counter()
        seq[scatter] = []
        for e in engines:
            e.initialize()
            e.run()
            seq[scatter].append(e.seismograms())
        seq_s[scatter] = time.perf_counter() - t
        if scatter == "gather":
            break
        for e in engines:
            e.close()
    ptsteps = float(sum(e.nglob * w.nstep for e, w in zip(engines, ws)))
    b = Batch(engines)
    choice = b.tune(tune_steps)
    t = time.perf_counter()
    b.run()
    batch_wall_s = time.perf_counter() - t
    info = b.info()
    ref = seq["gather" if info["scatter"] == "gather" else "colored"]
    worst = 0.0
    for i in range(nsim):
        got = b.seismograms(i).astype(np.float64)
        r = ref[i].astype(np.float64)
        worst = max(worst, float(np.linalg.norm(got - r) / max(np.linalg.norm(r), 1e-30)))
    b.close()
    sum_nglob = int(sum(e.nglob for e in engines))
    for e in engines:
        e.close()
    best_seq = min(seq_s.values())
    return {"workload": f"{nsim} 2-D simulations, 40^2..240^2 quads of 50 m, seed 4",
            "point_steps": ptsteps, "sum_nglob": sum_nglob,
            "batched_pt_steps_per_s": ptsteps / batch_wall_s, "batched_s": batch_wall_s,
            "batched_device_ms": info["run_ms"], "launches_per_step": info["launches_per_step"],
            "sequential_pt_steps_per_s": ptsteps / best_seq, "sequential_s": best_seq,
            "sequential_s_by_path": {k: round(v, 3) for k, v in seq_s.items()},
            "speedup": best_seq / batch_wall_s, "tuned": choice, "max_rel_l2_vs_sequential": worst,
            "setup_s": round(setup_s, 1)}


def adjoint_leg(snapshot_every: int = 55) -> dict:
    """Config 5: 200^2 2-D gradient of the L2 waveform misfit (8 sources, 100
    receivers, 3000 steps, checkpoints every 55 steps) against a plain forward
    run of the same model: forward record, combined adjoint + backward loop."""
    import numpy as np

    from specfempp_amd import Engine2D, meshes

    true, start = meshes.config5()
    ty = ("displacement",)
    et = Engine2D(true.mesh, true.sources, true.receivers, true.dt, true.nstep, 1, ty)
    et.initialize()
    et.run()
    data = et.seismograms()[:, 0].astype(np.float64)
    et.close()
    eng = Engine2D(start.mesh, start.sources, start.receivers, start.dt, start.nstep, 1, ty)
    eng.initialize()
    eng.run()  # warm-up
    eng.initialize()
    t = time.perf_counter()
    eng.run()
    plain_s = time.perf_counter() - t
    eng.initialize()
    t = time.perf_counter()
    eng.forward_record(snapshot_every)
    record_s = time.perf_counter() - t
    u = eng.seismograms()[:, 0].astype(np.float64)
    res = u - data
    misfit = 0.5 * float(np.sum(res ** 2)) * start.dt
    eng.set_adjoint_sources(start.receivers, np.transpose(res, (1, 0, 2)))
    t = time.perf_counter()
    eng.adjoint_run()
    adjoint_s = time.perf_counter() - t
    K = eng.kernels()
    finite = bool(all(np.isfinite(v).all() for v in K.values()))
    pts = eng.nglob * start.nstep
    eng.close()
    return {"workload": "2-D 200^2 quads of 20 m, checkerboard +-5 %, 8 sources, 100 receivers, 3000 steps",
            "plain_forward_s": plain_s, "forward_record_s": record_s, "adjoint_run_s": adjoint_s,
            "adjoint_over_forward": adjoint_s / plain_s,
            "gradient_over_forward": (record_s + adjoint_s) / plain_s,
            "combined_pt_steps_per_s": pts / adjoint_s, "forward_pt_steps_per_s": pts / plain_s,
            "misfit": misfit, "snapshot_every": snapshot_every, "kernels_finite": finite}


GEOMETRY_BYTES_PER_ELEMENT_2D = {"point": 25 * 20, "element": 32, "control_nodes_4": 4 * 8, "control_nodes_9": 9 * 8}


def algorithmic_bytes_2d(nspec: int, nglob: int, geometry: str = "point", element_properties: bool = False,
                         scatter: str = "colored") -> dict:
    """Bytes each kernel of the 2-D path must move per step (FP32):
      stiffness_2d (all colors): gather index, kappa, mu per local point
                   (12 B x 25 / element) and the geometry (the engine's
                   layout): xix, xiz, gammax, gammaz, J per local point
                   (20 B x 25 / element), one 32-byte record per element
                   (affine elements), or the element's 4 or 9 control nodes
                   (8 B each) from which the gather path recomputes it; the
                   displacement gathered once per point (8 B), the
                   acceleration read-modify-written once per point (16 B)
      update_fused: a, M^-1, u, v read (32 B) and u, v, a written (24 B) per point
    With per-point geometry this is the data-model figure of 130 B per
    point-step (25/16 local points per point); model_per_pt_step keeps it.
    The gather path instead stores the element forces (8 B per local point)
    and its update reads them once with the inverse index (4 B per local
    point and 4 B per point), and u, v, M^-1 (24 B), writes u, v (16 B; the
    acceleration is recomputed only when a caller asks for it):
      elemforce_2d: gather index per local point, kappa and mu per local
                    point or (element_properties) one pair per element, the
                    geometry, the displacement (8 B per point), the forces
                    written
      assemble_update_2d: index and forces read, the update as above."""
    geo = nspec * GEOMETRY_BYTES_PER_ELEMENT_2D[geometry]
    props = nspec * 8 if element_properties else nspec * 25 * 8
    if scatter == "gather":
        stiff = nspec * 25 * (4 + 8) + props + geo + nglob * 8
        upd = nspec * 25 * (8 + 4) + nglob * (4 + 24 + 16)
    else:
        stiff = nspec * 25 * 4 + props + geo + nglob * 24
        upd = nglob * 56
    model = nspec * 25 * 32 + nglob * 24 + nglob * 56  # per-point geometry, colored scatter, fused update
    return {"stiffness_2d": stiff, "update_fused": upd, "per_pt_step": (stiff + upd) / nglob,
            "model_per_pt_step": model / nglob}


def config2_leg(n: int, steps: int, warmup: int, sample_steps: int = 5, scatter: str = "gather") -> dict:
    """Config 2: the heterogeneous 2-D workload at n^2 quads of 2.5 m (full
    size n = 4000).  Graph-launched steps are timed for the throughput; a few
    single steps give the per-kernel times for the roofline."""
    import numpy as np
    import torch

    from specfempp_amd import Engine2D, meshes

    t = time.perf_counter()
    w = meshes.config2(n=n, L=2.5 * n, nstep=warmup + steps + sample_steps)
    mesh_s = time.perf_counter() - t
    print(f"[config2] mesh {n}^2 built in {mesh_s:.1f} s", file=sys.stderr, flush=True)
    t = time.perf_counter()
    eng = Engine2D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, ("displacement",), scatter=scatter)
    setup_s = time.perf_counter() - t
    print(f"[config2] engine ({eng.nglob} points) built in {setup_s:.1f} s", file=sys.stderr, flush=True)
    nglob, nspec = eng.nglob, eng.nspec
    eng.initialize()
    if warmup:
        eng.run(0, warmup)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(warmup, steps)
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    acc = {"stiffness_ms": 0.0, "update_ms": 0.0, "boundary_source_seismo_ms": 0.0, "step_ms": 0.0}
    for i in range(sample_steps):
        eng.run(warmup + steps + i, 1)
        for k, v in eng.timings().items():
            acc[k] += v / sample_steps
    finite = bool(np.isfinite(eng.seismograms()).all())
    lay = eng.layout()
    eng.close()
    value = nglob * steps / elapsed
    # the bytes of the layout the engine chose (geometry source, properties
    # per point or per element)
    ab = algorithmic_bytes_2d(nspec, nglob, lay["geometry"], lay["element_properties"], scatter=scatter)
    if scatter == "gather":
        kern = {"elemforce_2d_kernel": (ab["stiffness_2d"], acc["stiffness_ms"]),
                "assemble_update_2d_kernel": (ab["update_fused"], acc["update_ms"])}
    else:
        kern = {"stiffness_2d_kernel": (ab["stiffness_2d"], acc["stiffness_ms"]),
                "update_fused_kernel": (ab["update_fused"], acc["update_ms"])}
    dominant = max(kern, key=lambda k: kern[k][1])
    nbytes, ms = kern[dominant]
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    step_gbs = ab["per_pt_step"] * value / 1e9
    return {
        "metric": "GLL pt-steps/s, 2D elastic heterogeneous NGLL=5 FP32 forward (Gaussian random field)",
        "value": value, "unit": "GLL pt-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"2d_grf_{n}x{n}_ngll5", "nspec": nspec, "nglob": nglob, "h": 2.5, "dt": w.dt,
                   "dt_stated": w.dt_stated, "nrec": len(w.receivers), "mesh_s": round(mesh_s, 1),
                   "setup_s": round(setup_s, 1)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant,
                     "algorithmic_bytes_per_launch": nbytes, "avg_ms": ms},
        "kernels": {k: {"avg_ms": v[1], "algorithmic_bytes": v[0],
                        "GBs": (v[0] / (v[1] * 1e-3) / 1e9 if v[1] > 0 else 0.0)} for k, v in kern.items()},
        "kernels_ms": acc, "step_hbm_gbs": step_gbs, "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "algorithmic_bytes_per_pt_step": ab["per_pt_step"],
        "model_bytes_per_pt_step": ab["model_per_pt_step"],
        "model_step_hbm_frac": ab["model_per_pt_step"] * value / 1e9 / HBM_PEAK_GBS,
        "layout": lay, "scatter": scatter, "seismograms_finite": finite,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--scatter", default="auto", choices=["auto", "block", "gather", "colored", "atomic"])
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-config2", action="store_true",
                    help="skip the config-2 leg (2-D 4000^2 heterogeneous, ~30 s of host setup)")
    ap.add_argument("--c2-steps", type=int, default=40, help="timed steps of the config-2 leg")
    ap.add_argument("--batch-sims", type=int, default=64)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (scale reference for HBM byte counters)")
    ap.add_argument("--config2", action="store_true",
                    help="measure the 2-D heterogeneous workload instead (minutes of host setup at full size)")
    ap.add_argument("--c2-n", type=int, default=4000, help="--config2: quads per side (2.5 m each)")
    ap.add_argument("--c2-scatter", default="gather", choices=["gather", "colored", "atomic"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine is single-GPU).",
              file=sys.stderr)
        return 2
    if args.config2:
        print(json.dumps(config2_leg(args.c2_n, args.steps, args.warmup, scatter=args.c2_scatter)))
        return 0

    import numpy as np
    import torch

    from specfempp_amd import Engine3D, meshes

    w = meshes.config3(n=args.n, nstep=args.warmup + args.steps)
    t = time.perf_counter()
    eng = Engine3D(w.mesh, w.sources, w.receivers, w.dt, w.nstep, 1, ("displacement",), scatter=args.scatter)
    setup_s = time.perf_counter() - t
    scatter = eng.scatter
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
    lay = eng.layout()
    seis = eng.seismograms()
    finite = bool(np.isfinite(seis).all())
    nglob, nspec = eng.nglob, eng.nspec
    eng.close()

    ms_per_step = 1e3 * elapsed / args.steps
    value = nglob * args.steps / elapsed
    if scatter == "block":
        ab = algorithmic_bytes_block(args.n)
        kern = {"block_step_3d_kernel": (ab["block_step_3d_kernel"], tm["stiffness_ms"]),
                "surface_update_3d_kernel": (ab["surface_update_3d_kernel"], tm["update_ms"])}
    else:
        # the GATHER update reads one M^-1 component off the absorbing faces; the
        # colored and atomic updates read all three
        minv = minv_bytes_3d(args.n, args.n, args.n, w.mesh["absorbing"]) if scatter == "gather" else None
        ab = algorithmic_bytes_3d(nspec, nglob, minv, element_properties=lay["element_properties"])
        kern = {"elemforce_3d_kernel": (ab["elemforce"], tm["stiffness_ms"]),
                "assemble_update_3d_kernel": (ab["update"], tm["update_ms"])}
    dominant = max(kern, key=lambda k: kern[k][1])
    nbytes, ms = kern[dominant]
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    per_kernel = {k: {"avg_ms": v[1], "algorithmic_bytes": v[0],
                      "GBs": (v[0] / (v[1] * 1e-3) / 1e9 if v[1] > 0 else 0.0)} for k, v in kern.items()}
    step_gbs = ab["per_pt_step"] * value / 1e9
    survey_gbs = ab["survey_per_pt_step"] * value / 1e9
    valu = None
    if scatter != "block" and tm["stiffness_ms"] > 0:
        flops = elemforce_3d_flops(nspec)
        tflops = flops / (tm["stiffness_ms"] * 1e-3) / 1e12
        valu = {"kernel": "elemforce_3d_kernel", "algorithmic_flops_per_launch": flops, "achieved": tflops,
                "unit": "TFLOP/s", "peak": FP32_VALU_PEAK_TFLOPS, "frac": tflops / FP32_VALU_PEAK_TFLOPS,
                "peak_unpacked": FP32_VALU_UNPACKED_TFLOPS, "frac_unpacked": tflops / FP32_VALU_UNPACKED_TFLOPS,
                "counted_flops_per_launch": measured_valu_flops(f"3d_layered_cube_{args.n}^3_ngll5", scatter,
                                                                "elemforce_3d_kernel"),
                "counted_source": "profiles/gather3d_64cube_valu.json (rocprofv3 SQ_INSTS_VALU_*_F32)"}
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
                   "scatter": scatter, "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"3d_layered_cube_{args.n}^3_ngll5", scatter, dominant),
                     "traffic_source": "profiles/gather3d_64cube_pmc.json (rocprofv3 FETCH_SIZE, WRITE_SIZE)",
                     "kernel": dominant, "algorithmic_bytes_per_launch": nbytes, "avg_ms": ms},
        "kernels": per_kernel,
        "kernels_ms": tm,
        "valu": valu,
        "step_hbm_gbs": step_gbs,
        "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "algorithmic_bytes_per_pt_step": ab["per_pt_step"],
        "survey_bytes_per_pt_step": ab["survey_per_pt_step"],
        "survey_step_hbm_frac": survey_gbs / HBM_PEAK_GBS,
        "layout": lay,
        "seismograms_finite": finite,
    }
    # the practical HBM ceiling beside the nominal peak: a plain 1 GiB device copy
    cc = copy_calibration()
    out["roofline"]["copy_GBs"] = cc["copy_GBs"]
    out["roofline"]["frac_of_copy"] = achieved / cc["copy_GBs"] if cc["copy_GBs"] > 0 else None
    if args.calibrate_copy:
        out["copy_calibration"] = cc
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_steps)
    if not args.no_batch:
        out["batched"] = batched_leg(args.batch_sims, 5)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
    if not args.no_config2:
        # the 2-D heterogeneous configuration at full size (256 M points), gather path
        out["config2"] = config2_leg(args.c2_n, args.c2_steps, 3, scatter=args.c2_scatter)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
