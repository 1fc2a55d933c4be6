"""Run the call sequences of the GPU tests on a host (ASan) build of the
library at small sizes: the forward loop seam by seam (2-D colored and
gather, 3-D gather), the combined loop seam by seam and fused, the Stacey
record and point-coordinate getters and setters, wavefield snapshots between
fused steps (2-D colored and gather, 3-D gather, colored, atomic), a 2-D
moment-tensor source, storage chunks 2 and 64, and a small batch.  Checks
against the FP32 oracle where a result exists there.
Usage: check_paths2d.py LIB.so"""
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
from conftest import rel_l2  # noqa: E402
from specfempp_amd import Batch, Engine2D, Engine3D, meshes  # noqa: E402
from specfempp_amd.solver import Newmark, TimeMarching  # noqa: E402

failures = []


def check(name, ok, detail=""):
    print(f"{'ok  ' if ok else 'FAIL'} {name} {detail}", flush=True)
    if not ok:
        failures.append(name)


N, H, DT, NSTEP = 6, 50.0, 1e-3, 8
TYPES = ("displacement", "velocity", "acceleration")
SRC = [dict(x=110.0, z=160.0, angle=30.0, f0=15.0, factor=1e10)]
REC = [dict(x=200.0, z=250.0), dict(x=80.0, z=100.0)]


def mesh2d(n=N):
    L = n * H
    return meshes.quad_mesh(n, n, 0, L, 0, L, materials=[meshes.elastic(2400.0, 3000.0, 1732.0)],
                            absorbing=("bottom", "right", "left", "top"))


mesh = mesh2d()
ref, _ = oracle.Oracle2D(mesh, 4, 1, DT).forward(SRC, REC, NSTEP, DT, 1, TYPES)

# forward loop seam by seam against the fused run (2-D colored, gather; 3-D gather)
for scatter in ("colored", "gather"):
    a = Engine2D(mesh, SRC, REC, DT, NSTEP, 2, TYPES, scatter=scatter)
    a.initialize()
    a.run()
    fused = a.seismograms().astype(np.float64)
    a.close()
    b = Engine2D(mesh, SRC, REC, DT, NSTEP, 2, TYPES, scatter=scatter)
    ts = Newmark(NSTEP, 2, DT, "forward")
    ts.link_assembly(b)
    TimeMarching(b, ts).run()
    seam = b.seismograms().astype(np.float64)
    b.close()
    check(f"seams 2-D {scatter}", rel_l2(seam, fused) <= 1e-5 and rel_l2(fused[:, 0], ref[::2, 0]) <= 1e-5,
          f"{rel_l2(seam, fused):.2e} {rel_l2(fused[:, 0], ref[::2, 0]):.2e}")

w3 = meshes.config3(n=3, nstep=NSTEP, h=100.0, nrec_side=2)
a = Engine3D(w3.mesh, w3.sources, w3.receivers, w3.dt, NSTEP, 2, TYPES, scatter="gather")
a.initialize()
a.run()
fused3 = a.seismograms().astype(np.float64)
a.close()
b = Engine3D(w3.mesh, w3.sources, w3.receivers, w3.dt, NSTEP, 2, TYPES, scatter="gather")
ts = Newmark(NSTEP, 2, w3.dt, "forward")
ts.link_assembly(b)
TimeMarching(b, ts).run()
check("seams 3-D gather", rel_l2(b.seismograms().astype(np.float64), fused3) <= 1e-5)
b.close()

# combined loop: fused adjoint run, then seam by seam from the record, the
# Stacey record and the coordinates through the getters and setters
rng = np.random.default_rng(7)
stf = 1e-6 * rng.standard_normal((len(REC), NSTEP, 2))
fused = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, ("displacement",), scatter="colored")
fused.initialize()
fused.forward_record(0)
record = fused.stacey_record()
coords = fused.point_coordinates()
final = [fused.wavefield_field("forward", k) for k in ("displacement", "velocity", "acceleration")]
fused.set_adjoint_sources(REC, stf)
fused.adjoint_run()
k_fused = fused.kernels()
fused.close()
eng = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, ("displacement",), scatter="colored")
eng.initialize()
eng.set_stacey_record(record)
for k, v in zip(("displacement", "velocity", "acceleration"), final):
    eng.set_wavefield_field("buffer", k, v)
eng.set_adjoint_sources(REC, stf)
ts = Newmark(NSTEP, 1, DT, "combined")
ts.link_assembly(eng)
TimeMarching(eng, ts).run()
k_seam = eng.kernels()
eng.close()
worst = max(rel_l2(k_seam[n].astype(np.float64), k_fused[n].astype(np.float64)) for n in k_fused)
check("combined seams vs fused (record through the getters)", worst <= 1e-4 and coords.shape == (N * N, 5, 5, 2),
      f"{worst:.2e}")
# checkpoints every 3 steps
eng = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, ("displacement",), scatter="atomic")
eng.initialize()
eng.forward_record(3)
eng.set_adjoint_sources(REC, stf)
eng.adjoint_run()
check("adjoint run with checkpoints (atomic)", all(np.isfinite(v).all() for v in eng.kernels().values()))
eng.close()

# wavefield snapshots between fused steps (the test_gpu_wavefields sequence)


def snapshots(eng, steps):
    eng.initialize()
    done, out = 0, []
    for k in steps:
        if k - 1 > done:
            eng.run(done, k - 1 - done)
        eng.update_wavefields(k - 1)
        eng.apply_corrector_phase()
        eng.compute_seismograms(k - 1)
        out.append([eng.field(x).astype(np.float64) for x in TYPES])
        eng.apply_predictor_phase()
        done = k
    return out


steps = (3, 5, NSTEP)
for scatter in ("colored", "gather"):
    eng = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, ("displacement",), scatter=scatter)
    snap = snapshots(eng, steps)
    eng.close()
    _, _, f = oracle.Oracle2D(mesh, 4, 1, DT).forward(SRC, REC, NSTEP, DT, 1, ("displacement",), return_fields=True)
    check(f"snapshots 2-D {scatter}", rel_l2(snap[-1][0], f[0]) <= 1e-5, f"{rel_l2(snap[-1][0], f[0]):.2e}")
for scatter in ("gather", "colored", "atomic"):
    eng = Engine3D(w3.mesh, w3.sources, w3.receivers, w3.dt, NSTEP, 1, ("displacement",), scatter=scatter)
    snap = snapshots(eng, steps)
    eng.close()
    _, _, f = oracle.Oracle3D(w3.mesh, 4, w3.dt).forward(w3.sources, w3.receivers, NSTEP, w3.dt, return_fields=True)
    check(f"snapshots 3-D {scatter}", rel_l2(snap[-1][0], f[0]) <= 1e-5, f"{rel_l2(snap[-1][0], f[0]):.2e}")

# 2-D moment tensor; storage chunks 2 and 64
mt = [dict(type="moment_tensor", x=160.0, z=140.0, mxx=1.0, mzz=-0.5, mxz=0.3, f0=15.0, factor=1e10)]
r_mt, _ = oracle.Oracle2D(mesh, 4, 1, DT).forward(mt, REC, NSTEP, DT, 1, ("displacement",))
for scatter in ("colored", "gather"):
    eng = Engine2D(mesh, mt, REC, DT, NSTEP, 1, ("displacement",), scatter=scatter)
    eng.initialize()
    eng.run()
    e = rel_l2(eng.seismograms().astype(np.float64), r_mt)
    eng.close()
    check(f"moment tensor 2-D {scatter}", e <= 1e-5, f"{e:.2e}")
# the reference's 9-node 80 x 60 fixture on the gather path: each point's
# geometry recomputed from the element's 9 control nodes, kappa and mu read
# once per element; the displacement after a few steps against the oracle's
from conftest import load_case  # noqa: E402

m9, cfg9 = load_case("elastic_stacey_80x60")
eng = Engine2D(m9, cfg9["sources"], cfg9["stations"], cfg9["dt"], 6, 1, ("displacement",), scatter="gather")
lay = eng.layout()
d9 = snapshots(eng, (6,))[-1][0]
eng.close()
_, _, f9 = oracle.Oracle2D(m9, 4, 1, cfg9["dt"]).forward(cfg9["sources"], cfg9["stations"], 6, cfg9["dt"], 1,
                                                          ("displacement",), return_fields=True)
check("9-node mesh, gather (geometry from control nodes)",
      lay["geometry"] == "control_nodes_9" and lay["element_properties"] and rel_l2(d9, f9[0]) <= 1e-5
      and np.abs(d9).max() > 0, f"{lay} {rel_l2(d9, f9[0]):.2e}")
for chunk in (2, 64):
    eng = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, ("displacement",), chunk_size=chunk)
    o = oracle.Oracle2D(mesh, 4, chunk, DT)
    ibool, midx, c2m = o.numbering()
    eng.initialize()
    eng.run()
    r, _ = o.forward(SRC, REC, NSTEP, DT, 1, ("displacement",))
    ok = np.array_equal(eng.index_mapping(), ibool) and np.array_equal(eng.field_index(), midx)
    e = rel_l2(eng.seismograms().astype(np.float64), r)
    eng.close()
    check(f"chunk {chunk}", ok and e <= 1e-5, f"{e:.2e}")

# a small batch of unequal meshes and step counts, every scatter: the
# one-launch-per-step scatter against the engines' gather path bit for bit,
# the colored one against their colored path, the atomic one within 1e-6
SIZES = ((4, 5), (6, 8), (3, 8))
single = {}
for scatter in ("gather", "colored"):
    single[scatter] = []
    for n, ns in SIZES:
        e = Engine2D(mesh2d(n), SRC, REC, DT, ns, 1, TYPES, scatter=scatter)
        e.initialize()
        e.run()
        single[scatter].append(e.seismograms())
        e.close()
for scatter, epb, graphs in (("gather", 2, True), ("gather", 5, False), ("gather", 10, True), ("gather", 5, True),
                             ("colored", 2, True), ("atomic", 5, False)):
    engs = [Engine2D(mesh2d(n), SRC, REC, DT, ns, 1, TYPES, host_only=True) for n, ns in SIZES]
    b = Batch(engs)
    for e in engs:
        e.close()
    b.set_config(epb, graphs, scatter)
    b.run()
    ref = single["colored" if scatter == "atomic" else scatter]
    got = [b.seismograms(i) for i in range(len(SIZES))]
    if scatter == "atomic":
        ok = all(rel_l2(got[i].astype(np.float64), ref[i].astype(np.float64)) <= 1e-6 for i in range(len(SIZES)))
    else:
        ok = all(np.array_equal(got[i], ref[i]) for i in range(len(SIZES)))
    info = b.info()
    b.close()
    check(f"batch {scatter} E={epb} graphs={graphs} vs single runs", ok and info["scatter"] == scatter,
          f"launches/step {info['launches_per_step']}")

# a colored batch of engines built with the atomic scatter (one "color" in
# compute order): the batch colors the meshes itself
engs = [Engine2D(mesh2d(n), SRC, REC, DT, ns, 1, TYPES, scatter="atomic", host_only=True) for n, ns in SIZES]
b = Batch(engs)
for e in engs:
    e.close()
b.set_config(5, False, "colored")
b.run()
check("colored batch of atomic-scatter engines", all(np.array_equal(b.seismograms(i), single["colored"][i])
                                                     for i in range(len(SIZES))))
b.close()

# samples every third step, simulations ending at different steps
ref3 = []
for n, ns in SIZES:
    e = Engine2D(mesh2d(n), SRC, REC, DT, ns, 3, TYPES, scatter="gather")
    e.initialize()
    e.run()
    ref3.append(e.seismograms())
    e.close()
engs = [Engine2D(mesh2d(n), SRC, REC, DT, ns, 3, TYPES, host_only=True) for n, ns in SIZES]
b = Batch(engs)
for e in engs:
    e.close()
b.set_config(5, True, "gather")
b.run()
check("batch gather, every third step sampled", all(np.array_equal(b.seismograms(i), ref3[i]) for i in range(len(SIZES))))
b.close()

print("failures:", failures, flush=True)
sys.exit(1 if failures else 0)
