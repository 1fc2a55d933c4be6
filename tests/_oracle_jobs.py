"""Oracle runs for the GPU tests, executed in CPU worker processes (spawned,
so they never touch the GPU) so that independent oracle runs of one test
proceed in parallel on the host's cores."""
from __future__ import annotations

import concurrent.futures as cf
import multiprocessing as mp
import os

import numpy as np


def host_workers(cap: int = 16) -> int:
    return max(1, min(cap, len(os.sched_getaffinity(0))))


def pool(n: int | None = None) -> cf.ProcessPoolExecutor:
    return cf.ProcessPoolExecutor(max_workers=n or host_workers(), mp_context=mp.get_context("spawn"))


def forward_2d(mesh, prec, sources, receivers, nstep, dt, types=("displacement",), props=None, nsb=1):
    """Seismograms (nsig, ntypes, nrec, 2) of one oracle forward run;
    props = (rho, kappa, mu) per point (mesh order) replaces the mesh's."""
    import oracle

    o = oracle.Oracle2D(mesh, prec, 1, dt)
    if props is not None:
        o.set_properties(*props)
    return o.forward(sources, receivers, nstep, dt, nsb, types)[0]


def misfit_2d(mesh, sources, receivers, nstep, dt, data, props):
    """0.5 sum (u - d)^2 dt of the FP64 oracle's displacement (the L2
    waveform misfit), for per-point properties props = (rho, kappa, mu)."""
    u = forward_2d(mesh, 8, sources, receivers, nstep, dt, props=props)[:, 0]
    return 0.5 * float(np.sum((u - data) ** 2)) * dt


def adjoint_2d(mesh, prec, sources, receivers, nstep, dt, stf, snapshot_every, types=("displacement",)):
    """Forward record then combined run of the oracle with adjoint sources
    at the receivers: (forward seismograms, kernels (6, nspec, 5, 5), backward
    seismograms)."""
    import oracle

    o = oracle.Oracle2D(mesh, prec, 1, dt)
    seis, _, fields, record, *snaps = o.forward_record(sources, receivers, nstep, dt, 1, types,
                                                       snapshot_every=snapshot_every)
    kern, bseis = o.combined(sources, receivers, stf, receivers, nstep, dt, fields, record, 1, types,
                             snapshot_every=snapshot_every, snapshots=snaps[0] if snaps else None)
    return seis, kern, bseis


def forward_3d(mesh, prec, sources, receivers, nstep, dt, nthreads=1):
    """(seismograms, fields (3, 3, nglob)) of a 3-D oracle forward run."""
    import oracle

    seis, _, fields = oracle.Oracle3D(mesh, prec, dt).forward(sources, receivers, nstep, dt, return_fields=True,
                                                              nthreads=nthreads)
    return seis, fields
