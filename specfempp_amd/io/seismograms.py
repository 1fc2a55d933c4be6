"""Seismogram read-out in the reference's station format.

The engine returns seismograms as recorded, (nsig, ntypes, nrec, 2) before
station rotation (the layout of the reference's ``seismogram_components``).
This module applies the read-out the reference does when it writes them:

* rotation by each station's angle and the sample times
  (``include/compute/receivers/receivers.hpp:126-141``: time = isig * dt *
  nstep_between_samples + t0; x' = cos a x - sin a z, z' = sin a x + cos a z,
  with sin and cos of pi / 180 * angle in single precision as
  ``compute_receivers.cpp:81-85`` forms them);
* ASCII files ``NET.STA.S2.BX{X,Z}.sem{d,v,a}`` and ``NET.STA.S2.PRE.semp``
  with "time value" lines in scientific notation
  (``src/io/seismogram/writer.cpp:23-88``).
"""
from __future__ import annotations

import os

import numpy as np

_EXT = {"displacement": "semd", "velocity": "semv", "acceleration": "sema", "pressure": "semp"}


def station_traces(seis, t0: float, dt: float, nstep_between_samples: int, angles):
    """Rotated traces (nsig, ntypes, nrec, 2) as float32 and the times (nsig,)."""
    seis = np.asarray(seis, dtype=np.float32)
    ang = np.asarray(angles, dtype=np.float32)
    pi = np.float32(np.pi)
    s = np.sin(pi / np.float32(180) * ang).astype(np.float32)
    c = np.cos(pi / np.float32(180) * ang).astype(np.float32)
    x, z = seis[..., 0], seis[..., 1]
    out = np.empty_like(seis)
    out[..., 0] = c * x - s * z
    out[..., 1] = s * x + c * z
    nsig = seis.shape[0]
    times = (np.arange(nsig, dtype=np.float32) * np.float32(dt) * np.float32(nstep_between_samples)
             + np.float32(t0))
    return out, times


def write_ascii(folder: str, stations, seis, types, t0: float, dt: float, nstep_between_samples: int = 1):
    """Write one file per station, type and component.  stations: sequence of
    (name, network, angle).  Returns the paths written."""
    os.makedirs(folder, exist_ok=True)
    traces, times = station_traces(seis, t0, dt, nstep_between_samples, [s[2] for s in stations])
    paths = []
    for it, t in enumerate(types):
        for r, (name, network, _) in enumerate(stations):
            if t == "pressure":
                comps = [("PRE", 0)]
            else:
                comps = [("BXX", 0), ("BXZ", 1)]
            for code, ci in comps:
                path = os.path.join(folder, f"{network}.{name}.S2.{code}.{_EXT[t]}")
                with open(path, "w") as fh:
                    for k in range(len(times)):
                        fh.write(f"{float(times[k]):e} {float(traces[k, it, r, ci]):e}\n")
                paths.append(path)
    return paths
