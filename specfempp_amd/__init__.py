"""specfempp_amd — MI355X-native SPECFEM++ Newmark engine.

The compute path is the HIP library built from ``csrc/`` (C ABI in
``include/specfempp_amd.h``); this package is its host-side binding:

* :mod:`specfempp_amd.engine` — ``Engine2D`` (elastic, acoustic and coupled
  meshes), ``Engine3D`` (structured hexahedra), ``Engine3DDatabase`` (3-D mesh
  databases, general hexahedra), ``Batch`` (many 2-D simulations in the same
  launches)
* :mod:`specfempp_amd.autograd` — differentiable 2-D forward model (adjoint run)
* :mod:`specfempp_amd.meshes` — structured mesh generators for the benchmark
  configurations
* :mod:`specfempp_amd.solver` — the reference's time scheme and time loops
  driven seam by seam
* :mod:`specfempp_amd.io` — 2-D and 3-D SPECFEM mesh databases, stations,
  sources, configuration files, seismogram and wavefield/kernel files
"""
from ._native import build, load  # noqa: F401
from .engine import (ACCELERATION, DISPLACEMENT, VELOCITY, Batch, Engine2D, Engine3D,  # noqa: F401
                     Engine3DDatabase, SfxError)

__version__ = "0.1.0"
