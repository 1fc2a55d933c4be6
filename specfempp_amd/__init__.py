"""specfempp_amd — MI355X-native SPECFEM++ Newmark engine.

The compute path is the HIP library built from ``csrc/`` (C ABI in
``include/specfempp_amd.h``); this package is its host-side binding:

* :mod:`specfempp_amd.engine` — ``Engine2D`` / ``Engine3D`` (solver seams),
  ``Batch`` (many 2-D simulations in the same launches)
* :mod:`specfempp_amd.autograd` — differentiable 2-D forward model (adjoint run)
* :mod:`specfempp_amd.meshes` — structured mesh generators for the benchmark
  configurations
* :mod:`specfempp_amd.io` — reader of the 2-D SPECFEM mesh database
"""
from ._native import build, load  # noqa: F401
from .engine import ACCELERATION, DISPLACEMENT, VELOCITY, Batch, Engine2D, Engine3D, SfxError  # noqa: F401

__version__ = "0.1.0"
