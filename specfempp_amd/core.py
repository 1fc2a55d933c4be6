"""The entry points of the reference's Python module ``specfempp_core._core``
(src/python/core.cpp:80-112), over this library:

* ``_initialize(argv)`` loads the HIP library (the reference initializes MPI
  and Kokkos); True on the first call, False when already initialized.
* ``_execute(parameter_string, default_string)``: the YAML text of a
  ``specfem_config.yaml`` and of the defaults file (``_default_file_path``);
  runs the simulation it describes (src/execute.cpp: mesh database, sources,
  receivers, Newmark loop, seismogram writer).  False before ``_initialize``.
* ``_finalize()``: True when it ends an initialized session.

Relative paths in the configuration resolve against the current directory,
as the reference's do.
"""
from __future__ import annotations

import os

import yaml

__version__ = "0.2.0"
_default_file_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "io", "default.yaml")

_session = None


def _initialize(argv) -> bool:
    global _session
    if _session is not None:
        return False
    from . import _native

    lib = _native.load()
    _session = {"argv": [str(a) for a in argv], "version": lib.sfx_version()}
    return True


def _execute(parameter_string: str, default_string: str) -> bool:
    if _session is None:
        return False
    from .io import config

    doc = config.merge_defaults(yaml.safe_load(parameter_string), yaml.safe_load(default_string))
    run = config.parse_config(doc, os.getcwd())
    if run.simulation == "forward":
        config.run_config(run, out_dir=run.writer_directory)
    else:
        config.run_combined_config(run)
    return True


def _finalize() -> bool:
    global _session
    if _session is None:
        return False
    _session = None
    return True
