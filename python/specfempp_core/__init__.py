"""Drop-in for the reference's ``specfempp_core`` package
(python/specfempp_core/__init__.py re-exporting the pybind11 ``_core``): the
same names, served by the MI355X engine (``specfempp_amd.core``).  Put this
directory on ``sys.path`` in place of the reference's build."""
import os
import sys

_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)

from specfempp_amd.core import (  # noqa: E402
    __doc__,
    __version__,
    _default_file_path,
    _execute,
    _finalize,
    _initialize,
)

__all__ = ["__version__", "__doc__", "_default_file_path", "_initialize", "_execute", "_finalize"]
