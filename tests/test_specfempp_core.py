"""The drop-in ``specfempp_core`` package (python/specfempp_core) keeps the
reference module's names and session semantics (src/python/core.cpp:13-78):
_execute refuses to run before _initialize, _initialize and _finalize return
False when repeated, and the defaults file exists."""
import os
import sys

import yaml

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_core_names_and_session():
    sys.path.insert(0, os.path.join(ROOT, "python"))
    try:
        import specfempp_core as core
    finally:
        sys.path.pop(0)
    for name in ("__version__", "__doc__", "_default_file_path", "_initialize", "_execute", "_finalize"):
        assert hasattr(core, name), name
    assert os.path.exists(core._default_file_path)
    defaults = open(core._default_file_path).read()
    assert yaml.safe_load(defaults)["default-parameters"]["quadrature"]["quadrature-type"] == "GLL4"
    assert core._execute("parameters: {}", defaults) is False
    assert core._finalize() is False
    assert core._initialize(["specfem"]) is True
    assert core._initialize(["specfem"]) is False
    assert core._finalize() is True
    assert core._finalize() is False
