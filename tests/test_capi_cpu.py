"""The C-ABI library builds, loads and exports every declared symbol (no GPU needed)."""
import ctypes
import os
import subprocess

import pytest

from specfempp_amd import _native


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    declared = _native.declared_symbols()
    assert len(declared) >= 20
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported


def test_version_and_error_string():
    lib = _native.load()
    assert lib.sfx_version() >= 100
    assert isinstance(lib.sfx_last_error(), bytes)


def test_null_arguments_are_rejected_without_touching_the_gpu():
    lib = _native.load()
    out = ctypes.c_void_p()
    rc = lib.sfx_engine_create_2d(None, 0, None, 0, None, None, ctypes.byref(out))
    assert rc == -1
    assert b"null" in lib.sfx_last_error()
    sizes = (ctypes.c_int64 * 7)()
    assert lib.sfx_get_sizes(None, sizes) == -1


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU failure path")
def test_engine_fails_loudly_without_gpu():
    from specfempp_amd import Engine2D, SfxError, meshes

    w = meshes.config1(nstep=4, n=4, h=1000.0)
    with pytest.raises(SfxError):
        Engine2D(w.mesh, w.sources, w.receivers, w.dt, w.nstep)
