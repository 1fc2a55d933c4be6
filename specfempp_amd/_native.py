"""Build and load the HIP C-ABI library ``_lib/libspecfempp_amd.so``.

``build()`` compiles every source under ``csrc/`` for gfx950 with hipcc (no
GPU needed).  ``load()`` only loads what is built; it never falls back to
anything else.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libspecfempp_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "specfempp_amd.h")

SOURCES = ["engine.hip", "kernels.hip", "acoustic.hip", "general3d.hip", "block3d.hip", "batch.hip",
           "assembly.cpp"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off: some of their forms have given wrong
    # results on this hardware without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

_lib = None


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-4000:])
    if verbose:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def stale_sources():
    """Sources (and the header) newer than the built library."""
    if not os.path.exists(LIB_PATH):
        return []
    built = os.path.getmtime(LIB_PATH)
    paths = [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [HEADER]
    return sorted(os.path.basename(p) for p in paths if os.path.getmtime(p) > built)


def load():
    """Load the built library (building it first only if it is absent).  A
    library older than its sources is loaded as it is, with a warning: run
    build() (``__graft_entry__.build()``) after changing a source."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    stale = stale_sources()
    if stale:
        import warnings

        warnings.warn(f"{LIB_PATH} is older than {', '.join(stale)}: rebuild with specfempp_amd.build()",
                      RuntimeWarning, stacklevel=2)
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, dp = C.c_void_p, C.c_int, C.c_int64, C.c_void_p
    sig = {
        "sfx_version": ([], i32),
        "sfx_last_error": ([], C.c_char_p),
        "sfx_engine_create_2d": ([vp, i32, vp, i32, vp, vp, vp], i32),
        "sfx_engine_create_3d": ([vp, i32, vp, i32, vp, vp, vp], i32),
        "sfx_engine_create_host_2d": ([vp, i32, vp, i32, vp, vp, vp], i32),
        "sfx_engine_create_host_3d": ([vp, i32, vp, i32, vp, vp, vp], i32),
        "sfx_engine_create_3d_db": ([vp, i32, vp, i32, vp, vp, vp], i32),
        "sfx_engine_create_host_3d_db": ([vp, i32, vp, i32, vp, vp, vp], i32),
        "sfx_engine_destroy": ([vp], None),
        "sfx_get_sizes": ([vp, vp], i32),
        "sfx_get_media_sizes": ([vp, vp], i32),
        "sfx_get_layout": ([vp, vp], i32),
        "sfx_get_field_index_acoustic": ([vp, vp], i32),
        "sfx_get_medium_field": ([vp, i32, i32, vp], i32),
        "sfx_set_medium_field": ([vp, i32, i32, vp], i32),
        "sfx_get_t0": ([vp, vp], i32),
        "sfx_get_index_mapping": ([vp, vp], i32),
        "sfx_get_field_index": ([vp, vp], i32),
        "sfx_get_compute_to_mesh": ([vp, vp], i32),
        "sfx_get_locations": ([vp, i32, vp, vp], i32),
        "sfx_initialize": ([vp], i32),
        "sfx_predictor": ([vp], i32),
        "sfx_update_wavefields": ([vp, i32], i32),
        "sfx_corrector": ([vp], i32),
        "sfx_compute_seismograms": ([vp, i32], i32),
        "sfx_initialize_wavefield": ([vp, i32], i32),
        "sfx_predictor_phase": ([vp, i32, i32, i32], i32),
        "sfx_update_wavefields_medium": ([vp, i32, i32, i32], i32),
        "sfx_corrector_phase": ([vp, i32, i32, i32], i32),
        "sfx_compute_seismograms_wavefield": ([vp, i32, i32], i32),
        "sfx_copy_wavefield": ([vp, i32, i32], i32),
        "sfx_compute_derivatives": ([vp, C.c_double], i32),
        "sfx_get_wavefield": ([vp, i32, i32, vp], i32),
        "sfx_set_wavefield": ([vp, i32, i32, vp], i32),
        "sfx_step": ([vp, i32], i32),
        "sfx_run": ([vp, i32, i32], i32),
        "sfx_run_noalloc": ([vp, i32, i32], i32),
        "sfx_synchronize": ([vp], i32),
        "sfx_get_seismograms": ([vp, vp], i32),
        "sfx_get_field": ([vp, i32, vp], i32),
        "sfx_set_field": ([vp, i32, vp], i32),
        "sfx_get_timings": ([vp, vp], i32),
        "sfx_forward_record": ([vp, i32], i32),
        "sfx_set_properties": ([vp, vp, vp, vp], i32),
        "sfx_set_adjoint_sources": ([vp, i32, vp, vp], i32),
        "sfx_adjoint_run": ([vp], i32),
        "sfx_get_kernels": ([vp, vp], i32),
        "sfx_get_point_weights": ([vp, vp], i32),
        "sfx_get_point_coordinates": ([vp, vp], i32),
        "sfx_get_stacey_record": ([vp, vp], i32),
        "sfx_set_stacey_record": ([vp, vp], i32),
        "sfx_batch_create": ([i32, vp, vp], i32),
        "sfx_batch_destroy": ([vp], None),
        "sfx_batch_tune": ([vp, i32, vp], i32),
        "sfx_batch_set_config": ([vp, i32, i32, i32], i32),
        "sfx_batch_run": ([vp], i32),
        "sfx_batch_get_seismograms": ([vp, i32, vp], i32),
        "sfx_batch_get_info": ([vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def declared_symbols():
    """Function names declared in include/specfempp_amd.h."""
    import re

    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sfx_[a-z0-9_]+)\s*\(", text)))
