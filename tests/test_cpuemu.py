"""Host (ASan + bounds) build of the HIP kernels, with AddressSanitizer
watching every kernel's memory accesses and -fsanitize=bounds every index
into a fixed-size array (kernel-argument structs included), against the FP32
oracle:
- the 3-D GATHER path on a 3^3 cube and the 3-D BLOCK path on a 4^3 cube,
  fused and seam calls;
- the 2-D and 3-D call sequences of the seam, combined-loop, checkpoint,
  snapshot, moment-tensor, storage-chunk and batch GPU tests at small sizes;
- the 3-D mesh-database path (general hexahedra) on a small irregular mesh;
- the acoustic and coupled elastic-acoustic fixtures for a few steps (the
  regression case of an empty elastic receiver list, which once launched the
  elastic seismogram kernel on acoustic receivers).
It takes tens of minutes, so it runs only when SFX_CPUEMU=1 is set."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

pytestmark = pytest.mark.skipif(os.environ.get("SFX_CPUEMU") != "1", reason="set SFX_CPUEMU=1 to run")


@pytest.fixture(scope="module")
def emu_lib(tmp_path_factory):
    lib = str(tmp_path_factory.mktemp("emu") / "libemu.so")
    subprocess.run(["bash", os.path.join(HERE, "cpuemu", "build.sh"), lib], check=True, timeout=1200)
    return lib


@pytest.mark.parametrize("check", ["check_gather3d.py", "check_block3d.py", "check_paths2d.py",
                                   "check_acoustic2d.py", "check_database3d.py"])
def test_3d_paths_under_asan(emu_lib, check):
    lib = emu_lib
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    extra = ["2"] if check == "check_acoustic2d.py" else []  # steps of each acoustic run
    r = subprocess.run([sys.executable, os.path.join(HERE, "cpuemu", check), lib] + extra, env=env,
                       capture_output=True, text=True, timeout=3600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
