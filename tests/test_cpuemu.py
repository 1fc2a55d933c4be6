"""Host (ASan) build of the HIP kernels: the 3-D GATHER path on a 3^3 cube,
fused and seam calls, against the FP32 oracle, with AddressSanitizer watching
every kernel's accesses.  It takes a few minutes, so it runs only when
SFX_CPUEMU=1 is set."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

pytestmark = pytest.mark.skipif(os.environ.get("SFX_CPUEMU") != "1", reason="set SFX_CPUEMU=1 to run")


def test_gather3d_under_asan(tmp_path):
    lib = str(tmp_path / "libemu.so")
    subprocess.run(["bash", os.path.join(HERE, "cpuemu", "build.sh"), lib], check=True, timeout=1200)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(HERE, "cpuemu", "check_gather3d.py"), lib], env=env,
                       capture_output=True, text=True, timeout=1800)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "ERROR: AddressSanitizer" not in r.stderr
