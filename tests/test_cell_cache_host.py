"""The register cell cache of the advect kernels, built for the host.

dpa_amd/csrc/device_math.hpp keeps the corners of a particle's last cell in
registers and, when the particle moves into a face neighbour, keeps the four
shared corners and loads four.  tests/native/cell_cache_check.cpp compiles that
header for the CPU and checks that every sample through a warm cache equals the
sample through an empty one, bit for bit, over random walks on steady and
time-variant grids (a wrong corner shuffle would change the interpolated value).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_warm_cache_equals_full_reload(tmp_path):
    exe = tmp_path / "cell_cache_check"
    src = os.path.join(HERE, "native", "cell_cache_check.cpp")
    subprocess.run([HIPCC, "-O2", "-ffp-contract=off", "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-o", str(exe), src], check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    words = out.stdout.split()
    samples, shifts = int(words[1]), int(words[3])
    assert samples > 100000 and shifts > 10000, out.stdout
