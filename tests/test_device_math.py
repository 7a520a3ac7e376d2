"""Host-side checks of the device helpers in chllmodel_amd/csrc/chll_device.h:
fast integer division (exact for all tested numerators), the device Philox
against the published known answers, and the reciprocal-corrected quotient
against IEEE division.  Compiles tests/device_math_check.hip with hipcc for the
host and runs it (no GPU needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "device_math_check.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_device_math_helpers(tmp_path):
    exe = str(tmp_path / "device_math_check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    f"-I{os.path.join(ROOT, 'chllmodel_amd', 'csrc')}", SRC, "-o", exe],
                   check=True, capture_output=True)
    res = subprocess.run([exe, "20000000"], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0 and "OK" in res.stdout, res.stdout
