"""Build libchll_mod.so in-tree with hipcc for gfx950.

    python -m chllmodel_amd.build

Flags: -ffp-contract=off keeps the reference-compatible arithmetic free of
fused multiply-adds (the FP64 mode writes its fma() calls explicitly), and the
packed-FP32 target feature is disabled for every HIP translation unit.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libchll_mod.so")
SOURCES = ["chll_kernels.hip", "chll_api.hip", "chll_ensemble.hip", "chll_grad.hip", "chll_march.hip",
           "chll_analysis.hip"]
HEADERS = ["chll_device.h", "chll_kernels.h", "chll_site.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CHLL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def build(out: str = LIB, extra: list[str] | None = None, quiet: bool = True) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC, *FLAGS, f"-I{os.path.join(ROOT, 'include')}", *srcs, "-o", out + ".tmp"]
    if extra:
        cmd[1:1] = extra
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
    if not quiet:
        sys.stderr.write(res.stderr)
    os.replace(out + ".tmp", out)
    return out


def build_selftest() -> str:
    """tests/libdevice_selftest.so: GPU checks of hardware-dependent device
    helpers (tests/device_selftest.hip), built with the library's flags."""
    src = os.path.join(ROOT, "tests", "device_selftest.hip")
    out = os.path.join(ROOT, "tests", "libdevice_selftest.so")
    cmd = [HIPCC, *FLAGS, f"-I{CSRC}", src, "-o", out + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(quiet=False))
    print(build_selftest())
