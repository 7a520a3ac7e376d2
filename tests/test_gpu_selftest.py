"""GPU checks of device helpers that rest on hardware instructions.

accept_exp (chllmodel_amd/csrc/chll_device.h) decides the Metropolis test
u <= exp(x) from an fp32 estimate of exp(x) built on v_exp_f32 and falls back
to the binary64 exp only inside an error bracket.  Its decisions equal the
binary64 test only if the bracket holds exp(x); this measures the estimate's
error against the margin over x in [-90, 0] and checks the decisions at and
around the bracket edges and at u = exp(x) itself."""
import ctypes as ct
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "libdevice_selftest.so")


@pytest.mark.gpu
def test_accept_exp_bracket():
    if not os.path.exists(SO):
        pytest.fail(f"{SO} not built: run __graft_entry__.build()")
    lib = ct.CDLL(SO)
    lib.chll_selftest_exp.argtypes = [ct.c_int, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)]
    n = 1 << 20
    ratio = np.zeros(n)
    bad = np.zeros(n, dtype=np.int32)
    print(f"accept_exp bracket check over {n} points", flush=True)
    assert lib.chll_selftest_exp(n, ratio.ctypes.data_as(ct.POINTER(ct.c_double)),
                                 bad.ctypes.data_as(ct.POINTER(ct.c_int))) == 0
    print(f"worst error / margin {ratio.max():.4f}, failed checks {int(bad.sum())}")
    assert int(bad.sum()) == 0
    # the assumed error bound is 5x below the margin; the hardware must stay under it
    assert ratio.max() < 0.2


@pytest.mark.gpu
def test_bandwidth_probe_rates():
    """chll_bandwidth_probe, which the bench prices the sweep kernels against:
    every pattern runs and reports a rate between 1 TB/s and twice the 8 TB/s
    peak, the best pass at least the median, and a pass that only reads is
    not much slower per byte than the copy that also writes (measured 6-10 %
    faster; the bound leaves room for box-to-box spread)."""
    import torch  # noqa: F401  (one HIP runtime: torch first, as the library's users do)

    from chllmodel_amd._lib import BW_PATTERNS, bandwidth_probe

    rates = {}
    for name in BW_PATTERNS:
        print(f"bandwidth probe {name} on 2 x 512 MiB", flush=True)
        best, med = bandwidth_probe(name, 512, 5)
        print(f"  best {best:.0f} GB/s, median {med:.0f} GB/s", flush=True)
        assert 1000.0 < med <= best < 16000.0
        rates[name] = best
    assert rates["read"] > 0.9 * rates["copy"]
