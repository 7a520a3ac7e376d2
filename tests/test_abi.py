"""C-ABI of libchll_mod.so, checked without a GPU: every symbol that
include/chll_mod.h declares is exported, the host-only entry point ``init``
matches the reference's sublattice lists (chll/chll_mod.f90:9-38, restated by
the oracle), and invalid inputs are rejected with a message before any device
work (chll/chll.py's argtypes reject wrong dtypes/layouts as in the reference).
"""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

from chllmodel_amd import _lib
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chll_mod.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"#.*", "", txt)
    txt = re.sub(r"typedef struct \w+ \{.*?\} \w+;", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{}]*\)\s*;", txt)))


def test_header_symbols_exported():
    names = declared_functions()
    assert "init" in names and "run" in names and len(names) > 20
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n)


@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 3, 5), (8, 8, 8), (6, 1, 1), (3, 2, 2)])
def test_init_lists_match_oracle(dims):
    ni, nj, nk = dims
    N = ni * nj * nk
    nsub = N // 2
    f = lambda dt: np.zeros(dims, dtype=dt, order="F")
    sl1 = np.zeros(nsub, dtype=np.int32)
    sl2 = np.zeros(nsub, dtype=np.int32)
    L = _lib.lib()
    L.init(f(np.float64), f(np.float64), f(np.float64), f(np.int32), f(np.int32), sl1, sl2,
           ct.c_int(ni), ct.c_int(nj), ct.c_int(nk), ct.c_int(nsub))
    assert _lib.last_error() == ""
    o1, o2 = Oracle().init(ni, nj, nk)
    assert np.array_equal(sl1, o1[:nsub]) and np.array_equal(sl2, o2[:nsub])


def test_init_rejects_odd_volume():
    L = _lib.lib()
    dims = (3, 3, 3)
    f = lambda dt: np.zeros(dims, dtype=dt, order="F")
    sl = np.full(13, -7, dtype=np.int32)
    L.init(f(np.float64), f(np.float64), f(np.float64), f(np.int32), f(np.int32), sl, sl.copy(),
           ct.c_int(3), ct.c_int(3), ct.c_int(3), ct.c_int(13))
    assert "N even" in _lib.last_error()
    assert np.all(sl == -7)  # nothing written


def _run_args(dims, nstart=1, nstop=4, nout=2, order="F", dtype=np.float64):
    nx = np.zeros(dims, dtype=dtype, order=order)
    s = np.zeros(dims, dtype=np.int32, order=order)
    rows = max((1 + nstop - nstart) // nout, 1) if nout > 0 else 1
    return [ct.c_int(nstart), ct.c_int(nstop), ct.c_int(nout), nx, nx.copy(order="A"), nx.copy(order="A"), s,
            s.copy(order="A"), np.zeros(max(np.prod(dims) // 2, 1), np.int32),
            np.zeros(max(np.prod(dims) // 2, 1), np.int32), ct.c_double(0.5), ct.c_double(1.0),
            ct.c_double(0.3), ct.c_double(0.7), ct.c_int(dims[0]), ct.c_int(dims[1]), ct.c_int(dims[2]),
            ct.c_int(0), np.zeros(rows, np.int32), np.zeros((rows, 5), np.float64, order="F")]


def test_run_rejects_bad_inputs_before_device_work():
    L = _lib.lib()
    # odd N: the reference's init overflows its lists (odd ni with N even is accepted)
    L.run(*_run_args((3, 3, 3)))
    assert "must be even" in _lib.last_error()
    # nout = 0 would divide by zero in the reference (mod(istep, 0))
    L.run(*_run_args((4, 4, 4), nout=0))
    assert "nout must be positive" in _lib.last_error()
    # periodic boundaries need even ni, nj, nk
    _lib.check(L.chll_configure(0, 0, 1, 0))
    try:
        L.run(*_run_args((4, 3, 4)))
        assert "periodic" in _lib.last_error()
        L.run(*_run_args((3, 4, 4)))
        assert "periodic" in _lib.last_error()
    finally:
        _lib.check(L.chll_configure(0, 0, 0, 1))


def test_argtypes_reject_wrong_layout_like_reference():
    L = _lib.lib()
    with pytest.raises(ct.ArgumentError):
        L.run(*_run_args((4, 4, 4), order="C"))
    with pytest.raises(ct.ArgumentError):
        L.run(*_run_args((4, 4, 4), dtype=np.float32))


def test_configure_validates():
    L = _lib.lib()
    assert L.chll_configure(7, 0, 0, 1) == -1 and "rng" in _lib.last_error()
    assert L.chll_configure(0, 9, 0, 1) == -1 and "arith" in _lib.last_error()
    assert L.chll_configure(0, 0, 5, 1) == -1 and "boundary" in _lib.last_error()
    assert L.chll_configure(0, 0, 0, 1) == 0


def test_bandwidth_probe_validates():
    """The streaming probe refuses bad patterns, sizes and repeat counts
    before touching the device."""
    from chllmodel_amd import _lib

    L = _lib.lib()
    assert L.chll_bandwidth_probe(4, 1 << 30, 3, None, None) == -1 and "pattern" in _lib.last_error()
    assert L.chll_bandwidth_probe(0, 255 << 20, 3, None, None) == -1 and "256 MiB" in _lib.last_error()
    assert L.chll_bandwidth_probe(0, 17 << 30, 3, None, None) == -1 and "16 GiB" in _lib.last_error()
    assert L.chll_bandwidth_probe(0, 1 << 30, 0, None, None) == -1 and "reps" in _lib.last_error()
