"""ctypes binding of libchll_mod.so (the HIP library built by
``chllmodel_amd.build``).

The drop-in symbols ``init`` and ``run`` get exactly the argtypes the
reference declares (chll/chll.py:9-36); the ``chll_*`` extensions are declared
from include/chll_mod.h.  Loading fails loudly when the library is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np
from numpy import ctypeslib

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CHLL_LIB", os.path.join(PKG, "libchll_mod.so"))

RNG = {"philox": 0, "stream": 1, "buffer": 2}
ARITH = {"compat": 0, "fp64": 1}
BC = {"open": 0, "periodic": 1}
ROW_COLS = 16
ROW_NAMES = ("istep", "E", "eta", "paccept", "faccept", "d", "E_true", "eta_exact",
             "Qxx", "Qyy", "Qzz", "Qxy", "Qxz", "Qyz", "naccept", "nflip")
TIMING_KINDS = ("director", "chirality", "measure", "finalize", "fill", "fused", "fused_next")

_f64p = ct.POINTER(ct.c_double)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)


class ChllParams(ct.Structure):
    """Mirror of ``chll_params`` (include/chll_mod.h)."""

    _fields_ = [
        ("g", ct.c_double), ("KK", ct.c_double), ("kbt", ct.c_double),
        ("rng", ct.c_int32), ("arith", ct.c_int32),
        ("seed", ct.c_uint64), ("stream_pos", ct.c_uint64),
        ("buffer", _f64p), ("buffer_len", ct.c_int64),
        ("run_index", ct.c_uint32), ("replica", ct.c_uint32),
        ("adapt", ct.c_int32), ("pad", ct.c_int32),
    ]


class ChllError(RuntimeError):
    pass


_lib = None


def _load_torch_first() -> None:
    """Import torch, when installed, before this library is loaded.  Both link
    the HIP runtime under one soname, so whichever loads first serves the whole
    process; torch's ROCm build finds no GPU ("No HIP GPUs are available") on
    a runtime of another ROCm release, while libchll_mod.so runs on torch's."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> ct.CDLL:
    """Load libchll_mod.so once and declare every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ChllError(f"{LIB_PATH} not built: run `python -m chllmodel_amd.build` (hipcc, gfx950)")
    _load_torch_first()
    L = ct.CDLL(LIB_PATH)
    # drop-in symbols, argtypes as chll/chll.py:9-36
    L.init.argtypes = [ctypeslib.ndpointer(dtype=np.float64, ndim=3, flags="F_CONTIGUOUS"),
                       ctypeslib.ndpointer(dtype=np.float64, ndim=3, flags="F_CONTIGUOUS"),
                       ctypeslib.ndpointer(dtype=np.float64, ndim=3, flags="F_CONTIGUOUS"),
                       ctypeslib.ndpointer(dtype=np.int32, ndim=3, flags="F_CONTIGUOUS"),
                       ctypeslib.ndpointer(dtype=np.int32, ndim=3, flags="F_CONTIGUOUS"),
                       ctypeslib.ndpointer(dtype=np.int32, ndim=1, flags="F_CONTIGUOUS"),
                       ctypeslib.ndpointer(dtype=np.int32, ndim=1, flags="F_CONTIGUOUS"),
                       ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
    L.init.restype = None
    L.run.argtypes = [ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                      ctypeslib.ndpointer(dtype=np.float64, ndim=3, flags="F_CONTIGUOUS"),
                      ctypeslib.ndpointer(dtype=np.float64, ndim=3, flags="F_CONTIGUOUS"),
                      ctypeslib.ndpointer(dtype=np.float64, ndim=3, flags="F_CONTIGUOUS"),
                      ctypeslib.ndpointer(dtype=np.int32, ndim=3, flags="F_CONTIGUOUS"),
                      ctypeslib.ndpointer(dtype=np.int32, ndim=3, flags="F_CONTIGUOUS"),
                      ctypeslib.ndpointer(dtype=np.int32, ndim=1, flags="F_CONTIGUOUS"),
                      ctypeslib.ndpointer(dtype=np.int32, ndim=1, flags="F_CONTIGUOUS"),
                      ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), ct.POINTER(ct.c_double),
                      ct.POINTER(ct.c_double),
                      ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                      ctypeslib.ndpointer(dtype=np.int32, ndim=1, flags="F_CONTIGUOUS"),
                      ctypeslib.ndpointer(dtype=np.float64, ndim=2, flags="F_CONTIGUOUS")]
    L.run.restype = None
    # extensions
    L.chll_last_error.argtypes = []
    L.chll_last_error.restype = ct.c_char_p
    L.chll_configure.argtypes = [ct.c_int32] * 4
    L.chll_set_seed.argtypes = [ct.c_uint64, ct.c_uint32]
    L.chll_set_stream.argtypes = [ct.c_uint64, ct.c_uint64]
    L.chll_get_stream_pos.argtypes = []
    L.chll_get_stream_pos.restype = ct.c_uint64
    L.chll_set_uniform_buffer.argtypes = [_f64p, ct.c_int64, ct.c_uint64]
    L.chll_last_rows.argtypes = [_f64p, ct.c_int32]
    L.chll_last_rows.restype = ct.c_int32
    L.chll_lattice_create.argtypes = [ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, ct.POINTER(ct.c_void_p)]
    L.chll_lattice_destroy.argtypes = [ct.c_void_p]
    L.chll_lattice_set_stream.argtypes = [ct.c_void_p, ct.c_void_p]
    L.chll_lattice_upload.argtypes = [ct.c_void_p, _f64p, _f64p, _f64p, _i32p]
    L.chll_lattice_download.argtypes = [ct.c_void_p, _f64p, _f64p, _f64p, _i32p]
    L.chll_lattice_set_step.argtypes = [ct.c_void_p, ct.c_double]
    L.chll_lattice_get_step.argtypes = [ct.c_void_p, _f64p]
    L.chll_lattice_sweeps.argtypes = [ct.c_void_p, ct.POINTER(ChllParams), ct.c_int32, ct.c_int32, ct.c_int32,
                                      _f64p, ct.c_int32, _i64p]
    L.chll_lattice_measure.argtypes = [ct.c_void_p, ct.c_double, ct.c_double, _f64p]
    L.chll_lattice_sync.argtypes = [ct.c_void_p]
    L.chll_lattice_set_timing.argtypes = [ct.c_void_p, ct.c_int32]
    L.chll_lattice_timing.argtypes = [ct.c_void_p, _f64p, _i64p, ct.c_int32]
    L.chll_lattice_set_kernel.argtypes = [ct.c_void_p] + [ct.c_int32] * 5
    L.chll_lattice_get_kernel.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, _i32p, _i32p, _i32p]
    L.chll_lattice_set_fusion.argtypes = [ct.c_void_p, ct.c_int32]
    L.chll_lattice_get_fusion.argtypes = [ct.c_void_p, _i32p]
    for name in ("chll_configure", "chll_set_seed", "chll_set_stream", "chll_set_uniform_buffer",
                 "chll_lattice_create", "chll_lattice_destroy", "chll_lattice_set_stream",
                 "chll_lattice_upload", "chll_lattice_download", "chll_lattice_set_step",
                 "chll_lattice_get_step", "chll_lattice_sweeps", "chll_lattice_measure", "chll_lattice_sync",
                 "chll_lattice_set_timing", "chll_lattice_timing", "chll_lattice_set_kernel",
                 "chll_lattice_get_kernel", "chll_lattice_set_fusion", "chll_lattice_get_fusion"):
        getattr(L, name).restype = ct.c_int
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        raise ChllError(lib().chll_last_error().decode())
    return rc


def last_error() -> str:
    return lib().chll_last_error().decode()
