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
TIMING_KINDS = ("director", "chirality", "measure", "finalize", "fill", "fused", "fused_next", "held")
INIT_KINDS = {"isotropic": 0, "aligned_x": 1, "aligned_xz": 2}

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
        ("kbt_end", ct.c_double),
    ]


class ChllError(RuntimeError):
    pass


_lib = None


def _mapped(fragment: str) -> list[str]:
    """Paths of the shared objects mapped into this process whose path
    contains `fragment` (Linux /proc/self/maps)."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({ln.split()[-1] for ln in f if fragment in ln and "/" in ln})
    except OSError:
        return []


def _torch_hip_runtime() -> str | None:
    """Path of the HIP runtime bundled with an installed torch, found without
    importing torch."""
    import importlib.util

    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    return path if os.path.exists(path) else None


def _one_hip_runtime() -> None:
    """Make the process hold one HIP runtime whatever is imported first.

    libchll_mod.so needs libamdhip64.so.7.  An installed torch bundles its own
    HIP runtime under that soname; when the system's ROCm runtime is loaded
    first, torch imported afterwards finds no GPU ("No HIP GPUs are
    available").  So when torch is installed its runtime is loaded here, by
    path and globally, before libchll_mod.so, without importing torch: the
    library's dependency then resolves to it, and a later `import torch` maps
    the same file again (same inode), so both share one runtime.  When another
    HIP runtime is already mapped and torch's is not, this raises with the
    reason instead of leaving a process whose torch would see no GPU."""
    path = _torch_hip_runtime()
    if path is None:
        return  # no torch: the system ROCm runtime serves the library alone
    held = _mapped("libamdhip64")
    real = os.path.realpath(path)
    if held and all(os.path.realpath(h) != real for h in held):
        raise ChllError(
            f"this process already holds the HIP runtime {held[0]}, not torch's {path}: torch imported "
            "afterwards would find no GPU.  Load chllmodel_amd (or import torch) before any other HIP library.")
    if not held:
        ct.CDLL(path, mode=ct.RTLD_GLOBAL)


def lib() -> ct.CDLL:
    """Load libchll_mod.so once and declare every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ChllError(f"{LIB_PATH} not built: run `python -m chllmodel_amd.build` (hipcc, gfx950)")
    _one_hip_runtime()
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
    L.chll_lattice_init_state.argtypes = [ct.c_void_p, ct.c_int32, ct.c_uint64, ct.c_uint32, ct.c_double]
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
    L.chll_lattice_set_screen.argtypes = [ct.c_void_p, ct.c_double]
    L.chll_lattice_held_trials.argtypes = [ct.c_void_p, _i64p]
    L.chll_bandwidth_probe.argtypes = [ct.c_int32, ct.c_int64, ct.c_int32, _f64p, _f64p]
    for name in ("chll_configure", "chll_set_seed", "chll_set_stream", "chll_set_uniform_buffer",
                 "chll_lattice_create", "chll_lattice_destroy", "chll_lattice_set_stream",
                 "chll_lattice_upload", "chll_lattice_download", "chll_lattice_init_state", "chll_lattice_set_step",
                 "chll_lattice_get_step", "chll_lattice_sweeps", "chll_lattice_measure", "chll_lattice_sync",
                 "chll_lattice_set_timing", "chll_lattice_timing", "chll_lattice_set_kernel",
                 "chll_lattice_get_kernel", "chll_lattice_set_fusion", "chll_lattice_get_fusion", "chll_lattice_set_screen",
                 "chll_lattice_held_trials", "chll_bandwidth_probe"):
        getattr(L, name).restype = ct.c_int
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        raise ChllError(lib().chll_last_error().decode())
    return rc


def last_error() -> str:
    return lib().chll_last_error().decode()


BW_PATTERNS = {"copy": 0, "read2_write1": 1, "read": 2, "write": 3}


def bandwidth_probe(pattern: str, mib: int = 2048, reps: int = 10) -> tuple[float, float]:
    """Best and median GB/s of a plain streaming kernel with one mix of loads
    and stores (chll_bandwidth_probe) on two arrays of `mib` MiB each."""
    best, med = ct.c_double(), ct.c_double()
    check(lib().chll_bandwidth_probe(BW_PATTERNS[pattern], int(mib) << 20, int(reps), ct.byref(best), ct.byref(med)))
    return best.value, med.value
