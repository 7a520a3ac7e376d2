"""Device-resident lattice: the state stays in HBM across sweep calls.

This is the Python face of the ``chll_lattice_*`` extension of
libchll_mod.so.  The reference keeps no device state between ``run`` calls
(chll/chll_mod.f90:61, 111); ``Lattice`` is what a caller uses to sweep, measure
and time without a host round trip per call.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import ARITH, BC, INIT_KINDS, RNG, ROW_COLS, ROW_NAMES, TIMING_KINDS, ChllParams, check, lib

_f64p = ct.POINTER(ct.c_double)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class Lattice:
    """An (ni, nj, nk) chiral Lebwohl-Lasher lattice in GPU memory."""

    def __init__(self, ni: int, nj: int, nk: int, bc: str = "open"):
        self.shape = (int(ni), int(nj), int(nk))
        self.bc = bc
        self.N = ni * nj * nk
        h = ct.c_void_p()
        check(lib().chll_lattice_create(ni, nj, nk, BC[bc], ct.byref(h)))
        self._h = h
        self._buf_keep = None

    # -- state ---------------------------------------------------------------
    def upload(self, nx, ny, nz, s) -> None:
        nx, ny, nz = (_fortran(a, np.float64) for a in (nx, ny, nz))
        s = _fortran(s, np.int32)
        for a in (nx, ny, nz, s):
            if a.shape != self.shape:
                raise ValueError(f"field shape {a.shape} != lattice {self.shape}")
        check(lib().chll_lattice_upload(self._h, nx.ctypes.data_as(_f64p), ny.ctypes.data_as(_f64p),
                                        nz.ctypes.data_as(_f64p), s.ctypes.data_as(_i32p)))

    def upload_state(self, st: dict) -> None:
        self.upload(st["nx"], st["ny"], st["nz"], st["s"])

    def download(self) -> dict:
        out = {k: np.zeros(self.shape, dtype=np.float64, order="F") for k in ("nx", "ny", "nz")}
        out["s"] = np.zeros(self.shape, dtype=np.int32, order="F")
        check(lib().chll_lattice_download(self._h, out["nx"].ctypes.data_as(_f64p), out["ny"].ctypes.data_as(_f64p),
                                          out["nz"].ctypes.data_as(_f64p), out["s"].ctypes.data_as(_i32p)))
        return out

    def init_state(self, kind: str = "isotropic", seed: int = 0, replica: int = 0, p_flip: float = 0.5) -> None:
        """Generate the initial state on the device, keyed by (seed, replica,
        site): 'isotropic' directors (the scripts' init_random formula),
        'aligned_x' n = (1,0,0) or 'aligned_xz' n = (1,0,1)/sqrt(2); spins -1
        with probability p_flip (init_rand_s).  No host arrays involved."""
        check(lib().chll_lattice_init_state(self._h, INIT_KINDS[kind], int(seed), int(replica), float(p_flip)))

    def set_step(self, d: float) -> None:
        check(lib().chll_lattice_set_step(self._h, float(d)))

    def get_step(self) -> float:
        v = ct.c_double()
        check(lib().chll_lattice_get_step(self._h, ct.byref(v)))
        return v.value

    def set_stream(self, hip_stream_ptr: int | None) -> None:
        """Launch on a caller-owned HIP stream (e.g. torch.cuda.Stream().cuda_stream)."""
        check(lib().chll_lattice_set_stream(self._h, ct.c_void_p(hip_stream_ptr or 0)))

    KINDS = ("director", "chirality")

    def set_kernel(self, variant: str = "march", tile: int = 0, nslab: int = 1, *, kinds=KINDS,
                   ariths=("compat", "fp64")) -> None:
        """Half-sweep kernel for the given trial kinds and arithmetic modes.
        'site': one site per thread, tile bit 0 = one band of rows per XCD
        (else blocks in lattice order), bit 1 = two sites per thread with pair
        loads, bit 2 = the fused colour-0 FP64 launch walks its blocks in
        reverse.  'march': tile 0/1/2 = 64x4/32x8/16x16 packed
        sites x rows marching along k in nslab slabs, 3/4/5 = the same tiles one
        plane per block."""
        for kind in ([kinds] if isinstance(kinds, str) else kinds):
            for ar in ([ariths] if isinstance(ariths, str) else ariths):
                check(lib().chll_lattice_set_kernel(self._h, self.KINDS.index(kind), ARITH[ar],
                                                    {"site": 0, "march": 1}[variant], int(tile), int(nslab)))

    def get_kernel(self, kind: str = "director", arith: str = "fp64") -> tuple:
        v, t, n = ct.c_int32(), ct.c_int32(), ct.c_int32()
        check(lib().chll_lattice_get_kernel(self._h, self.KINDS.index(kind), ARITH[arith], ct.byref(v), ct.byref(t),
                                            ct.byref(n)))
        return ("site", "march")[v.value], t.value, n.value

    def set_fusion(self, on: bool) -> None:
        """FP64 sweeps: run the colour-1 chirality trials inside the colour-1
        director launch (same results, one lattice pass fewer)."""
        check(lib().chll_lattice_set_fusion(self._h, 1 if on else 0))

    def get_fusion(self) -> bool:
        v = ct.c_int32()
        check(lib().chll_lattice_get_fusion(self._h, ct.byref(v)))
        return bool(v.value)

    def set_screen(self, scale: float) -> None:
        """Reference arithmetic: decide each trial from its binary64 energy
        difference when the binary32 roundings of the two site energies
        (at most scale x their magnitude, 2^-21 by default) cannot change the
        outcome; 0 rounds both energies of every trial.  Results are the same
        at every scale."""
        check(lib().chll_lattice_set_screen(self._h, float(scale)))

    def held_trials(self) -> int:
        """Reference-arithmetic chirality trials of colour 1 held since the
        lattice was created: those whose outcome the screen left open in a
        fused launch, decided once colour 0's chirality outcomes were known."""
        n = ct.c_int64(0)
        check(lib().chll_lattice_held_trials(self._h, ct.byref(n)))
        return int(n.value)

    def sync(self) -> None:
        check(lib().chll_lattice_sync(self._h))

    # -- sweeps --------------------------------------------------------------
    def params(self, *, g, KK, kbt, rng="philox", arith="compat", seed=0, stream_pos=0, buffer=None,
               run_index=0, replica=0, adapt=True, kbt_end=0.0) -> ChllParams:
        """Sweep parameters; kbt_end > 0 anneals kT linearly from kbt at the
        call's first sweep to kbt_end at its last (anneal_kt gives each
        sweep's value)."""
        p = ChllParams()
        p.g, p.KK, p.kbt, p.kbt_end = float(g), float(KK), float(kbt), float(kbt_end)
        p.rng, p.arith = RNG[rng], ARITH[arith]
        p.seed, p.stream_pos = int(seed), int(stream_pos)
        if buffer is not None:
            self._buf_keep = np.ascontiguousarray(buffer, dtype=np.float64)
            p.buffer = self._buf_keep.ctypes.data_as(_f64p)
            p.buffer_len = self._buf_keep.size
        p.run_index, p.replica = int(run_index), int(replica)
        p.adapt = 1 if adapt else 0
        return p

    def sweeps(self, nstart: int, nstop: int, nout: int = 0, *, params: ChllParams | None = None,
               want_rows: bool = True, want_counts: bool = False, **kw) -> dict:
        """Run sweeps istep = nstart..nstop.  Returns {'rows': (R,16) array,
        'counts': (nsweeps,2) array or None, 'stream_pos': int}."""
        p = params if params is not None else self.params(**kw)
        nsw = max(nstop - nstart + 1, 0)
        max_rows = 0
        if nout > 0 and nsw > 0:
            max_rows = nstop // nout - (nstart - 1) // nout
        rows = np.zeros((max(max_rows, 1), ROW_COLS), dtype=np.float64) if want_rows else None
        counts = np.zeros((max(nsw, 1), 2), dtype=np.int64) if want_counts else None
        got = check(lib().chll_lattice_sweeps(
            self._h, ct.byref(p), int(nstart), int(nstop), int(nout),
            rows.ctypes.data_as(_f64p) if rows is not None else None, max_rows,
            counts.ctypes.data_as(_i64p) if counts is not None else None))
        return {"rows": rows[:got] if rows is not None else None,
                "counts": counts[:nsw] if counts is not None else None,
                "stream_pos": p.stream_pos}

    def measure(self, g: float, KK: float) -> dict:
        row = np.zeros(ROW_COLS, dtype=np.float64)
        check(lib().chll_lattice_measure(self._h, float(g), float(KK), row.ctypes.data_as(_f64p)))
        return dict(zip(ROW_NAMES, row.tolist()))

    # -- timing --------------------------------------------------------------
    def set_timing(self, on: bool) -> None:
        check(lib().chll_lattice_set_timing(self._h, 1 if on else 0))

    def timing(self) -> dict:
        ms = np.zeros(len(TIMING_KINDS), dtype=np.float64)
        n = np.zeros(len(TIMING_KINDS), dtype=np.int64)
        check(lib().chll_lattice_timing(self._h, ms.ctypes.data_as(_f64p), n.ctypes.data_as(_i64p), len(TIMING_KINDS)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(TIMING_KINDS)}

    # -- lifetime ------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().chll_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def anneal_kt(kbt: float, kbt_end: float, istep: int, nstart: int, nstop: int) -> float:
    """kT of sweep istep in a call with a ramp (chll_params.kbt_end > 0): the
    library's binary64 expression, operation for operation."""
    if not (kbt_end > 0.0 and nstop > nstart):
        return float(kbt)
    if istep == nstop:
        return float(kbt_end)
    return float(kbt) + (float(kbt_end) - float(kbt)) * float(istep - nstart) / float(nstop - nstart)


def rows_to_dict(rows: np.ndarray) -> dict:
    return {name: rows[:, i].copy() for i, name in enumerate(ROW_NAMES)}
