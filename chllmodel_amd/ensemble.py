"""Batched replicas with on-device parallel tempering (the ``chll_ensemble_*``
extension of libchll_mod.so).

The reference maps phase diagrams by running one ChLLSim per (K, kT) or
(g, kT) point, one after another (scripts/phase_diagram_large1.py:30-37,
scripts/gfield0.0.py:10-19).  An ``Ensemble`` holds every replica of such a
scan in HBM and sweeps all of them with one launch per half-sweep; every K
sweeps it measures each replica and attempts parallel-tempering swaps between
neighbouring temperatures of each ladder.  One block of K sweeps is captured in
a hipGraph and replayed.
"""
from __future__ import annotations

import ctypes as ct
import json
import os
import time

import numpy as np

from ._lib import ARITH, BC, INIT_KINDS, lib
from .states import isotropic_state

ENS_COLS = 16
ENS_NAMES = ("istep", "kT", "E_true", "eta", "Qxx", "Qyy", "Qzz", "Qxy", "Qxz", "Qyz",
             "paccept", "faccept", "d", "slot", "E_ref", "reserved")

_f64p = ct.POINTER(ct.c_double)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)


class ReplicaDesc(ct.Structure):
    _fields_ = [("ni", ct.c_int32), ("nj", ct.c_int32), ("nk", ct.c_int32), ("ladder", ct.c_int32),
                ("g", ct.c_double), ("KK", ct.c_double), ("kbt", ct.c_double), ("d0", ct.c_double)]


def _declare(L):
    if getattr(L, "_ens_declared", False):
        return L
    L.chll_ensemble_last_error.restype = ct.c_char_p
    L.chll_ensemble_create.argtypes = [ct.POINTER(ReplicaDesc), ct.c_int32, ct.c_int32, ct.c_int32,
                                       ct.POINTER(ct.c_void_p)]
    L.chll_ensemble_destroy.argtypes = [ct.c_void_p]
    L.chll_ensemble_set_stream.argtypes = [ct.c_void_p, ct.c_void_p]
    L.chll_ensemble_set_block.argtypes = [ct.c_void_p, ct.c_int32]
    L.chll_ensemble_set_fusion.argtypes = [ct.c_void_p, ct.c_int32]
    L.chll_ensemble_set_chunk.argtypes = [ct.c_void_p, ct.c_int64]
    L.chll_ensemble_chunks.argtypes = [ct.c_void_p]
    L.chll_ensemble_chunks.restype = ct.c_int32
    L.chll_ensemble_set_branches.argtypes = [ct.c_void_p, ct.c_int32]
    L.chll_ensemble_blocks.argtypes = [ct.c_void_p]
    L.chll_ensemble_upload.argtypes = [ct.c_void_p, ct.c_int32, _f64p, _f64p, _f64p, _i32p]
    L.chll_ensemble_download.argtypes = [ct.c_void_p, ct.c_int32, _f64p, _f64p, _f64p, _i32p]
    L.chll_ensemble_run.argtypes = [ct.c_void_p, ct.c_uint64, ct.c_uint32, ct.c_int32, ct.c_int32, ct.c_int32,
                                    ct.c_int32, _f64p, ct.c_int32]
    L.chll_ensemble_sync.argtypes = [ct.c_void_p]
    L.chll_ensemble_init_state.argtypes = [ct.c_void_p, ct.c_int32, ct.c_uint64, ct.c_double]
    L.chll_ensemble_state.argtypes = [ct.c_void_p, _i32p, _f64p, _f64p, _i64p, _i64p, ct.POINTER(ct.c_uint32)]
    for n in ("chll_ensemble_create", "chll_ensemble_destroy", "chll_ensemble_set_stream", "chll_ensemble_set_block",
              "chll_ensemble_set_fusion", "chll_ensemble_set_chunk", "chll_ensemble_set_branches",
              "chll_ensemble_blocks", "chll_ensemble_upload", "chll_ensemble_download", "chll_ensemble_run",
              "chll_ensemble_sync", "chll_ensemble_state", "chll_ensemble_init_state"):
        getattr(L, n).restype = ct.c_int
    L._ens_declared = True
    return L


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().chll_ensemble_last_error().decode())
    return rc


class Ensemble:
    """Replicas ``[{'ni','nj','nk','ladder','g','KK','kbt','d0'}, ...]``:
    replicas of one ladder share lattice, g and K and are listed contiguously in
    ascending kbt; each replica starts at its own temperature slot."""

    def __init__(self, replicas, bc: str = "open", arith: str = "compat"):
        L = _declare(lib())
        self.replicas = [dict(r) for r in replicas]
        arr = (ReplicaDesc * len(replicas))()
        for a, r in zip(arr, replicas):
            a.ni, a.nj, a.nk, a.ladder = int(r["ni"]), int(r["nj"]), int(r["nk"]), int(r["ladder"])
            a.g, a.KK, a.kbt, a.d0 = float(r["g"]), float(r["KK"]), float(r["kbt"]), float(r.get("d0", 0.3))
        h = ct.c_void_p()
        _check(L.chll_ensemble_create(arr, len(replicas), BC[bc], ARITH[arith], ct.byref(h)))
        self._h = h
        self.bc, self.arith = bc, arith
        self.nrep = len(replicas)
        self.block_size = 256

    @property
    def sites(self):
        return [r["ni"] * r["nj"] * r["nk"] for r in self.replicas]

    def upload(self, r, st):
        if not 0 <= r < self.nrep:
            raise IndexError(f"replica {r} outside 0..{self.nrep - 1}")
        rr = self.replicas[r]
        shape = (rr["ni"], rr["nj"], rr["nk"])
        a = [np.asfortranarray(st[k], dtype=np.float64) for k in ("nx", "ny", "nz")]
        s = np.asfortranarray(st["s"], dtype=np.int32)
        for x in (*a, s):
            if x.shape != shape:  # the library reads ni nj nk elements of each
                raise ValueError(f"field shape {x.shape} != replica {r}'s lattice {shape}")
        _check(lib().chll_ensemble_upload(self._h, r, *[x.ctypes.data_as(_f64p) for x in a], s.ctypes.data_as(_i32p)))

    def download(self, r) -> dict:
        if not 0 <= r < self.nrep:
            raise IndexError(f"replica {r} outside 0..{self.nrep - 1}")
        rr = self.replicas[r]
        shape = (rr["ni"], rr["nj"], rr["nk"])
        out = {k: np.zeros(shape, dtype=np.float64, order="F") for k in ("nx", "ny", "nz")}
        out["s"] = np.zeros(shape, dtype=np.int32, order="F")
        _check(lib().chll_ensemble_download(self._h, r, *[out[k].ctypes.data_as(_f64p) for k in ("nx", "ny", "nz")],
                                            out["s"].ctypes.data_as(_i32p)))
        return out

    def init_isotropic(self, seed: int):
        for r, rr in enumerate(self.replicas):
            self.upload(r, isotropic_state(rr["ni"], rr["nj"], rr["nk"], seed + r))

    def init_state(self, kind: str = "isotropic", seed: int = 0, p_flip: float = 0.5):
        """Every replica's initial state generated on the device
        (Lattice.init_state with replica index r)."""
        _check(lib().chll_ensemble_init_state(self._h, INIT_KINDS[kind], int(seed), float(p_flip)))

    def set_block(self, bs: int):
        _check(lib().chll_ensemble_set_block(self._h, int(bs)))
        self.block_size = int(bs)

    def set_fusion(self, on: bool, pairs: bool = True):
        """Fused director + chirality launches within each block of sweeps
        (same results, fewer passes over the replicas); with `pairs`, two
        sites per thread in those launches when every replica's ni is a
        multiple of 4.  Reference arithmetic fuses only with `pairs`, where
        every replica allows them, through held chirality trials."""
        _check(lib().chll_ensemble_set_fusion(self._h, (1 if pairs else 2) if on else 0))

    def set_chunk(self, nbytes: int):
        """Run each block of sweeps chunk by chunk, at most `nbytes` of lattice
        state per chunk (0: all replicas at once); results are unchanged."""
        _check(lib().chll_ensemble_set_chunk(self._h, int(nbytes)))

    def chunks(self) -> int:
        return int(lib().chll_ensemble_chunks(self._h))

    def set_branches(self, n: int):
        """Run the chunks on 1 to 3 concurrent streams (graph branches);
        results are unchanged."""
        _check(lib().chll_ensemble_set_branches(self._h, int(n)))

    def set_stream(self, ptr):
        _check(lib().chll_ensemble_set_stream(self._h, ct.c_void_p(ptr or 0)))

    def run(self, nblocks: int, K: int = 10, *, seed: int = 0, run_index: int = 0, adapt: bool = True,
            use_graph: bool = True, want_rows: bool = True, max_rows: int | None = None):
        """nblocks x K sweeps; returns rows (nblocks, nrep, 16) or None."""
        if max_rows is None:
            max_rows = nblocks if want_rows else 0
        rows = np.zeros((max(max_rows, 1), self.nrep, ENS_COLS), dtype=np.float64) if want_rows else None
        _check(lib().chll_ensemble_run(self._h, int(seed), int(run_index), int(nblocks), int(K), 1 if adapt else 0,
                                       1 if use_graph else 0, rows.ctypes.data_as(_f64p) if rows is not None else None,
                                       int(max_rows)))
        return rows[: min(nblocks, max_rows)] if rows is not None else None

    def sync(self):
        _check(lib().chll_ensemble_sync(self._h))

    def state(self) -> dict:
        n = self.nrep
        rep_slot = np.zeros(n, np.int32)
        kbt = np.zeros(n)
        d = np.zeros(n)
        att = np.zeros(n, np.int64)
        acc = np.zeros(n, np.int64)
        ist = ct.c_uint32()
        _check(lib().chll_ensemble_state(self._h, rep_slot.ctypes.data_as(_i32p), kbt.ctypes.data_as(_f64p),
                                         d.ctypes.data_as(_f64p), att.ctypes.data_as(_i64p), acc.ctypes.data_as(_i64p),
                                         ct.byref(ist)))
        return {"rep_slot": rep_slot, "slot_kbt": kbt, "slot_d": d, "swap_att": att, "swap_acc": acc,
                "istep": ist.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().chll_ensemble_destroy(self._h)
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


def phase_scan_replicas(K=1.0, d0=0.3, kT_min=0.2, kT_max=2.0,
                        ladders=((0.0, 32, 13), (0.5, 48, 13), (1.0, 64, 13), (2.0, 96, 13), (3.0, 128, 12))):
    """The batched phase-diagram scan: one parallel-tempering ladder per field g
    with its own lattice size L, kT geometric in [kT_min, kT_max]
    (default: 64 replicas, g in {0, 0.5, 1, 2, 3}, L in {32, 48, 64, 96, 128})."""
    reps = []
    for lad, (g, L, n) in enumerate(ladders):
        for kT in np.geomspace(kT_min, kT_max, n):
            reps.append({"ni": L, "nj": L, "nk": L, "ladder": lad, "g": g, "KK": K, "kbt": float(kT), "d0": d0})
    return reps


def series_by_temperature(rows: np.ndarray, replicas) -> dict:
    """Regroup ensemble rows (T, nrep, 16) into per-temperature-slot time
    series: slot -> {'kT', 'E', 'eta', 'Q', 'replica'}."""
    T, n, _ = rows.shape
    out = {}
    for s in range(n):
        out[s] = {"E": np.empty(T), "eta": np.empty(T), "Q": np.empty((T, 6)), "replica": np.empty(T, np.int32)}
    for t in range(T):
        for r in range(n):
            s = int(rows[t, r, 13])
            o = out[s]
            o["E"][t] = rows[t, r, 2]
            o["eta"][t] = rows[t, r, 3]
            o["Q"][t] = rows[t, r, 4:10]
            o["replica"][t] = r
            o["kT"] = rows[t, r, 1]
    return out


# ----------------------------------------------------------------------------
# Autotuner: the threads-per-block of the batched kernels.

CACHE = os.environ.get("CHLL_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "chll_autotune.json"))


def _signature(replicas, bc, arith, K):
    sizes = sorted((r["ni"], r["nj"], r["nk"]) for r in replicas)
    return json.dumps({"sizes": sizes, "bc": bc, "arith": arith, "K": K}, sort_keys=True)


def autotune(replicas, *, bc="open", arith="compat", K=10, blocks=3, candidates=(64, 128, 256, 512),
             chunks=(0, 96 << 20, 128 << 20, 160 << 20, 192 << 20, 256 << 20, 320 << 20), seed=12345,
             use_cache=True, verbose=False) -> dict:
    """Time each candidate block size, then each chunk budget (bytes of state
    per chunk, 0 = all replicas at once) with the best block size, each on one
    stream and, where it makes enough chunks, on two and three (set_branches), on a
    scratch ensemble of the same shape (so the caller's ensemble is not
    advanced).  Returns {'best': bs, 'chunk': bytes, 'branches': 1 to 3,
    'ms_per_sweep': {...}, 'ms_per_sweep_chunk': {'<bytes>' or
    '<bytes>x<branches>': ms}}; cached per shape in CACHE.  Apply with set_block(res['best']),
    set_chunk(res['chunk']), set_branches(res['branches'])."""
    key = _signature(replicas, bc, arith, K)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if key in cache and "branches" in cache[key]:
            return cache[key]
    times, ctimes = {}, {}
    with Ensemble(replicas, bc=bc, arith=arith) as ens:
        ens.init_state("isotropic", seed=seed)

        def time_blocks():
            ens.run(1, K, seed=seed, want_rows=False, use_graph=True)  # capture + warm up
            ens.sync()
            t0 = time.perf_counter()
            ens.run(blocks, K, seed=seed, want_rows=False, use_graph=True)
            ens.sync()
            return 1e3 * (time.perf_counter() - t0) / (blocks * K)

        for bs in candidates:
            ens.set_block(bs)
            times[bs] = time_blocks()
            if verbose:
                print(f"autotune bs={bs}: {times[bs]:.3f} ms/sweep", flush=True)
        best = min(times, key=times.get)
        ens.set_block(best)
        for cb in chunks:
            ens.set_chunk(cb)
            for nb in (1, 2, 3):
                if nb > ens.chunks():
                    continue
                ens.set_branches(nb)
                ctimes[(cb, nb)] = time_blocks()
                if verbose:
                    print(f"autotune bs={best} chunk={cb >> 20} MiB ({ens.chunks()} chunks, {nb} branches): "
                          f"{ctimes[(cb, nb)]:.3f} ms/sweep", flush=True)
            ens.set_branches(1)
    chunk, branches = min(ctimes, key=ctimes.get)
    res = {"best": int(best), "chunk": int(chunk), "branches": int(branches),
           "ms_per_sweep": {str(k): v for k, v in times.items()},
           "ms_per_sweep_chunk": {(str(c) if b == 1 else f"{c}x{b}"): v for (c, b), v in ctimes.items()}}
    if use_cache:
        cache[key] = res
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            with open(CACHE, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return res
