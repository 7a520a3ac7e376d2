"""Project autotuner for the single-lattice half-sweep kernels.

For one lattice shape, boundary condition and arithmetic it times every
candidate kernel (one or two sites per thread, in lattice order or in XCD row
bands, or the k-march with each tile shape and a few slab counts) for the director and
the chirality half-sweeps separately, with HIP events on the launch stream, and
applies the fastest.  It then times whole sweeps of those kernels against
sweeps with the fused colour-1 director + chirality launch (in reference
arithmetic with held chirality trials, which need site pairs for both kinds)
and keeps the faster.  Sweeping with
any candidate produces the same lattice (the trial arithmetic and random
counters do not depend on the kernel shape), so tuning on a scratch lattice of
the same shape is safe.  Results are cached per (shape, bc, arith) in the same
JSON file as the ensemble tuner (``CHLL_AUTOTUNE_CACHE``).
"""
from __future__ import annotations

import json
import os
import time

from .ensemble import CACHE
from .lattice import Lattice
from .states import isotropic_state


def candidates(ni, nj, nk):
    if ni % 2:
        return [("site", 0, 1)]  # odd ni: the one-site kernel in lattice order only
    out = [("site", 0, 1), ("site", 1, 1)]
    if ni % 4 == 0:
        out += [("site", 2, 1), ("site", 3, 1)]  # tile | 4 is tried on whole fused sweeps
    half = ni // 2
    for tile, (th, tj) in enumerate(((64, 4), (32, 8), (16, 16))):
        if th > 2 * half and tile < 2:
            continue
        tiles = -(-half // th) * -(-nj // tj)
        for target in (512, 1024, 2048, 4096):
            ns = max(1, min(nk, -(-target // tiles)))
            if -(-nk // ns) < 4:
                continue
            c = ("march", tile, ns)
            if c not in out:
                out.append(c)
        if th <= 2 * half or tile == 2:
            out.append(("march", tile + 3, 1))
    return out


def _timed_sweeps(lat, p, sweeps, spin_s=0.04, passes=2):
    """Per-kind kernel times {kind: (ms, launches)} of the fastest of `passes`
    passes of `sweeps` sweeps, after untimed sweeps for at least spin_s
    seconds (a lattice upload leaves the GPU idle long enough for its clocks
    to drop; a pass timed right after it runs up to ~10 % slow)."""
    nxt, t0 = 1, time.perf_counter()
    while True:
        lat.sweeps(nxt, nxt + 9, 0, params=p, want_rows=False)
        lat.sync()
        nxt += 10
        if time.perf_counter() - t0 >= spin_s or nxt > 2000:
            break
    best = None
    for _ in range(passes):
        lat.set_timing(True)
        lat.sweeps(nxt, nxt + sweeps - 1, 0, params=p, want_rows=False)
        t = lat.timing()
        lat.set_timing(False)
        nxt += sweeps
        if best is None or sum(ms for ms, _ in t.values()) < sum(ms for ms, _ in best.values()):
            best = t
    return best


def tune_lattice(ni, nj, nk, *, bc="open", arith="fp64", sweeps=20, seed=7, use_cache=True, verbose=False):
    """Returns {'director': (variant, tile, nslab), 'chirality': ..., 'fuse':
    bool (fused colour-1 director + chirality launch), 'us': {...}}."""
    key = json.dumps({"lattice": [ni, nj, nk], "bc": bc, "arith": arith}, sort_keys=True)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]
    res = {"us": {}}
    st0 = isotropic_state(ni, nj, nk, seed)
    with Lattice(ni, nj, nk, bc=bc) as lat:
        # each half-sweep kernel on its own first: no fused launches
        lat.set_fusion(False)
        best = {"director": (None, 1e30), "chirality": (None, 1e30)}
        best_site = (None, 1e30)
        best_pair = {"director": (None, 1e30), "chirality": (None, 1e30)}
        for cand in candidates(ni, nj, nk):
            # every candidate sweeps the same initial state: the cost of a sweep
            # depends on how many trials are accepted
            lat.upload_state(st0)
            lat.set_kernel(*cand, ariths=arith)
            lat.set_step(0.3)
            p = lat.params(g=0.5, KK=1.0, kbt=0.7, arith=arith, seed=seed)
            t = _timed_sweeps(lat, p, sweeps)
            for kind in ("director", "chirality"):
                ms, n = t[kind]
                us = 1e3 * ms / max(n, 1)
                res["us"][f"{kind} {cand[0]} tile={cand[1]} nslab={cand[2]}"] = round(us, 2)
                if us < best[kind][1]:
                    best[kind] = (cand, us)
                if kind == "director" and cand[0] == "site" and us < best_site[1]:
                    best_site = (cand, us)
                if cand[0] == "site" and cand[1] & 2 and us < best_pair[kind][1]:
                    best_pair[kind] = (cand, us)
                if verbose:
                    print(f"autotune {ni}x{nj}x{nk} {arith} {kind} {cand}: {us:.1f} us", flush=True)
        res["director"] = list(best["director"][0])
        res["chirality"] = list(best["chirality"][0])
        res["fuse"] = False
        # whole sweeps: the fastest kernels in four half-sweeps against the
        # fused launches, which need the one-site director kernel (FP64) or
        # site pairs for both kinds (reference arithmetic); with site pairs
        # also the form whose colour-0 launch walks its blocks in reverse (it
        # starts on the planes the colour-1 launch finished with, which the
        # memory-side cache still holds)
        trials = [(False, best["director"][0], best["chirality"][0])]
        if arith == "fp64":
            fused = (best_site[0], best["chirality"][0])
        elif best_pair["director"][0] is not None and best_pair["chirality"][0] is not None:
            fused = (best_pair["director"][0], best_pair["chirality"][0])
        else:
            fused = None
        if fused is not None:
            dc, cc = fused
            trials.append((True, dc, cc))
            if dc[1] & 2:
                trials.append((True, (dc[0], dc[1] | 4, dc[2]), cc))
        sweep_us = {}
        best_fused = (None, None, 1e30)
        for fuse, dcand, ccand in trials:
            lat.upload_state(st0)
            lat.set_kernel(*dcand, kinds="director", ariths=arith)
            lat.set_kernel(*ccand, kinds="chirality", ariths=arith)
            lat.set_fusion(fuse)
            lat.set_step(0.3)
            p = lat.params(g=0.5, KK=1.0, kbt=0.7, arith=arith, seed=seed)
            t = _timed_sweeps(lat, p, sweeps)
            us = 1e3 * sum(ms for ms, _ in t.values()) / sweeps
            res["us"][f"sweep fuse={fuse} director tile={dcand[1]} chirality tile={ccand[1]}"] = round(us, 2)
            if fuse and us < best_fused[2]:
                best_fused = (dcand, ccand, us)
            elif not fuse:
                sweep_us[False] = us
        if best_fused[2] < sweep_us[False]:
            res["fuse"] = True
            res["director"] = list(best_fused[0])
            res["chirality"] = list(best_fused[1])
    if use_cache:
        cache[key] = res
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            with open(CACHE, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return res


def apply(lat: Lattice, tuned: dict, arith: str) -> None:
    for kind in ("director", "chirality"):
        v, t, n = tuned[kind]
        lat.set_kernel(v, t, n, kinds=kind, ariths=arith)
    lat.set_fusion(bool(tuned.get("fuse", True)))
