#!/usr/bin/env python
"""Benchmark: chiral Lebwohl-Lasher checkerboard sweeps on one MI355X.

A "step" is one full sweep (director half-sweeps on both colours, then
chirality half-sweeps on both colours, then the on-device step adaptation and,
every nout sweeps, the observable reduction) of an L^3 FP64 lattice already
resident in HBM.  The metric is site-updates/s = L^3 * sweeps / s.

Default workload: L = 256 (16,777,216 sites), seeded isotropic directors and
+-1 spins, eps = 1, K = 1.0, g = 0.5, J = 0, kT = 0.7, d0 = 0.3, nout = 50, open
boundaries (the reference's), binary64 energies, site-keyed Philox random
numbers; the reference-compatible arithmetic (binary32-rounded site energies),
periodic boundaries and a run without measurement sweeps are timed beside it
under "variants".  A second section times the 64-replica
batched phase-diagram scan against the same replicas run one after another.

Prints ONE JSON line.  The roofline object is for the dominant kernel (the
one with the largest total time: the director half-sweep, or with FP64 the
fused colour-1 director + chirality launch), timed with HIP events on the
stream it runs on, priced at its algorithmic traffic per site (BYTES below).
The CPU baseline is
the C restatement of the reference's sweep in its own arithmetic and draw
schedule, timed on bounded samples of the same workload on one core and on
every host core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "site-updates/s at L=256 FP64 + 64-replica batched scan; % of HBM roofline"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per lattice site per launch of each sweep kernel:
# director half-sweep: read all directors 24 B and all spins 1 B, write the
# active half's directors 12 B; chirality half-sweep: read all directors 24 B
# and the active half's spins 0.5 B (FP64; the reference-compatible energies
# also read the other half's spins, 0.5 B more), write those spins 0.5 B (upper
# bound: every flip written); fused colour-1 director + chirality, and fused
# colour-0 chirality + next sweep's colour-0 director: the director traffic
# plus the active half's spin writes 0.5 B; in reference arithmetic the
# colour-1 launch also writes its flip bitmap, one bit per active site
# (1/16 B per site).
BYTES = {"director": 37.0, "chirality": 25.0, "fused": 37.5, "fused_next": 37.5}
BYTES_CHIRALITY_COMPAT = 25.5
BYTES_FUSED_COMPAT = 37.5 + 1.0 / 16.0
KERNEL_NAMES = {"director": "director half-sweep", "chirality": "chirality half-sweep",
                "fused": "k_pair<2> (colour-1 director + chirality, fused)",
                "fused_next": "k_pair<3> (colour-0 chirality + next sweep's colour-0 director, fused)"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> int:
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where the host sets it (the GPU host's CPU share)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(L, st0, args):
    """The C restatement of the reference's sweep (oracle/chll_oracle.c,
    oracle_run_mt), timed on bounded samples of the same L^3 workload in the
    reference's own arithmetic (binary32-rounded site energies) and draw
    schedule: each sweep first fills its 10 nsub host uniforms, as the
    reference's random_number calls do (chll/chll_mod.f90:66-80), then runs
    its four half-sweeps.  Once on one core, once on every host core this
    process may use (OpenMP over planes, the analogue of the reference's
    -acc=multicore build, chll/Makefile:8,10); the headline is the all-cores
    figure.  oracle/calibration.json (written by oracle/calibrate.py where
    the reference builds) gives reference time / restatement time on one
    core, measured on one host against the reference compiled from source."""
    from oracle.oracle import RNG_STREAM, Oracle, copy_state

    orc = Oracle()
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    cpu_model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    N = L ** 3

    def timed(nsweeps, threads):
        st = copy_state(st0)
        t0 = time.perf_counter()
        # nstart = 2: no header; nout past the range: no rows
        orc.run(st, 2, 1 + nsweeps, 10 ** 9, args.g, args.K, args.d0, args.kT, rng=RNG_STREAM, seed=1,
                threads=threads, hostfill=True)
        return time.perf_counter() - t0

    nth = host_threads()
    n1, nm = args.cpu_sweeps, args.cpu_sweeps_mt
    log(f"bench: CPU baseline: {n1} sweeps on 1 core, {nm} sweeps on {nth} threads, L={L}")
    t1 = timed(n1, 1)
    tm = timed(nm, nth)
    calib = None
    cf = os.path.join(ROOT, "oracle", "calibration.json")
    if os.path.exists(cf):
        with open(cf) as f:
            calib = json.load(f)
    one = {"value": N * n1 / t1, "unit": "site-updates/s", "cores": 1,
           "sample": f"{n1} sweeps of the L={L} workload, {t1:.2f} s wall"}
    out = {"value": N * nm / tm, "unit": "site-updates/s", "cores": nth, "kind": "port",
           "sample": f"{nm} sweeps of the L={L} workload (K={args.K} g={args.g} kT={args.kT} d0={args.d0}, open "
                     f"boundaries, reference arithmetic, host-filled uniforms), {tm:.2f} s wall on {nth} threads "
                     f"of {cpu_model}",
           "single_core": one, "cpu_model": cpu_model,
           "threads_scaling": (N * nm / tm) / one["value"]}
    if calib:
        r = calib["reference_over_restatement"]
        out["calibration"] = {"reference_over_restatement_1core": r, "measured_on": calib["cpu"],
                              "workload": calib["workload"],
                              "reference_estimate_1core_site_updates_per_s": one["value"] / r}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--L", type=int, default=256)
    ap.add_argument("--nout", type=int, default=50)
    ap.add_argument("--arith", default="fp64", choices=["compat", "fp64"])
    ap.add_argument("--bc", default="open", choices=["open", "periodic"])
    ap.add_argument("--rng", default="philox", choices=["philox", "stream"])
    ap.add_argument("--g", type=float, default=0.5)
    ap.add_argument("--K", type=float, default=1.0)
    ap.add_argument("--kT", type=float, default=0.7)
    ap.add_argument("--d0", type=float, default=0.3)
    ap.add_argument("--cpu-sweeps", type=int, default=2, help="CPU baseline sweeps on one core")
    ap.add_argument("--cpu-sweeps-mt", type=int, default=10, help="CPU baseline sweeps on all host cores")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events")
    ap.add_argument("--no-variants", action="store_true", help="skip the other arithmetic mode")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-replica scan")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in run() timing")
    ap.add_argument("--no-grad", action="store_true", help="skip the energy-gradient timing")
    ap.add_argument("--no-analysis", action="store_true", help="skip the analysis-kernel timing")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the other single-lattice workload configurations (L=32, L=128 anneal, L=480)")
    ap.add_argument("--analysis-only", action="store_true",
                    help="run only the analysis section (and its CPU figures unless --no-cpu)")
    ap.add_argument("--dropin-sweeps", type=int, default=20, help="sweeps per drop-in run() call")
    ap.add_argument("--spinup-s", type=float, default=0.3,
                    help="untimed sweeps for at least this long before the warm-up (GPU clocks leave idle)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default kernels instead of autotuning (clean profiles)")
    ap.add_argument("--batched-blocks", type=int, default=20, help="timed blocks of 10 sweeps")
    ap.add_argument("--batched-config", default=None, metavar="BLOCK,CHUNK_MIB,BRANCHES",
                    help="run the scan with this block size, chunk budget (MiB, 0 = one chunk) and number of graph "
                         "branches instead of autotuning (e.g. under a profiler, which changes the timings the "
                         "autotuner compares)")
    ap.add_argument("--batched-obs-blocks", type=int, default=1000,
                    help="blocks of 10 sweeps for the scan's per-slot observables (0: skip); 1000 = the scan "
                         "workload's 10k sweeps")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run one plain elementwise copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; runs on more than one GPU are out of scope "
              "(the lattice would shard into k-slabs with one halo plane per phase)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from chllmodel_amd.lattice import Lattice
    from chllmodel_amd.states import isotropic_state

    if args.analysis_only:
        torch.cuda.init()
        res = analysis_rates(args)
        if not args.no_cpu:
            analysis_cpu_baseline(res)
        print(json.dumps({"analysis": res}), flush=True)
        return

    L = args.L
    N = L ** 3
    log(f"bench: L={L} N={N} bc={args.bc} arith={args.arith} rng={args.rng} steps={args.steps} warmup={args.warmup}")
    st0 = isotropic_state(L, L, L, 20261203)
    torch.cuda.init()
    stream = torch.cuda.Stream()
    if args.copy_calib_mib > 0:
        n = args.copy_calib_mib * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)  # reads and writes copy_calib_mib MiB once
        torch.cuda.synchronize()
        del src, dst
        log(f"bench: calibration copy of {args.copy_calib_mib} MiB done")

    ceilings = hbm_ceilings()
    copy_rate = ceilings["copy"]["best_GBps"]
    mix_rate = ceilings["read2_write1"]["best_GBps"]
    tuned = {}
    held_rate = {}  # reference arithmetic: held chirality trials per sweep

    def time_single(arith, timing, bc=None, nout=None):
        """Autotune the half-sweep kernels for this shape (once per arithmetic;
        runs with another boundary condition or measurement interval reuse the
        choice), then untimed sweeps -- at least W, and at least
        args.spinup_s seconds of them, so that the GPU has left its idle
        clocks -- and K timed sweeps bracketed by synchronisation.  The timed
        window ends on a measurement sweep (its last sweep is a multiple of
        nout), so even a short window contains one."""
        from chllmodel_amd.autotune import apply, tune_lattice

        bc = args.bc if bc is None else bc
        nout = args.nout if nout is None else nout
        lat = Lattice(L, L, L, bc=bc)
        if arith in tuned:
            if not args.no_tune:
                apply(lat, tuned[arith], arith)
        elif args.no_tune:
            tuned[arith] = {k: list(lat.get_kernel(k, arith)) for k in ("director", "chirality")}
            tuned[arith]["fuse"] = lat.get_fusion()
            log(f"bench: {arith}: default kernels {tuned[arith]}")
        else:
            tuned[arith] = tune_lattice(L, L, L, bc=args.bc, arith=arith, use_cache=False)
            log(f"bench: {arith}: tuned director {tuned[arith]['director']} chirality {tuned[arith]['chirality']}")
            apply(lat, tuned[arith], arith)
        lat.set_stream(stream.cuda_stream)
        lat.upload_state(st0)
        lat.set_step(args.d0)
        p = lat.params(g=args.g, KK=args.K, kbt=args.kT, rng=args.rng, arith=arith, seed=20261203, adapt=True)
        W, K = args.warmup, args.steps
        # untimed sweeps 1 .. t0 - 1: chunks of 50 until spinup_s has passed, then W more
        done, t_spin = 0, time.perf_counter()
        while time.perf_counter() - t_spin < args.spinup_s:
            lat.sweeps(done + 1, done + 50, nout, params=p, want_rows=False)
            lat.sync()
            done += 50
        first = done + W + 1
        if nout > 0:
            first += (-(first + K - 1)) % nout  # last timed sweep is a multiple of nout
        lat.sweeps(done + 1, first - 1, nout, params=p, want_rows=False)
        lat.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lat.sweeps(first, first + K - 1, nout, params=p, want_rows=False)
        lat.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        log(f"bench: {arith} bc={bc} nout={nout}: {K} sweeps {first}..{first + K - 1} in {wall:.4f} s "
            f"after {first - 1} untimed")
        kernels = {}
        if timing:
            # per-kernel HIP events in a pass of their own, after the timed
            # region (events around every launch add gaps between kernels):
            # the same number of sweeps, the same alignment on nout
            KT = min(K, 100)
            nxt = first + K
            if nout > 0:
                nxt += (-(nxt + KT - 1)) % nout
                lat.sweeps(first + K, nxt - 1, nout, params=p, want_rows=False)
            lat.set_timing(True)
            lat.sweeps(nxt, nxt + KT - 1, nout, params=p, want_rows=False)
            kt = lat.timing()
            lat.set_timing(False)
            for name, (ms, cnt) in kt.items():
                if cnt:
                    kernels[name] = {"launches": cnt, "avg_us": 1e3 * ms / cnt, "total_ms": ms, "sweeps": KT}
        held = lat.held_trials()
        last_sweep = nxt + KT - 1 if timing else first + K - 1
        held_rate[arith] = held / last_sweep
        if held:
            log(f"bench: {arith}: {held} chirality trials held over {last_sweep} sweeps")
        row = lat.measure(args.g, args.K)
        lat.close()
        log(f"bench: {arith}: E_true={row['E_true']:.6f} eta={row['eta_exact']:.6f}")
        return N * K / wall, wall, kernels

    W, K = args.warmup, args.steps
    sups, wall, kernels = time_single(args.arith, not args.no_timing)
    roof = None
    if args.arith == "compat":
        BYTES["chirality"] = BYTES_CHIRALITY_COMPAT
        BYTES["fused"] = BYTES_FUSED_COMPAT
    sweep_bytes = None
    gap_us = None
    sweep_kinds = [k for k in BYTES if k in kernels]
    if sweep_kinds:
        for k in sweep_kinds:
            kernels[k]["GBps"] = BYTES[k] * N / (kernels[k]["avg_us"] * 1e-6) / 1e9
        nsw_t = kernels[sweep_kinds[0]]["sweeps"]
        # bytes per site-update from the launches actually made (timing pass)
        sweep_bytes = sum(BYTES[k] * kernels[k]["launches"] for k in sweep_kinds) / nsw_t
        # time per sweep outside kernels: the timed sweep less the sum of the
        # kernels a sweep runs (event-timed pass; the finalize kernel's events
        # overstate a 4-5 us kernel by ~2 us)
        ker_us = sum(v["total_ms"] for v in kernels.values()) * 1e3 / nsw_t
        gap_us = 1e6 * wall / K - ker_us
        dom = max(sweep_kinds, key=lambda k: kernels[k]["total_ms"])
        ach = kernels[dom]["GBps"]
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"{dom}_L{L}_{args.bc}_{args.arith}_traffic.json")
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        # sweep level: the reference algorithm's 125 B per site-update (four
        # dependent full-lattice half-sweeps: 2 x 37 B + 2 x 25.5 B) at the
        # measured site-update rate; 70 % of 8 TB/s is 44.8 G site-updates/s
        ref_bytes = 125.0
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": traffic, "kernel": KERNEL_NAMES[dom],
                "copy_GBps": copy_rate, "frac_of_copy": ach / copy_rate,
                "ceiling": {"pattern": "two arrays read, one rewritten in place (the sweep's mix), plain "
                                       "streaming kernel on 2 GiB arrays", "GBps": mix_rate,
                            "frac_of_peak": mix_rate / HBM_PEAK_GBS, "frac": ach / mix_rate},
                "streaming_ceilings_GBps": {k: round(v["best_GBps"], 1) for k, v in ceilings.items()},
                "algorithmic_bytes_per_launch": BYTES[dom] * N,
                "algorithmic_bytes_per_site": BYTES[dom],
                "kernel_avg_us": kernels[dom]["avg_us"],
                "sweep_level": {"bytes_per_site_update": ref_bytes,
                                "accounting": "reference algorithm: four full-lattice half-sweeps per sweep",
                                "achieved": sups * ref_bytes / 1e9, "unit": "GB/s",
                                "frac": sups * ref_bytes / (HBM_PEAK_GBS * 1e9),
                                "threshold_site_updates_per_s": 0.7 * HBM_PEAK_GBS * 1e9 / ref_bytes,
                                "meets_70pct": sups * ref_bytes >= 0.7 * HBM_PEAK_GBS * 1e9}}
    variants = {}
    if not args.no_variants:
        other = "compat" if args.arith == "fp64" else "fp64"
        v_sups, v_wall, v_k = time_single(other, not args.no_timing)
        variants[other] = {"site_updates_per_s": v_sups, "ms_per_step": 1e3 * v_wall / K,
                           "kernels": {k: round(v["avg_us"], 2) for k, v in v_k.items()}}
        if other == "compat":
            variants[other]["held_trials_per_sweep"] = round(held_rate.get(other, 0.0), 1)
        # the reference algorithm's accounting (125 B per site-update) against 8 TB/s
        variants[other]["sweep_level_frac"] = v_sups * 125.0 / (HBM_PEAK_GBS * 1e9)
        # the same workload with periodic boundaries, and without measurement sweeps
        other_bc = "periodic" if args.bc == "open" else "open"
        for name, kw in ((f"{args.arith}_{other_bc}", {"bc": other_bc}), (f"{args.arith}_no_measure", {"nout": 0})):
            v_sups, v_wall, _ = time_single(args.arith, False, **kw)
            variants[name] = {"site_updates_per_s": v_sups, "ms_per_step": 1e3 * v_wall / K,
                              "bc": kw.get("bc", args.bc), "nout": kw.get("nout", args.nout)}
    out = {
        "metric": METRIC,
        "value": sups,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ChLL Metropolis sweep, L={L} single lattice, {args.bc} boundaries, "
                               f"{args.arith} arithmetic, {args.rng} RNG, K={args.K} g={args.g} kT={args.kT} "
                               f"d0={args.d0} nout={args.nout}",
                   "L": L, "sites": N, "bc": args.bc, "arith": args.arith, "rng": args.rng, "nout": args.nout},
        "sweeps_per_s": K / wall,
        "sweep_bytes_per_site_update": sweep_bytes,
        "sweep_hbm_frac": sups * sweep_bytes / (HBM_PEAK_GBS * 1e9) if sweep_bytes else None,
        "gap_us_per_sweep": gap_us,
        "roofline": roof,
        "kernels": kernels,
        "variants": variants,
        "kernel_choice": {a: {k: t.get(k) for k in ("director", "chirality", "fuse")} for a, t in tuned.items()},
        "cpu_baseline": None,
    }
    if not args.no_dropin:
        out["dropin_run"] = dropin_run(L, st0, args)
    if not args.no_grad:
        out["gradient"] = gradient_rate(args)
        # the gradient reads one set of fields and writes another: a copy's mix
        out["gradient"]["frac_of_copy"] = out["gradient"]["GBps"] / copy_rate
    if not args.no_workloads and args.arith == "fp64":
        out["workloads"] = workload_rates(stream)
    if not args.no_analysis:
        out["analysis"] = analysis_rates(args)
    if not args.no_batched:
        out["batched"] = batched_scan(args, stream)
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        out["cpu_baseline"] = cpu_baseline(L, st0, args)
        out["speedup_vs_cpu"] = sups / out["cpu_baseline"]["value"]
        out["speedup_vs_cpu_1core"] = sups / out["cpu_baseline"]["single_core"]["value"]
        if "analysis" in out:
            analysis_cpu_baseline(out["analysis"])
    print(json.dumps(out), flush=True)


def hbm_ceilings(mib=2048, reps=10):
    """Rates plain streaming kernels reach on this GPU for the load/store
    mixes the sweep moves (chll_bandwidth_probe, two arrays of `mib` MiB, far
    past the memory-side cache; best of `reps` passes, HIP events): a copy,
    the sweep's own mix (two arrays read, one rewritten in place), reads only
    and writes only.  These are the practical ceilings beside the 8 TB/s
    peak."""
    from chllmodel_amd._lib import BW_PATTERNS, bandwidth_probe

    out = {}
    for name in BW_PATTERNS:
        best, med = bandwidth_probe(name, mib, reps)
        out[name] = {"best_GBps": best, "median_GBps": med}
    log("bench: streaming ceilings (GB/s, best of %d on %d MiB arrays): %s" % (
        reps, mib, ", ".join(f"{k} {v['best_GBps']:.0f}" for k, v in out.items())))
    return out


def dropin_run(L, st0, args):
    """The reference's own entry point run() (chll/chll_mod.f90:40) on host
    arrays: each call uploads the Fortran-ordered fields over PCIe, sweeps on
    the device and downloads them again, as chll/chll.py:160-173 drives it
    (once per simulation).  Calls of n and 10 n sweeps give the fixed cost of
    a call (the two transfers and the layout changes) and the cost per sweep.
    Reported beside value, never as value."""
    import ctypes as ct

    import numpy as np

    from chllmodel_amd import _lib

    lib = _lib.lib()
    st = {k: np.asfortranarray(v.copy()) for k, v in st0.items()}
    n1 = args.dropin_sweeps
    n2 = 10 * n1
    nsub = L ** 3 // 2
    msteps = np.zeros(1, dtype=np.int32)
    m = np.zeros((1, 5), dtype=np.float64, order="F")
    sl1 = np.zeros(nsub, np.int32)
    sl2 = np.zeros(nsub, np.int32)
    arith = {"compat": 0, "fp64": 1}[args.arith]
    _lib.check(lib.chll_configure(0, arith, {"open": 0, "periodic": 1}[args.bc], 0))
    _lib.check(lib.chll_set_seed(20261203, 0))

    def call(n):
        lib.run(ct.c_int(1), ct.c_int(n), ct.c_int(n), st["nx"], st["ny"], st["nz"], st["s"], st["dope"], sl1, sl2,
                ct.c_double(args.g), ct.c_double(args.K), ct.c_double(args.d0), ct.c_double(args.kT), ct.c_int(L),
                ct.c_int(L), ct.c_int(L), ct.c_int(0), msteps, m)
        if _lib.last_error():
            raise RuntimeError(_lib.last_error())

    def best(n, calls=3):
        times = []
        for _ in range(calls):
            t0 = time.perf_counter()
            call(n)
            times.append(time.perf_counter() - t0)
        return min(times)

    try:
        call(n1)  # first call allocates and builds the lists
        t1, t2 = best(n1), best(n2)
    finally:
        _lib.check(lib.chll_configure(0, 0, 0, 1))
    per_sweep = (t2 - t1) / (n2 - n1)
    fixed = t1 - n1 * per_sweep
    log(f"bench: drop-in run(): {n1} sweeps per call in {t1:.4f} s, {n2} in {t2:.4f} s")
    return {"entry": "run (chll/chll_mod.f90:40), host arrays in and out per call",
            "sweeps_per_call": n1, "ms_per_call": 1e3 * t1, "site_updates_per_s": L ** 3 * n1 / t1,
            f"site_updates_per_s_{n2}_sweeps": L ** 3 * n2 / t2,
            "fixed_ms_per_call": 1e3 * fixed, "ms_per_sweep_in_call": 1e3 * per_sweep,
            "arith": args.arith, "statistic": "best of 3 calls at each length"}


def gradient_rate(args):
    """Energy and tangent director gradient (chll_energy_grad) of a batch of
    8 L=128 lattices (16.8 M sites, as L=256), timed with HIP events on
    torch's current stream.  Algorithmic traffic 49 B per site: read the
    directors 24 B and the spin 1 B, write the gradient 24 B."""
    import torch

    from chllmodel_amd.autograd import energy_and_grad

    B, Lg = 8, 128
    gen = torch.Generator(device="cuda").manual_seed(7)
    n = torch.randn(B, 3, Lg, Lg, Lg, device="cuda", dtype=torch.float64, generator=gen)
    n = n / n.norm(dim=1, keepdim=True)
    s = (torch.randint(0, 2, (B, Lg, Lg, Lg), device="cuda", generator=gen) * 2 - 1).to(torch.int8)
    K = torch.full((B,), args.K, device="cuda", dtype=torch.float64)
    g = torch.full((B,), args.g, device="cuda", dtype=torch.float64)
    eps = torch.ones(B, device="cuda", dtype=torch.float64)
    for _ in range(3):
        energy_and_grad(n, s, K, g, eps)
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        energy_and_grad(n, s, K, g, eps)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    sites = B * Lg ** 3
    gbps = 49.0 * sites / sec / 1e9
    log(f"bench: energy + gradient of {B} x {Lg}^3: {1e6 * sec:.1f} us per call")
    return {"workload": f"{B} lattices of {Lg}^3, open boundaries, energy + tangent gradient + parameter sums",
            "us_per_call": 1e6 * sec, "sites_per_s": sites / sec, "algorithmic_bytes_per_site": 49.0,
            "GBps": gbps, "hbm_frac": gbps / HBM_PEAK_GBS}


def analysis_rates(args):
    """Analysis kernels on device-resident inputs, timed per call with HIP
    events on torch's current stream (each call also copies its small host
    table to the device and waits for it, so host latency enters; the figure
    is the best of three groups of calls):
    * chirality structure factor of a 480^3 spin field on 7 planes (faces and
      body diagonals) x 500 kv in [0, 0.5], the stripe analysis of the L=480
      runs; the reference's form evaluates planes x kv x N phases;
    * nematic + chiral shell correlations of a 128^3 state, 8 shells; the
      reference's form visits every ordered pair (r, r + d) once.
    The CPU figures come from analysis_cpu_baseline (the CPU-baseline leg)."""
    import numpy as np
    import torch

    from chllmodel_amd.analysis import corr_counts, shell_correlations_device, structure_factors_device

    planes = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1]]

    def best_per_call(call, reps, groups=3):
        """Seconds per call: HIP events around `reps` calls, best of `groups`."""
        best = float("inf")
        for _ in range(groups):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                call()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3 / reps)
        return best

    gen = torch.Generator(device="cuda").manual_seed(11)
    res = {}

    Ls, bins = 480, 500
    N = Ls ** 3
    ds = torch.randint(0, 2, (N,), device="cuda", generator=gen, dtype=torch.int8) * 2 - 1
    kvs = np.linspace(0.0, 0.5, bins)
    for _ in range(2):
        structure_factors_device(ds, (Ls, Ls, Ls), kvs, planes)
    sec = best_per_call(lambda: structure_factors_device(ds, (Ls, Ls, Ls), kvs, planes), 5)
    terms = len(planes) * bins * N
    log(f"bench: structure factor {Ls}^3 x {len(planes)} planes x {bins} kv: {1e3 * sec:.2f} ms per call")
    res["structure_factor"] = {
        "workload": f"{Ls}^3 spins, {len(planes)} planes x {bins} kv in [0, 0.5]",
        "ms_per_call": 1e3 * sec, "reference_phase_terms_per_s": terms / sec,
        "phase_terms": terms, "histogram_pass_GBps": N / sec / 1e9}
    del ds

    Lr, ns = 128, 8
    N = Lr ** 3
    n = torch.randn(3, N, device="cuda", dtype=torch.float64, generator=gen)
    n /= n.norm(dim=0, keepdim=True)
    dn = [n[0].contiguous(), n[1].contiguous(), n[2].contiguous()]
    dsp = torch.randint(0, 2, (N,), device="cuda", generator=gen, dtype=torch.int8) * 2 - 1
    shell_correlations_device((Lr, Lr, Lr), ns, dn, dsp)
    sec = best_per_call(lambda: shell_correlations_device((Lr, Lr, Lr), ns, dn, dsp), 3)
    pairs = int(corr_counts((Lr, Lr, Lr), ns).sum())
    log(f"bench: shell correlations {Lr}^3, {ns} shells: {1e3 * sec:.2f} ms per call")
    res["shell_correlations"] = {
        "workload": f"{Lr}^3 directors + spins, {ns} shells, nematic and chiral",
        "nshells": ns, "ms_per_call": 1e3 * sec, "ordered_pairs": pairs, "pairs_per_s": pairs / sec}
    return res


def analysis_cpu_baseline(res):
    """CPU figures for the analysis section: the oracle's vectorised numpy sums
    (one core) on bounded samples, per unit of the same work (phase terms of
    the structure factor, ordered pairs of the correlations)."""
    import numpy as np

    from oracle import analysis as ora

    sf = res["structure_factor"]
    Lc = 96
    sc = np.where(np.random.default_rng(1).random((Lc, Lc, Lc)) < 0.5, -1, 1).astype(np.int32)
    t0 = time.perf_counter()
    for kv in (0.1, 0.2, 0.3):
        ora.get_sc(sc, kv * np.array([1.0, 1.0, 1.0]))
    rate = 3 * Lc ** 3 / (time.perf_counter() - t0)
    sf["cpu_oracle"] = {"phase_terms_per_s": rate, "cores": 1, "kind": "port",
                        "sample": f"numpy direct sum, {Lc}^3 spins, 3 kv, plane (1,1,1)",
                        "est_seconds_for_workload": sf["phase_terms"] / rate}
    sf["speedup_vs_cpu"] = sf["reference_phase_terms_per_s"] / rate
    cr = res["shell_correlations"]
    ns, Lc = cr["nshells"], 24
    rng = np.random.default_rng(2)
    v = rng.normal(size=(3, Lc, Lc, Lc))
    v /= np.linalg.norm(v, axis=0)
    scc = np.where(rng.random((Lc, Lc, Lc)) < 0.5, -1, 1).astype(np.int32)
    t0 = time.perf_counter()
    ora.calc_n_corr(v[0], v[1], v[2], ns)
    ora.calc_s_corr(scc, ns)
    rate = int(ora.corr_counts((Lc, Lc, Lc), ns).sum()) / (time.perf_counter() - t0)
    cr["cpu_oracle"] = {"pairs_per_s": rate, "cores": 1, "kind": "port",
                        "sample": f"numpy per-offset sums, {Lc}^3, {ns} shells, nematic and chiral"}
    cr["speedup_vs_cpu"] = cr["pairs_per_s"] / rate


def workload_rates(stream):
    """The other single-lattice workload configurations at their full sizes
    and parameters, FP64 sweeps with Philox draws and device-generated
    isotropic initial states, each timed on a bounded window after 0.3 s of
    untimed sweeps (clock spin-up):
    - L=32 periodic, K=0.5, g=0, kT=1, d0=0.3, nout=10: the whole 1000-sweep run;
    - L=128, K=1, g=0.5, the linear anneal kT 1.5 -> 0.3 over 20,000 sweeps
      (one call with chll_params.kbt_end, every sweep at its own kT, a row
      every 100 sweeps): its first 2000 sweeps;
    - L=480, K=1, g=0.25, kT=0.5, nout=50 (coarsening): 100 sweeps."""
    import torch

    from chllmodel_amd.lattice import Lattice, anneal_kt

    cases = [
        ("L=32 periodic K=0.5 g=0 kT=1 nout=10, 1000 sweeps", 32, "periodic",
         dict(g=0.0, KK=0.5, kbt=1.0), 10, 1000),
        ("L=128 anneal kT 1.5->0.3 over 20000 sweeps (K=1 g=0.5, nout=100): sweeps 1-2000", 128, "open",
         dict(g=0.5, KK=1.0, kbt=1.5, kbt_end=0.3), 100, 2000),
        ("L=480 coarsening K=1 g=0.25 kT=0.5 nout=50: 100 sweeps", 480, "open",
         dict(g=0.25, KK=1.0, kbt=0.5), 50, 100),
    ]
    out = []
    for name, L, bc, par, nout, nsw in cases:
        with Lattice(L, L, L, bc=bc) as lat:
            lat.set_stream(stream.cuda_stream)
            lat.init_state("isotropic", seed=20261203)
            lat.set_step(0.3)
            spin = lat.params(**{**par, "kbt_end": 0.0}, rng="philox", arith="fp64", seed=7, run_index=1)
            t0 = time.perf_counter()
            done = 0
            while time.perf_counter() - t0 < 0.3:
                lat.sweeps(done + 1, done + 50, 0, params=spin, want_rows=False)
                lat.sync()
                done += 50
            lat.init_state("isotropic", seed=20261203)
            lat.set_step(0.3)
            p = lat.params(**par, rng="philox", arith="fp64", seed=20261203)
            if "kbt_end" in par:
                # the anneal's own kT schedule over 20,000 sweeps, of which this
                # window runs the first nsw: sweep nsw at its scheduled kT
                p.kbt_end = anneal_kt(1.5, 0.3, nsw, 1, 20000)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            got = lat.sweeps(1, nsw, nout, params=p)["rows"]
            lat.sync()
            wall = time.perf_counter() - t0
        out.append({"workload": name, "sites": L ** 3, "sweeps": nsw, "seconds": wall,
                    "site_updates_per_s": L ** 3 * nsw / wall, "ms_per_sweep": 1e3 * wall / nsw,
                    "rows": int(got.shape[0]), "E_true_last": float(got[-1, 6])})
        log(f"bench: workload {name}: {1e3 * wall / nsw:.4f} ms per sweep, {L ** 3 * nsw / wall / 1e9:.2f} G")
    return out


def batched_scan(args, stream):
    """64 replicas in 5 parallel-tempering ladders (g = 0, 0.5, 1, 2, 3 with
    L = 32, 48, 64, 96, 128; kT geometric in [0.2, 2]; K = 1): one hipGraph
    replay per block of 10 sweeps, against the same replicas swept one after
    another with the single-lattice path (measurement every 10 sweeps)."""
    import torch

    from chllmodel_amd.ensemble import Ensemble, autotune, phase_scan_replicas
    from chllmodel_amd.lattice import Lattice

    reps = phase_scan_replicas()
    sites = sum(r["ni"] * r["nj"] * r["nk"] for r in reps)
    Kb, nb = 10, args.batched_blocks
    if args.batched_config:
        bs, mib, br = (int(v) for v in args.batched_config.split(","))
        tune = {"best": bs, "chunk": mib << 20, "branches": br, "ms_per_sweep": None, "ms_per_sweep_chunk": None}
        log(f"bench: batched scan, {len(reps)} replicas, {sites} sites, fixed configuration {tune}")
    else:
        log(f"bench: batched scan, {len(reps)} replicas, {sites} sites, autotuning")
        tune = autotune(reps, bc="open", arith=args.arith, K=Kb, use_cache=False)
    ens = Ensemble(reps, bc="open", arith=args.arith)
    ens.set_stream(stream.cuda_stream)
    ens.set_block(tune["best"])
    ens.set_chunk(tune["chunk"])
    ens.set_branches(tune["branches"])
    ens.init_state("isotropic", seed=20261203)
    ens.run(2, Kb, seed=20261203, want_rows=False, use_graph=True)
    ens.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(nb, Kb, seed=20261203, want_rows=False, use_graph=True)
    ens.sync()
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    stt = ens.state()
    nchunks = ens.chunks()
    ens.close()
    nsw = nb * Kb
    log(f"bench: batched {nsw} sweeps in {t_b:.4f} s")
    # the same replicas one after another, single-lattice path
    t_seq = 0.0
    for i, r in enumerate(reps):
        lat = Lattice(r["ni"], r["nj"], r["nk"])
        lat.set_stream(stream.cuda_stream)
        lat.init_state("isotropic", seed=20261203, replica=i)
        lat.set_step(r["d0"])
        p = lat.params(g=r["g"], KK=r["KK"], kbt=r["kbt"], arith=args.arith, seed=20261203, replica=i)
        lat.sweeps(1, 2, Kb, params=p, want_rows=False)
        lat.sync()
        t0 = time.perf_counter()
        lat.sweeps(3, 2 + nsw, Kb, params=p, want_rows=False)
        lat.sync()
        t_seq += time.perf_counter() - t0
        lat.close()
    log(f"bench: sequential {nsw} sweeps x {len(reps)} replicas in {t_seq:.4f} s")
    att, acc = stt["swap_att"], stt["swap_acc"]
    obs = batched_observables(args, stream, reps, Kb, tune)
    return {"workload": "64 replicas, 5 PT ladders g={0,0.5,1,2,3} x L={32,48,64,96,128}, kT in [0.2,2] "
                        "geometric, K=1, swaps every 10 sweeps, one hipGraph replay per 10 sweeps",
            "replicas": len(reps), "sites": sites, "sweeps": nsw, "arith": args.arith,
            "site_updates_per_s": sites * nsw / t_b, "ms_per_sweep": 1e3 * t_b / nsw,
            "sequential_site_updates_per_s": sites * nsw / t_seq, "sequential_ms_per_sweep": 1e3 * t_seq / nsw,
            "speedup_vs_sequential": t_seq / t_b, "block_size": tune["best"],
            "chunk_bytes": tune["chunk"], "chunks": nchunks, "branches": tune["branches"],
            "autotune_ms_per_sweep": tune["ms_per_sweep"], "autotune_ms_per_sweep_chunk": tune["ms_per_sweep_chunk"],
            "swap_attempts": int(att.sum()), "swap_acceptance": float(acc.sum() / max(att.sum(), 1)),
            "swap_note": "rungs a factor 1.21 apart in kT on lattices of 32^3-128^3 sites: beta-gaps times "
                         "energy gaps are far above 1, so swaps are attempted (and timed) but almost never "
                         "accepted; acceptance on small lattices is covered by the ensemble tests",
            "observables": obs}


def batched_observables(args, stream, reps, Kb, tune):
    """Per-temperature-slot observables of the scan after the timed part: a
    fresh ensemble of the same replicas, args.batched_obs_blocks blocks of Kb
    sweeps, one measurement row per replica and block; the first quarter is
    dropped as equilibration.  Canonical C_v = N var(e)/kT^2 and
    chi_c = N var(eta)/kT with jackknife errors, means with blocking errors
    (observables.summarize), and the notebooks' own forms: C_v =
    var(E[-100:])/kT of the reported energy per site
    (notebooks/paper_plots2.ipynb, calc_specific_heat) and chi_c =
    N var(eta[-500:])/kT (notebooks/.ipynb_checkpoints/
    paper_plots-checkpoint.ipynb:440-452)."""
    import numpy as np

    from chllmodel_amd.ensemble import Ensemble, series_by_temperature
    from chllmodel_amd.observables import chi_notebook, cv_notebook, summarize

    nb = args.batched_obs_blocks
    if nb <= 0:
        return None
    ens = Ensemble(reps, bc="open", arith=args.arith)
    ens.set_stream(stream.cuda_stream)
    ens.set_block(tune["best"])
    ens.set_chunk(tune["chunk"])
    ens.set_branches(tune["branches"])
    ens.init_state("isotropic", seed=20261203)
    t0 = time.perf_counter()
    rows = ens.run(nb, Kb, seed=20261203, run_index=1, want_rows=True, use_graph=True)
    dt = time.perf_counter() - t0
    ens.close()
    ser = series_by_temperature(rows, reps)
    slots = []
    for sl in range(len(reps)):
        r0 = reps[sl]  # slot sl holds temperature reps[sl]['kbt'] of its ladder
        N = r0["ni"] * r0["nj"] * r0["nk"]
        sm = summarize(ser[sl], r0["kbt"], N)
        eref = np.array([rows[t, int(ser[sl]["replica"][t]), 14] for t in range(rows.shape[0])])
        ent = {"slot": sl, "ladder": r0["ladder"], "L": r0["ni"], "g": r0["g"], "kT": r0["kbt"]}
        ent.update({k: float(f"{v:.6g}") for k, v in sm.items()})
        ent["Cv_notebook"] = float(f"{cv_notebook(eref, r0['kbt']):.6g}")
        ent["chi_notebook"] = float(f"{chi_notebook(ser[sl]['eta'], r0['kbt'], N):.6g}")
        slots.append(ent)
    log(f"bench: batched observables, {nb} blocks of {Kb} sweeps in {dt:.2f} s")
    return {"blocks": nb, "sweeps_per_block": Kb, "samples_per_slot": int(rows.shape[0]),
            "discarded_fraction": 0.25, "seconds": dt, "slots": slots,
            "definitions": "E, eta, P2: means with blocking errors; Cv = N var(E/N)/kT^2, chi = N var(eta)/kT with "
                           "jackknife errors; Cv_notebook = var(E_ref[-100:])/kT; chi_notebook = N var(eta[-500:])/kT"}


if __name__ == "__main__":
    main()
