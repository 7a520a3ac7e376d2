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
the reference's own run() compiled from source (serial), timed on a bounded
sample of the same workload.
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
# plus the active half's spin writes 0.5 B.
BYTES = {"director": 37.0, "chirality": 25.0, "fused": 37.5, "fused_next": 37.5}
BYTES_CHIRALITY_COMPAT = 25.5
KERNEL_NAMES = {"director": "k_dir (director half-sweep)", "chirality": "chirality half-sweep",
                "fused": "k_dir_spin (colour-1 director + chirality, fused)",
                "fused_next": "k_spin_dir (colour-0 chirality + next sweep's colour-0 director, fused)"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(L, st0, args):
    """Time the reference's run() (oracle/_ref, built from the reference
    source) on a bounded sample: args.cpu_sweeps sweeps of the same L^3
    workload, serial, host random_number included.  Falls back to the C
    oracle when the reference build is absent."""
    import numpy as np

    from oracle.oracle import REF_SO, RNG_PHILOX, Oracle, Reference, copy_state

    st = copy_state(st0)
    n = args.cpu_sweeps
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    cpu_model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    if os.path.exists(REF_SO):
        ref = Reference(stream=False)
        t0 = time.perf_counter()
        # nstart = 2 skips the header line; nout beyond the range prints no rows
        ref.run(st, 2, 1 + n, 10 ** 9, args.g, args.K, args.d0, args.kT)
        dt = time.perf_counter() - t0
        kind, what = "reference", "reference chll_mod.f90 run() built from source with flang -O2 (serial)"
    else:
        t0 = time.perf_counter()
        Oracle().run(st, 2, 1 + n, 10 ** 9, args.g, args.K, args.d0, args.kT, rng=RNG_PHILOX, seed=1)
        dt = time.perf_counter() - t0
        kind, what = "port", "C restatement of the reference (oracle), serial"
    N = L ** 3
    return {"value": N * n / dt, "unit": "site-updates/s", "cores": 1, "kind": kind,
            "sample": f"{n} sweeps of the L={L} workload ({what}); {dt:.2f} s wall on {cpu_model}"}


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
    ap.add_argument("--cpu-sweeps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events")
    ap.add_argument("--no-variants", action="store_true", help="skip the other arithmetic mode")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-replica scan")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in run() timing")
    ap.add_argument("--no-grad", action="store_true", help="skip the energy-gradient timing")
    ap.add_argument("--no-analysis", action="store_true", help="skip the analysis-kernel timing")
    ap.add_argument("--analysis-only", action="store_true",
                    help="run only the analysis section (and its CPU figures unless --no-cpu)")
    ap.add_argument("--dropin-sweeps", type=int, default=20, help="sweeps per drop-in run() call")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default kernels instead of autotuning (clean profiles)")
    ap.add_argument("--batched-blocks", type=int, default=20, help="timed blocks of 10 sweeps")
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

    tuned = {}

    def time_single(arith, timing, bc=None, nout=None):
        """Autotune the half-sweep kernels for this shape (once per arithmetic;
        runs with another boundary condition or measurement interval reuse the
        choice), then W warm-up sweeps and K timed sweeps bracketed by
        synchronisation."""
        from chllmodel_amd.autotune import apply, tune_lattice

        bc = args.bc if bc is None else bc
        nout = args.nout if nout is None else nout
        lat = Lattice(L, L, L, bc=bc)
        if arith in tuned:
            if not args.no_tune:
                apply(lat, tuned[arith], arith)
        elif args.no_tune:
            tuned[arith] = {k: list(lat.get_kernel(k, arith)) for k in ("director", "chirality")}
            tuned[arith]["fuse"] = lat.get_fusion() if arith == "fp64" else False
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
        lat.sweeps(1, W, nout, params=p, want_rows=False)
        lat.sync()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(stream)
        lat.sweeps(W + 1, W + K, nout, params=p, want_rows=False)
        ev1.record(stream)
        lat.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        log(f"bench: {arith} bc={bc} nout={nout}: {K} sweeps in {wall:.4f} s (events {ev0.elapsed_time(ev1):.3f} ms)")
        kernels = {}
        if timing:
            # per-kernel HIP events in a pass of their own, after the timed
            # region: the events around every launch add gaps between kernels
            KT = min(K, 100)
            lat.set_timing(True)
            lat.sweeps(W + K + 1, W + K + KT, nout, params=p, want_rows=False)
            kt = lat.timing()
            lat.set_timing(False)
            for name, (ms, cnt) in kt.items():
                if cnt:
                    kernels[name] = {"launches": cnt, "avg_us": 1e3 * ms / cnt, "total_ms": ms, "sweeps": KT}
        row = lat.measure(args.g, args.K)
        lat.close()
        log(f"bench: {arith}: E_true={row['E_true']:.6f} eta={row['eta_exact']:.6f}")
        return N * K / wall, wall, kernels

    W, K = args.warmup, args.steps
    sups, wall, kernels = time_single(args.arith, not args.no_timing)
    roof = None
    if args.arith == "compat":
        BYTES["chirality"] = BYTES_CHIRALITY_COMPAT
    sweep_bytes = None
    sweep_kinds = [k for k in BYTES if k in kernels]
    if sweep_kinds:
        for k in sweep_kinds:
            kernels[k]["GBps"] = BYTES[k] * N / (kernels[k]["avg_us"] * 1e-6) / 1e9
        # bytes per site-update from the launches actually made (timing pass)
        sweep_bytes = sum(BYTES[k] * kernels[k]["launches"] for k in sweep_kinds) / kernels[sweep_kinds[0]]["sweeps"]
        dom = max(sweep_kinds, key=lambda k: kernels[k]["total_ms"])
        ach = kernels[dom]["GBps"]
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"{dom}_L{L}_{args.bc}_{args.arith}_traffic.json")
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": traffic, "kernel": KERNEL_NAMES[dom],
                "algorithmic_bytes_per_launch": BYTES[dom] * N}
    variants = {}
    if not args.no_variants:
        other = "compat" if args.arith == "fp64" else "fp64"
        v_sups, v_wall, v_k = time_single(other, not args.no_timing)
        variants[other] = {"site_updates_per_s": v_sups, "ms_per_step": 1e3 * v_wall / K,
                           "kernels": {k: round(v["avg_us"], 2) for k, v in v_k.items()}}
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
    if not args.no_analysis:
        out["analysis"] = analysis_rates(args)
    if not args.no_batched:
        out["batched"] = batched_scan(args, stream)
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        log(f"bench: CPU baseline, {args.cpu_sweeps} sweeps at L={L}")
        out["cpu_baseline"] = cpu_baseline(L, st0, args)
        out["speedup_vs_cpu"] = sups / out["cpu_baseline"]["value"]
        if "analysis" in out:
            analysis_cpu_baseline(out["analysis"])
    print(json.dumps(out), flush=True)


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


def batched_scan(args, stream):
    """64 replicas in 5 parallel-tempering ladders (g = 0, 0.5, 1, 2, 3 with
    L = 32, 48, 64, 96, 128; kT geometric in [0.2, 2]; K = 1): one hipGraph
    replay per block of 10 sweeps, against the same replicas swept one after
    another with the single-lattice path (measurement every 10 sweeps)."""
    import torch

    from chllmodel_amd.ensemble import Ensemble, autotune, phase_scan_replicas
    from chllmodel_amd.lattice import Lattice
    from chllmodel_amd.states import isotropic_state

    reps = phase_scan_replicas()
    sites = sum(r["ni"] * r["nj"] * r["nk"] for r in reps)
    Kb, nb = 10, args.batched_blocks
    log(f"bench: batched scan, {len(reps)} replicas, {sites} sites, autotuning")
    tune = autotune(reps, bc="open", arith=args.arith, K=Kb, use_cache=False)
    ens = Ensemble(reps, bc="open", arith=args.arith)
    ens.set_stream(stream.cuda_stream)
    ens.set_block(tune["best"])
    ens.set_chunk(tune["chunk"])
    ens.init_isotropic(20261203)
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
        lat.upload_state(isotropic_state(r["ni"], r["nj"], r["nk"], 20261203 + i))
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
    return {"workload": "64 replicas, 5 PT ladders g={0,0.5,1,2,3} x L={32,48,64,96,128}, kT in [0.2,2] "
                        "geometric, K=1, swaps every 10 sweeps, one hipGraph replay per 10 sweeps",
            "replicas": len(reps), "sites": sites, "sweeps": nsw, "arith": args.arith,
            "site_updates_per_s": sites * nsw / t_b, "ms_per_sweep": 1e3 * t_b / nsw,
            "sequential_site_updates_per_s": sites * nsw / t_seq, "sequential_ms_per_sweep": 1e3 * t_seq / nsw,
            "speedup_vs_sequential": t_seq / t_b, "block_size": tune["best"],
            "chunk_bytes": tune["chunk"], "chunks": nchunks,
            "autotune_ms_per_sweep": tune["ms_per_sweep"], "autotune_ms_per_sweep_chunk": tune["ms_per_sweep_chunk"],
            "swap_attempts": int(att.sum()), "swap_acceptance": float(acc.sum() / max(att.sum(), 1)),
            "swap_note": "rungs a factor 1.21 apart in kT on lattices of 32^3-128^3 sites: beta-gaps times "
                         "energy gaps are far above 1, so swaps are attempted (and timed) but almost never "
                         "accepted; acceptance on small lattices is covered by the ensemble tests"}


if __name__ == "__main__":
    main()
