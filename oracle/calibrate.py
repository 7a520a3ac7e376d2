"""Calibrate the CPU restatement against the reference build, on one core.

TEST INFRASTRUCTURE ONLY (the CPU-baseline leg of bench.py reads its output).

Times, on the same host and the same workload, the reference's own run()
(chll/chll_mod.f90:40-115, compiled from source by ``make -C oracle ref`` into
oracle/_ref/libchll_ref.so, serial, drawing from its own random_number) and the
C restatement (oracle_run_mt on one thread, compat arithmetic, every sweep's
10 nsub uniforms filled into a host array first as the reference's
random_number calls do).  Writes oracle/calibration.json with the ratio
t_reference / t_restatement, so that a restatement timing taken where no
reference build exists (the GPU host) can be stated in reference time.

Run in the build container (needs /root/reference):
    python oracle/calibrate.py
"""
from __future__ import annotations

import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, sys.path[0] is this directory, where oracle.py would shadow the package
sys.path[0] = os.path.dirname(HERE)

from oracle.oracle import RNG_STREAM, Oracle, Reference, copy_state, isotropic_state  # noqa: E402

OUT = os.path.join(HERE, "calibration.json")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main(L: int = 128, sweeps: int = 4, repeats: int = 3, g=0.5, K=1.0, d0=0.3, kT=0.7) -> dict:
    st0 = isotropic_state(L, L, L, 20261203)
    ref = Reference(stream=False)
    orc = Oracle()

    def t_ref():
        st = copy_state(st0)
        t0 = time.perf_counter()
        # nstart = 2 skips the header line; nout past the range prints no rows
        ref.run(st, 2, 1 + sweeps, 10 ** 9, g, K, d0, kT)
        return time.perf_counter() - t0

    def t_port():
        st = copy_state(st0)
        t0 = time.perf_counter()
        orc.run(st, 2, 1 + sweeps, 10 ** 9, g, K, d0, kT, rng=RNG_STREAM, seed=1, threads=1, hostfill=True)
        return time.perf_counter() - t0

    tr, tp = [], []
    for _ in range(repeats):  # interleaved, best of `repeats`
        tr.append(t_ref())
        tp.append(t_port())
    res = {
        "workload": f"L={L} open boundaries, K={K} g={g} kT={kT} d0={d0}, {sweeps} sweeps, compat arithmetic",
        "reference": "chll/chll_mod.f90 run() built by oracle/Makefile with flang -O2, serial, own random_number",
        "restatement": "oracle_run_mt, 1 thread, host-filled stream uniforms (10 nsub per sweep)",
        "cpu": cpu_model(),
        "t_reference_s": min(tr),
        "t_restatement_s": min(tp),
        "reference_over_restatement": min(tr) / min(tp),
        "site_updates_per_s_reference": L ** 3 * sweeps / min(tr),
        "site_updates_per_s_restatement": L ** 3 * sweeps / min(tp),
    }
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
        f.write("\n")
    print(json.dumps(res, indent=1))
    return res


if __name__ == "__main__":
    main()
