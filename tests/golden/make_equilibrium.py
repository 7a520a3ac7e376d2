"""Write equilibrium series under tests/golden/ from the reference itself.

The reference's chll/chll_mod.f90, compiled from source by oracle/Makefile
(``make -C oracle ref``) into oracle/_ref/libchll_ref.so, runs with its own
random_number.  Each case starts from a seeded isotropic state at sweep 10001,
so the step adaptation (istep < 10000, chll/chll_mod.f90:83-90) is off and the
chain is a plain Metropolis chain at fixed d, and keeps the reference's rows
(reported energy, eta, paccept, faccept) every nout sweeps after the first
`discard` rows.  tests/test_gpu_equilibrium.py compares the GPU's Philox
chains with these series within 3 standard errors.

Run in the build container only (needs /root/reference):
    python tests/golden/make_equilibrium.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Reference, isotropic_state  # noqa: E402

# name, dims, init seed, g, KK, d, kbt
CASES = [
    ("eq_10c_K1_g05_T07", (10, 10, 10), 21, 0.5, 1.0, 0.3, 0.7),
    ("eq_10c_K05_g0_T15", (10, 10, 10), 22, 0.0, 0.5, 0.4, 1.5),
    ("eq_12c_K15_g1_T10", (12, 12, 12), 23, 1.0, 1.5, 0.3, 1.0),
]
NSTART, NSWEEPS, NOUT, DISCARD = 10001, 20000, 10, 400


def main():
    ref = Reference(stream=False)
    for name, dims, seed, g, KK, d, kbt in CASES:
        st = isotropic_state(*dims, seed)
        msteps, m, d_out, _ = ref.run(st, NSTART, NSTART + NSWEEPS - 1, NOUT, g, KK, d, kbt)
        assert d_out == d, (name, d_out)
        series = m[DISCARD:, :4].copy()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), series=series, msteps=msteps[DISCARD:],
                            params=np.array([NSTART, NSWEEPS, NOUT, DISCARD, g, KK, d, kbt, seed], dtype=np.float64),
                            dims=np.array(dims, dtype=np.int32))
        print(name, "rows", series.shape[0], "means", series.mean(axis=0))


if __name__ == "__main__":
    main()
