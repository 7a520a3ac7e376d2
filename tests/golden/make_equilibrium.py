"""Write equilibrium series under tests/golden/ from the reference itself.

The reference's chll/chll_mod.f90, compiled from source by oracle/Makefile
(``make -C oracle ref``) into oracle/_ref/libchll_ref.so, runs with its own
random_number.  Each case starts from a seeded isotropic state at sweep 10001,
so the step adaptation (istep < 10000, chll/chll_mod.f90:83-90) is off and the
chain is a plain Metropolis chain at fixed d, and keeps the reference's rows
(reported energy, eta, paccept, faccept) every nout sweeps after the first
`discard` rows.  tests/test_gpu_equilibrium.py compares the GPU's Philox
chains with these series within 3 standard errors.

The reference records no order parameter, and its run() calls random_seed()
on entry, which in this flang build restarts the same sequence on every call,
so a chain cannot be cut into calls to read its directors.  For <P2> each case
also runs the reference built with the documented uniform stream
(oracle/_ref/libchll_ref_stream.so), one call per nout sweeps continuing the
stream, and keeps the means <n_a n_b> (xx, yy, zz, xy, xz, yz) of the state
after each call (q_series) beside that chain's rows (stream_series).

Run in the build container only (needs /root/reference):
    python tests/golden/make_equilibrium.py [case names]
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
    # the L=32 workload's couplings (K = 0.5, g = 0, kT = 1) at its full size,
    # with the reference's open boundaries (its sweep has no periodic mode)
    ("eq_32c_K05_g0_T10", (32, 32, 32), 24, 0.0, 0.5, 0.3, 1.0),
    # two temperatures of the batched scan's first ladder (L = 32, g = 0, K = 1,
    # 13 temperatures geometric in [0.2, 2]): slots 8 and 11, in its disordered
    # phase; the GPU test reads them off the ensemble slot by slot
    ("eq_32c_K1_g0_scan08", (32, 32, 32), 25, 0.0, 1.0, 0.3, float(np.geomspace(0.2, 2.0, 13)[8])),
    ("eq_32c_K1_g0_scan11", (32, 32, 32), 26, 0.0, 1.0, 0.3, float(np.geomspace(0.2, 2.0, 13)[11])),
]
NSTART, NSWEEPS, NOUT, DISCARD = 10001, 20000, 10, 400
STREAM_SEED = 0x51E4


def q_sums(st):
    nx, ny, nz = st["nx"], st["ny"], st["nz"]
    return [np.mean(nx * nx), np.mean(ny * ny), np.mean(nz * nz), np.mean(nx * ny), np.mean(nx * nz),
            np.mean(ny * nz)]


def stream_chain(dims, seed, g, KK, d, kbt):
    """The reference on the documented stream, one call per NOUT sweeps."""
    ref = Reference(stream=True)
    st = isotropic_state(*dims, seed)
    pos, qs, rows = 0, [], []
    for c in range(NSWEEPS // NOUT):
        a = NSTART + c * NOUT
        _, m, d_out, pos = ref.run(st, a, a + NOUT - 1, NOUT, g, KK, d, kbt, seed=STREAM_SEED + seed,
                                   stream_pos=pos)
        assert d_out == d
        qs.append(q_sums(st))
        rows.append(m[0, :4])
    return np.array(qs)[DISCARD:], np.array(rows)[DISCARD:]


def main(only=None):
    ref = Reference(stream=False)
    for name, dims, seed, g, KK, d, kbt in CASES:
        if only and name not in only:
            continue
        st = isotropic_state(*dims, seed)
        msteps, m, d_out, _ = ref.run(st, NSTART, NSTART + NSWEEPS - 1, NOUT, g, KK, d, kbt)
        assert d_out == d, (name, d_out)
        series = m[DISCARD:, :4].copy()
        q_series, stream_series = stream_chain(dims, seed, g, KK, d, kbt)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), series=series, msteps=msteps[DISCARD:],
                            params=np.array([NSTART, NSWEEPS, NOUT, DISCARD, g, KK, d, kbt, seed], dtype=np.float64),
                            dims=np.array(dims, dtype=np.int32), q_series=q_series, stream_series=stream_series)
        print(name, "rows", series.shape[0], "means", series.mean(axis=0), "stream chain", stream_series.mean(axis=0))


if __name__ == "__main__":
    main(sys.argv[1:])
