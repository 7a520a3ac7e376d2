"""Write the golden vectors under tests/golden/ from the reference itself.

The reference's chll/chll_mod.f90 is compiled from source by oracle/Makefile
(``make -C oracle ref``) into oracle/_ref/libchll_ref_stream.so, whose
random_number draws come from the documented uniform stream (or an explicit
buffer).  Each case runs the reference's own ``run`` on a seeded initial state
and stores inputs and outputs:

    nx0, ny0, nz0, s0          initial Fortran-ordered state
    nx, ny, nz, s              state after the run
    msteps, m                  the reference's measurement outputs
    d_out, stream_pos_out      adapted step, next draw index
    params                     [nstart, nstop, nout, g, KK, d, kbt, seed, stream_pos]

Run in the build container only (needs /root/reference):
    python tests/golden/make_golden.py [case names...]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Reference, copy_state, isotropic_state  # noqa: E402

# name, dims, init seed, nstart, nstop, nout, g, KK, d, kbt, stream seed, state kind
CASES = [
    ("ref_4x4x4_K1_g05", (4, 4, 4), 11, 1, 20, 5, 0.5, 1.0, 0.3, 0.7, 101, "iso"),
    ("ref_6x4x5_K05_g0", (6, 4, 5), 12, 1, 30, 10, 0.0, 0.5, 0.3, 1.0, 102, "iso"),
    ("ref_8x8x8_K2_g1", (8, 8, 8), 13, 1, 20, 4, 1.0, 2.0, 0.1, 0.3, 103, "iso"),
    ("ref_16x16x16_K1_g05", (16, 16, 16), 14, 1, 10, 5, 0.5, 1.0, 0.3, 1.5, 104, "iso"),
    ("ref_8x6x4_pole", (8, 6, 4), 15, 1, 12, 3, 0.25, 1.0, 0.05, 0.05, 105, "zpole"),
    ("ref_2x2x2", (2, 2, 2), 16, 1, 40, 8, 0.5, 1.5, 0.4, 0.5, 106, "iso"),
    ("ref_4x4x4_adapt_stop", (4, 4, 4), 17, 9991, 10010, 5, 0.5, 1.0, 0.3, 0.2, 107, "iso"),
    ("ref_10x2x3_negK", (10, 2, 3), 18, 1, 16, 4, -0.75, -1.0, 0.5, 0.9, 108, "iso"),
    # odd ni with N even: a row holds one more site of one colour than of the
    # other, so a site's sublattice rank is not l / 2 (chll/chll_mod.f90:23-36)
    ("ref_3x2x2_oddni", (3, 2, 2), 19, 1, 24, 6, 0.5, 1.0, 0.3, 0.7, 109, "iso"),
    ("ref_5x4x6_oddni", (5, 4, 6), 20, 1, 15, 5, 0.25, 1.5, 0.3, 0.6, 110, "iso"),
    ("ref_7x6x2_oddni", (7, 6, 2), 22, 1, 12, 4, 1.0, -0.8, 0.2, 1.1, 111, "iso"),
    ("ref_9x3x4_oddni", (9, 3, 4), 23, 1, 10, 5, 0.5, 1.0, 0.3, 0.8, 112, "iso"),
]

BUFFER_CASES = [
    # explicit host uniforms (numpy MT19937), reference draw schedule
    ("ref_buf_4x4x6", (4, 4, 6), 21, 1, 6, 2, 0.5, 1.0, 0.3, 0.8, 2024),
    ("ref_buf_5x3x4_oddni", (5, 3, 4), 24, 1, 6, 3, 0.5, 1.0, 0.3, 0.8, 2025),
]


def make_state(dims, seed, kind):
    st = isotropic_state(*dims, seed)
    if kind == "zpole":
        # near the +z / -z poles so both trial branches (|nz| > 0.999 and not) are taken
        rs = np.random.RandomState(seed + 1000)
        eps = rs.uniform(-0.03, 0.03, size=dims + (2,))
        sign = np.where(rs.random_sample(dims) < 0.5, -1.0, 1.0)
        nx, ny = eps[..., 0], eps[..., 1]
        nz = sign * np.sqrt(1.0 - nx * nx - ny * ny)
        st["nx"], st["ny"], st["nz"] = (np.asfortranarray(a) for a in (nx, ny, nz))
    return st


def main(only=None):
    """Write every case, or only the cases named in `only`."""
    ref = Reference(stream=True)
    for name, dims, iseed, nstart, nstop, nout, g, KK, d, kbt, sseed, kind in CASES:
        if only and name not in only:
            continue
        st0 = make_state(dims, iseed, kind)
        st = copy_state(st0)
        msteps, m, d_out, pos = ref.run(st, nstart, nstop, nout, g, KK, d, kbt, seed=sseed, stream_pos=0)
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            nx0=st0["nx"], ny0=st0["ny"], nz0=st0["nz"], s0=st0["s"],
                            nx=st["nx"], ny=st["ny"], nz=st["nz"], s=st["s"],
                            msteps=msteps, m=m, d_out=np.float64(d_out), stream_pos_out=np.uint64(pos),
                            params=np.array([nstart, nstop, nout, g, KK, d, kbt, sseed, 0], dtype=np.float64),
                            mode=np.array("stream"))
        print(name, "rows", len(msteps), "d_out", d_out)
    for name, dims, iseed, nstart, nstop, nout, g, KK, d, kbt, bseed in BUFFER_CASES:
        if only and name not in only:
            continue
        st0 = make_state(dims, iseed, "iso")
        st = copy_state(st0)
        nsub = int(np.prod(dims)) // 2
        buf = np.random.RandomState(bseed).random_sample(10 * nsub * (nstop - nstart + 1))
        msteps, m, d_out, pos = ref.run(st, nstart, nstop, nout, g, KK, d, kbt, buf=buf, stream_pos=0)
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            nx0=st0["nx"], ny0=st0["ny"], nz0=st0["nz"], s0=st0["s"],
                            nx=st["nx"], ny=st["ny"], nz=st["nz"], s=st["s"],
                            msteps=msteps, m=m, d_out=np.float64(d_out), stream_pos_out=np.uint64(pos),
                            params=np.array([nstart, nstop, nout, g, KK, d, kbt, bseed, 0], dtype=np.float64),
                            mode=np.array("buffer"), buffer=buf)
        print(name, "rows", len(msteps), "d_out", d_out)


if __name__ == "__main__":
    main(set(sys.argv[1:]))
