"""Parity of the HIP sweep (libchll_mod.so on the GPU) with the reference.

Bar (deterministic modes): spins and accept/flip counts bit-exact, directors
within 1e-12, msteps / eta / paccept / faccept / d bit-exact, the reported
energy within one binary32 ulp (the reference sums it serially in binary64,
the GPU in a tree, before both round to binary32).

* golden vectors written by the reference's compiled source (tests/golden);
* the oracle on seeded inputs, for every mode: the reference's draw schedule
  fed by the documented stream or an explicit host buffer, the production
  site-keyed Philox draws, reference-compatible and FP64 arithmetic, open and
  periodic boundaries;
* the drop-in ``run`` entry point called with the reference's ctypes signature.
"""
import ctypes as ct
import glob
import os

import numpy as np
import pytest

from chllmodel_amd import _lib
from chllmodel_amd.lattice import Lattice
from oracle.oracle import (ARITH_COMPAT, ARITH_FP64, BC_OPEN, BC_PERIODIC, RNG_BUFFER, RNG_PHILOX, RNG_STREAM,
                           Oracle, copy_state, isotropic_state)

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_*.npz")))
DIR_TOL = 1e-12


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def f32_ulp(x):
    return float(np.spacing(np.float32(abs(x))))


def assert_rows_match(rows_gpu, msteps, m):
    """rows_gpu: (R,16) GPU rows; msteps, m: reference layout (R,), (R,5)."""
    assert rows_gpu.shape[0] == len(msteps)
    assert np.array_equal(rows_gpu[:, 0].astype(np.int32), msteps)
    for r in range(len(msteps)):
        assert abs(rows_gpu[r, 1] - m[r, 0]) <= f32_ulp(m[r, 0]), (r, rows_gpu[r, 1], m[r, 0])
    assert np.array_equal(rows_gpu[:, 2:6], m[:, 1:5])


def assert_state_match(a, b, tol=DIR_TOL):
    assert np.array_equal(a["s"], b["s"])
    for f in ("nx", "ny", "nz"):
        err = np.abs(a[f] - b[f]).max()
        assert err <= tol, (f, err)


def gpu_run(st0, bc, nstart, nstop, nout, d, want_counts=True, **kw):
    with Lattice(*st0["nx"].shape, bc=bc) as lat:
        lat.upload_state(st0)
        lat.set_step(d)
        out = lat.sweeps(nstart, nstop, nout, want_counts=want_counts, **kw)
        out["state"] = lat.download()
        out["d"] = lat.get_step()
    return out


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_vectors(path):
    z = np.load(path, allow_pickle=False)
    st0 = {k: np.asfortranarray(z[k + "0"]) for k in ("nx", "ny", "nz", "s")}
    p = z["params"]
    nstart, nstop, nout = int(p[0]), int(p[1]), int(p[2])
    kw = dict(g=p[3], KK=p[4], kbt=p[6], arith="compat", seed=int(p[7]), stream_pos=0)
    if str(z["mode"]) == "buffer":
        kw.update(rng="buffer", buffer=z["buffer"])
    else:
        kw.update(rng="stream")
    print(f"golden {os.path.basename(path)} dims={st0['nx'].shape}", flush=True)
    out = gpu_run(st0, "open", nstart, nstop, nout, float(p[5]), want_counts=False, **kw)
    assert_rows_match(out["rows"], z["msteps"], z["m"])
    assert out["d"] == float(z["d_out"])
    assert out["stream_pos"] == int(z["stream_pos_out"])
    ref = {k: z[k] for k in ("nx", "ny", "nz", "s")}
    assert_state_match(out["state"], ref)


CASES = [
    # dims, bc, rng, arith, nsweeps, nout
    ((4, 4, 4), "open", "stream", "compat", 12, 3),
    ((6, 4, 5), "open", "stream", "compat", 8, 2),
    ((32, 16, 8), "open", "stream", "compat", 6, 3),
    ((16, 16, 16), "periodic", "stream", "compat", 6, 2),
    ((8, 6, 4), "periodic", "philox", "compat", 10, 5),
    ((32, 32, 32), "open", "philox", "compat", 6, 3),
    ((24, 20, 18), "periodic", "philox", "compat", 5, 5),
    ((16, 16, 16), "open", "philox", "fp64", 6, 3),
    ((12, 8, 10), "periodic", "philox", "fp64", 6, 2),
    ((16, 8, 8), "open", "stream", "fp64", 5, 5),
    ((10, 6, 8), "periodic", "stream", "fp64", 6, 3),
    ((64, 64, 64), "open", "philox", "compat", 3, 3),
    # odd ni (open boundaries): rows hold (ni +- 1) / 2 sites of a colour
    ((5, 4, 6), "open", "philox", "compat", 6, 3),
    ((7, 3, 4), "open", "philox", "fp64", 6, 2),
    ((9, 5, 2), "open", "stream", "fp64", 5, 5),
    ((3, 2, 2), "open", "stream", "compat", 8, 4),
    ((63, 20, 12), "open", "philox", "fp64", 4, 2),
    ((33, 17, 10), "open", "philox", "compat", 4, 4),
    # padded XCD bands: 480 x 480 planes hold 225 pair blocks (the last of
    # each plane's 8 bands short)
    ((480, 480, 2), "open", "philox", "fp64", 4, 2),
    ((480, 480, 2), "periodic", "philox", "compat", 3, 3),
]


@pytest.mark.parametrize("dims,bc,rng,arith,nsw,nout", CASES,
                         ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}-{c[2]}-{c[3]}" for c in CASES])
def test_matches_oracle(orc, dims, bc, rng, arith, nsw, nout):
    st0 = isotropic_state(*dims, 1000 + sum(dims))
    g, K, d, kbt = 0.45, 1.3, 0.3, 0.6
    seed, run_index, replica = 0x1234ABCD5678, 3, 0
    print(f"oracle parity dims={dims} bc={bc} rng={rng} arith={arith}", flush=True)
    out = gpu_run(st0, bc, 1, nsw, nout, d, g=g, KK=K, kbt=kbt, rng=rng, arith=arith, seed=seed,
                  run_index=run_index)
    st = copy_state(st0)
    o_rng = {"stream": RNG_STREAM, "philox": RNG_PHILOX}[rng]
    ms, m, d_out, pos, counts = orc.run(st, 1, nsw, nout, g, K, d, kbt, rng=o_rng, seed=seed, run_index=run_index,
                                        replica=replica, bc={"open": BC_OPEN, "periodic": BC_PERIODIC}[bc],
                                        arith={"compat": ARITH_COMPAT, "fp64": ARITH_FP64}[arith],
                                        want_counts=True)
    assert np.array_equal(out["counts"], counts)
    assert_state_match(out["state"], st)
    assert_rows_match(out["rows"], ms, m)
    assert out["d"] == d_out
    if rng == "stream":
        assert out["stream_pos"] == pos


# Observable shapes: half even (four sites per thread, marching along k in
# slabs of planes) with pairs wrapping in x at every row end (ni = 4), odd nj
# and nk, rows longer than a wave (ni = 260: 65 pairs per row), single rows and
# planes, slabs marching up and down with a partial last slab (132 x 100 x 90,
# 8 x 6 x 17), slab counts that are multiples of 8 (16 x 8 x 64, 64 x 32 x 128:
# the slabs of one XCD contiguous); half odd (one site pair per thread):
# 10 x 6 x 4; odd ni: 5 x 4 x 6, 7 x 3 x 4, 3 x 1 x 2.
MEASURE_SHAPES = [(12, 10, 8), (4, 3, 5), (260, 3, 2), (8, 1, 1), (10, 6, 4), (132, 100, 90), (5, 4, 6),
                  (7, 3, 4), (3, 1, 2), (16, 8, 64), (64, 32, 128), (8, 6, 17)]


@pytest.mark.parametrize("dims", MEASURE_SHAPES, ids=["x".join(map(str, d)) for d in MEASURE_SHAPES])
def test_measure_row(orc, dims):
    st = isotropic_state(*dims, 77)
    g, K = 0.35, 1.7
    for bc, obc in (("open", BC_OPEN), ("periodic", BC_PERIODIC)):
        if bc == "periodic" and (dims[0] % 2 or dims[1] % 2 or dims[2] % 2):
            continue  # periodic boundaries need even ni, nj and nk
        with Lattice(*dims, bc=bc) as lat:
            lat.upload_state(st)
            row = lat.measure(g, K)
        N = np.prod(dims)
        assert abs(row["E"] - orc.etot(st, g, K)) <= f32_ulp(row["E"])
        # the oracle sums N site terms serially; its rounding error grows with N
        # (one wrong bond moves the mean by ~1/N, far above this bound)
        assert abs(row["E_true"] - orc.energy_true(st, g, K, bc=obc) / N) < 1e-13 * max(1.0, N / 1e5)
        assert row["eta_exact"] == st["s"].sum() / N
        Q = {"Qxx": st["nx"] ** 2, "Qyy": st["ny"] ** 2, "Qzz": st["nz"] ** 2, "Qxy": st["nx"] * st["ny"],
             "Qxz": st["nx"] * st["nz"], "Qyz": st["ny"] * st["nz"]}
        for k, v in Q.items():
            assert abs(row[k] - v.mean()) < 1e-14


def test_upload_download_roundtrip():
    dims = (10, 6, 7)
    st = isotropic_state(*dims, 5)
    st["s"][0, 0, 0] = 5
    st["s"][9, 5, 6] = -100
    with Lattice(*dims) as lat:
        lat.upload_state(st)
        back = lat.download()
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(back[f], st[f])


def test_upload_rejects_wide_spins():
    dims = (4, 4, 4)
    st = isotropic_state(*dims, 5)
    st["s"][1, 2, 3] = 300
    with Lattice(*dims) as lat:
        with pytest.raises(_lib.ChllError, match="spin values"):
            lat.upload_state(st)


def _dropin_run(st, nstart, nstop, nout, g, KK, d, kbt):
    ni, nj, nk = st["nx"].shape
    nsub = ni * nj * nk // 2
    rows = max(int((1 + nstop - nstart) / nout), 0)  # Fortran integer division (chll/chll_mod.f90:47)
    msteps = np.zeros(max(rows, 1), dtype=np.int32)
    m = np.zeros((max(rows, 1), 5), dtype=np.float64, order="F")
    sl1 = np.zeros(nsub, np.int32)
    sl2 = np.zeros(nsub, np.int32)
    dd = ct.c_double(d)
    ns = ct.c_int(0)
    _lib.lib().run(ct.c_int(nstart), ct.c_int(nstop), ct.c_int(nout), st["nx"], st["ny"], st["nz"], st["s"],
                   st["dope"], sl1, sl2, ct.c_double(g), ct.c_double(KK), dd, ct.c_double(kbt),
                   ct.c_int(ni), ct.c_int(nj), ct.c_int(nk), ns, msteps, m)
    assert _lib.last_error() == ""
    return msteps[:rows], m[:rows], dd.value, ns.value


@pytest.mark.parametrize("dims,nstart,nstop,nout", [((6, 4, 5), 1, 7, 20), ((6, 4, 5), 5, 2, 4), ((2, 1, 1), 1, 6, 2),
                                                    ((2, 1, 3), 1, 9, 3), ((4, 3, 1), 3, 11, 4)])
def test_dropin_run_edge_cases(orc, dims, nstart, nstop, nout):
    """Calls the reference's interface allows: nout longer than the run (the
    sweeps run, no row is written), an empty range nstop < nstart (nothing
    changes), and the smallest lattices (2x1x1: one site of each colour; flat
    4x3x1), in the deterministic stream mode against the oracle."""
    L = _lib.lib()
    st0 = isotropic_state(*dims, 77 + sum(dims))
    a, b = copy_state(st0), copy_state(st0)
    _lib.check(L.chll_configure(1, 0, 0, 0))
    _lib.check(L.chll_set_stream(91, 0))
    try:
        ms, m, d_out, ns = _dropin_run(a, nstart, nstop, nout, 0.5, 1.0, 0.3, 0.7)
        oms, om, od, _ = orc.run(b, nstart, nstop, nout, 0.5, 1.0, 0.3, 0.7, rng=RNG_STREAM, seed=91, stream_pos=0)
    finally:
        _lib.check(L.chll_configure(0, 0, 0, 1))
    assert ns == np.prod(dims) // 2
    assert np.array_equal(ms, oms) and d_out == od
    assert np.array_equal(m[:, 1:], om[:, 1:])
    for r in range(m.shape[0]):
        assert abs(m[r, 0] - om[r, 0]) <= np.spacing(np.float32(abs(om[r, 0])))
    assert np.array_equal(a["s"], b["s"])
    for f in ("nx", "ny", "nz"):
        assert np.abs(a[f] - b[f]).max() <= DIR_TOL, f
    if nstop < nstart:
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(a[f], st0[f]), f
        assert d_out == 0.3


@pytest.mark.parametrize("name", ["ref_5x4x6_oddni", "ref_7x6x2_oddni"])
def test_dropin_run_odd_ni_golden(name):
    """The drop-in run() on an odd-ni lattice (chll/chll_mod.f90:23-36 builds
    its lists for any N even) in the deterministic stream mode reproduces the
    reference's own outputs on that shape."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"), allow_pickle=False)
    st = {k: np.asfortranarray(z[k + "0"].copy()) for k in ("nx", "ny", "nz", "s")}
    st["dope"] = np.zeros_like(st["s"])
    p = z["params"]
    L = _lib.lib()
    _lib.check(L.chll_configure(1, 0, 0, 0))
    _lib.check(L.chll_set_stream(int(p[7]), 0))
    print(f"drop-in run() {name}", flush=True)
    try:
        ms, m, d_out, ns = _dropin_run(st, int(p[0]), int(p[1]), int(p[2]), p[3], p[4], p[5], p[6])
    finally:
        _lib.check(L.chll_configure(0, 0, 0, 1))
    assert ns == st["s"].size // 2
    assert_rows_match(np.column_stack([ms, m, np.zeros((len(ms), 10))]), z["msteps"], z["m"])
    assert d_out == float(z["d_out"])
    assert_state_match(st, {k: z[k] for k in ("nx", "ny", "nz", "s")})


def test_dropin_run_stream_mode(orc):
    """The reference's own calling convention (chll/chll.py:160-173) in the
    deterministic stream mode, twice in a row: the second call continues the
    stream exactly as the reference build continues its random_number draws."""
    L = _lib.lib()
    dims = (8, 6, 10)
    st0 = isotropic_state(*dims, 404)
    a, b = copy_state(st0), copy_state(st0)
    _lib.check(L.chll_configure(1, 0, 0, 0))
    _lib.check(L.chll_set_stream(55, 0))
    try:
        for call in range(2):
            ms, m, d_out, ns = _dropin_run(a, 1, 12, 4, 0.5, 1.0, 0.3, 0.7)
            pos0 = 0 if call == 0 else pos
            oms, om, od, pos = orc.run(b, 1, 12, 4, 0.5, 1.0, 0.3, 0.7, rng=RNG_STREAM, seed=55, stream_pos=pos0)
            assert ns == np.prod(dims) // 2
            assert np.array_equal(ms, oms) and d_out == od
            for r in range(len(ms)):
                assert abs(m[r, 0] - om[r, 0]) <= f32_ulp(om[r, 0])
            assert np.array_equal(m[:, 1:], om[:, 1:])
            assert_state_match(a, b)
            assert L.chll_get_stream_pos() == pos
    finally:
        _lib.check(L.chll_configure(0, 0, 0, 1))


def test_dropin_run_philox_run_counter(orc):
    """Philox mode: consecutive run() calls use run_index 0, 1, ... after
    chll_set_seed, so call n equals the oracle with run_index n."""
    L = _lib.lib()
    dims = (6, 6, 6)
    st0 = isotropic_state(*dims, 405)
    a, b = copy_state(st0), copy_state(st0)
    _lib.check(L.chll_configure(0, 0, 0, 0))
    _lib.check(L.chll_set_seed(999, 0))
    try:
        for call in range(3):
            ms, m, d_out, _ = _dropin_run(a, 1, 6, 3, 0.2, 0.8, 0.25, 0.9)
            oms, om, od, _ = orc.run(b, 1, 6, 3, 0.2, 0.8, 0.25, 0.9, rng=RNG_PHILOX, seed=999, run_index=call)
            assert d_out == od and np.array_equal(m[:, 1:], om[:, 1:])
            assert_state_match(a, b)
    finally:
        _lib.check(L.chll_configure(0, 0, 0, 1))


SHAPES = [((256 // 8, 24, 20), "open"), ((48, 40, 30), "periodic"), ((10, 7, 5), "open"), ((130, 6, 9), "open"),
          ((18, 34, 4), "periodic"), ((64, 64, 8), "open"), ((128, 64, 6), "periodic"), ((256, 8, 6), "open"),
          ((256, 12, 4), "periodic"), ((480, 480, 2), "open"), ((200, 200, 4), "periodic")]


@pytest.mark.parametrize("dims,bc", SHAPES, ids=[f"{'x'.join(map(str, d))}-{b}" for d, b in SHAPES])
@pytest.mark.parametrize("arith", ["compat", "fp64"])
def test_kernel_shapes_identical(dims, bc, arith):
    """The one-site-per-thread kernels in both block orders, the two-sites-
    per-thread pair kernels (ni % 4 == 0; other shapes fall back to one site;
    tiles 6 and 7 walk the fused colour-0 launch's blocks in reverse),
    each with and without the fused launches (FP64), and every k-march tile /
    slab split give the same lattice, counts and rows (partial tiles, odd open
    dimensions; 64x64 and 128x64 planes split into XCD bands of 1 and 2
    blocks; 480x480 planes into padded bands, 225 pair blocks or 450 site
    blocks per plane with the last band short; 200x200 planes into padded
    bands of blocks that straddle planes)."""
    st0 = isotropic_state(*dims, 2000 + sum(dims))
    ref = None
    for variant, tile, nslab, fuse in [("site", 0, 1, False), ("site", 0, 1, True), ("site", 1, 1, True),
                                       ("site", 1, 1, False), ("site", 2, 1, True), ("site", 3, 1, True),
                                       ("site", 3, 1, False), ("site", 7, 1, True), ("site", 6, 1, True),
                                       ("site", 7, 1, False), ("march", 0, 1, True), ("march", 0, 3, True),
                                       ("march", 1, 2, True), ("march", 2, 1, True), ("march", 2, 64, True),
                                       ("march", 3, 1, True), ("march", 4, 1, True), ("march", 5, 1, True)]:
        print(f"kernel shapes dims={dims} bc={bc} {arith} {variant} tile={tile} nslab={nslab} fuse={fuse}",
              flush=True)
        with Lattice(*dims, bc=bc) as lat:
            lat.set_kernel(variant, tile, nslab)
            lat.set_fusion(fuse)
            lat.upload_state(st0)
            lat.set_step(0.3)
            out = lat.sweeps(1, 6, 3, g=0.4, KK=1.1, kbt=0.8, arith=arith, seed=77, want_counts=True)
            got = (lat.download(), out["counts"], out["rows"])
        if ref is None:
            ref = got
            continue
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(got[0][f], ref[0][f]), (variant, tile, nslab, fuse, f)
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])


@pytest.mark.parametrize("rng", ["philox", "stream"])
def test_fused_sweep_counts_and_timing(rng):
    """FP64 sweeps launch the fused colour-1 kernel once per sweep; with Philox
    draws the colour-0 chirality trials of sweeps 1-4 share launches with the
    director trials of sweeps 2-5 (sweep 5 is measured and last, so its
    chirality colour 0 runs alone).  The lattice, per-sweep counts and rows
    equal those of four separate half-sweeps.  Reference-arithmetic sweeps
    fuse the same way, each colour-1 launch followed by its held trials."""
    dims = (64, 32, 16)
    st0 = isotropic_state(*dims, 31)
    res = {}
    for fuse in (True, False):
        print(f"fused sweep rng={rng} fuse={fuse}", flush=True)
        with Lattice(*dims) as lat:
            lat.set_kernel("site", 1, 1, kinds="director")
            lat.set_fusion(fuse)
            lat.upload_state(st0)
            lat.set_step(0.3)
            lat.set_timing(True)
            out = lat.sweeps(1, 5, 5, g=0.5, KK=1.0, kbt=0.7, arith="fp64", rng=rng, seed=11, want_counts=True)
            t = lat.timing()
            res[fuse] = (lat.download(), out["counts"], out["rows"], t)
    t_on, t_off = res[True][3], res[False][3]
    launches = {k: t_on[k][1] for k in ("director", "chirality", "fused", "fused_next")}
    if rng == "philox":
        assert launches == {"director": 1, "chirality": 1, "fused": 5, "fused_next": 4}, launches
    else:
        assert launches == {"director": 5, "chirality": 5, "fused": 5, "fused_next": 0}, launches
    assert t_off["fused"][1] == 0 and t_off["fused_next"][1] == 0
    assert t_off["director"][1] == 10 and t_off["chirality"][1] == 10
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(res[True][0][f], res[False][0][f]), f
    assert np.array_equal(res[True][1], res[False][1]) and np.array_equal(res[True][2], res[False][2])
    with Lattice(*dims) as lat:
        lat.upload_state(st0)
        lat.set_timing(True)
        lat.sweeps(1, 2, 0, g=0.5, KK=1.0, kbt=0.7, arith="compat", rng=rng, seed=11, want_rows=False)
        t = lat.timing()
    launches = {k: t[k][1] for k in ("director", "chirality", "fused", "fused_next", "held")}
    if rng == "philox":
        assert launches == {"director": 1, "chirality": 1, "fused": 2, "fused_next": 1, "held": 2}, launches
    else:
        assert launches == {"director": 2, "chirality": 2, "fused": 2, "fused_next": 0, "held": 2}, launches


def test_pipelined_counts_across_calls():
    """Sweeps split over several calls (each call closes its last sweep) give
    the same lattice and per-sweep counts as one call, with and without the
    fused launches, measurement rows at every third sweep included."""
    dims = (32, 32, 16)
    st0 = isotropic_state(*dims, 77)
    kw = dict(g=0.3, KK=1.2, kbt=0.9, arith="fp64", rng="philox", seed=5)
    outs = []
    for fuse, splits in ((True, [(1, 10)]), (True, [(1, 4), (5, 5), (6, 10)]), (False, [(1, 10)])):
        print(f"pipelined counts fuse={fuse} splits={splits}", flush=True)
        with Lattice(*dims) as lat:
            lat.set_kernel("site", 1, 1, kinds="director")
            lat.set_fusion(fuse)
            lat.upload_state(st0)
            lat.set_step(0.3)
            counts, rows = [], []
            for a, b in splits:
                o = lat.sweeps(a, b, 3, want_counts=True, **kw)
                counts.append(o["counts"])
                rows.append(o["rows"])
            outs.append((lat.download(), np.concatenate(counts), np.concatenate(rows), lat.get_step()))
    for o in outs[1:]:
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(o[0][f], outs[0][0][f]), f
        assert np.array_equal(o[1], outs[0][1])
        assert np.array_equal(o[2], outs[0][2])
        assert o[3] == outs[0][3]


ANNEAL = [((32, 32, 16), "open", "philox", "fp64"), ((32, 32, 16), "open", "philox", "compat"),
          ((12, 10, 8), "periodic", "stream", "fp64"), ((10, 6, 8), "open", "stream", "compat"),
          ((7, 6, 4), "open", "philox", "fp64")]


@pytest.mark.parametrize("dims,bc,rng,arith", ANNEAL, ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}-{c[2]}-{c[3]}"
                                                           for c in ANNEAL])
def test_anneal_ramp(orc, dims, bc, rng, arith):
    """One call with a linear kT ramp (chll_params.kbt_end) equals one call
    per sweep at the temperatures anneal_kt gives, bit for bit (lattice, cone
    step, per-sweep counts, rows; with Philox draws and FP64 the ramped call
    keeps the fused launches across sweeps of different temperature), and
    equals the oracle run one sweep at a time at those temperatures, to the
    bar of the other oracle cases."""
    from chllmodel_amd.lattice import anneal_kt

    st0 = isotropic_state(*dims, 4000 + sum(dims))
    g, K, d, kt0, kt1, nsw, nout = 0.4, 1.2, 0.3, 1.5, 0.3, 24, 5
    seed, run_index = 0x5EED, 2
    kts = [anneal_kt(kt0, kt1, t, 1, nsw) for t in range(1, nsw + 1)]
    assert kts[0] == kt0 and kts[-1] == kt1 and all(a > b for a, b in zip(kts, kts[1:]))
    print(f"anneal ramp dims={dims} bc={bc} rng={rng} arith={arith}", flush=True)
    common = dict(g=g, KK=K, rng=rng, arith=arith, seed=seed, run_index=run_index)
    ramp = gpu_run(st0, bc, 1, nsw, nout, d, kbt=kt0, kbt_end=kt1, **common)
    # one call per sweep; a measured sweep asks for its row with nout = 1
    rows, counts, pos = [], [], 0
    with Lattice(*dims, bc=bc) as lat:
        lat.upload_state(st0)
        lat.set_step(d)
        for t in range(1, nsw + 1):
            out = lat.sweeps(t, t, 1 if t % nout == 0 else 10 ** 6, kbt=kts[t - 1], stream_pos=pos, want_counts=True,
                             **common)
            pos = out["stream_pos"]
            rows.extend(out["rows"])
            counts.append(out["counts"][0])
        per = {"state": lat.download(), "d": lat.get_step(), "rows": np.array(rows), "counts": np.array(counts)}
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(ramp["state"][f], per["state"][f]), f
    assert ramp["d"] == per["d"]
    assert np.array_equal(ramp["counts"], per["counts"]) and np.array_equal(ramp["rows"], per["rows"])
    # the oracle, one sweep at a time
    st = copy_state(st0)
    o = dict(rng={"stream": RNG_STREAM, "philox": RNG_PHILOX}[rng], seed=seed, run_index=run_index,
             bc={"open": BC_OPEN, "periodic": BC_PERIODIC}[bc],
             arith={"compat": ARITH_COMPAT, "fp64": ARITH_FP64}[arith], want_counts=True)
    dd, opos, ms_all, m_all, c_all = d, 0, [], [], []
    for t in range(1, nsw + 1):
        ms, m, dd, opos, c = orc.run(st, t, t, 1 if t % nout == 0 else 10 ** 6, g, K, dd, kts[t - 1],
                                     stream_pos=opos, **o)
        if t % nout == 0:
            ms_all.append(ms[0])
            m_all.append(m[0])
        c_all.append(np.asarray(c).reshape(-1)[:2])
    assert np.array_equal(ramp["counts"], np.array(c_all))
    assert_state_match(ramp["state"], st)
    assert_rows_match(ramp["rows"], np.array(ms_all, dtype=np.int32), np.array(m_all))
    assert ramp["d"] == dd


def test_anneal_ramp_rejects_bad_end():
    with Lattice(8, 8, 4) as lat:
        lat.set_step(0.3)
        for bad in (-1.0, float("inf"), float("nan")):
            with pytest.raises(_lib.ChllError, match="kbt_end"):
                lat.sweeps(1, 3, 0, g=0.5, KK=1.0, kbt=1.0, kbt_end=bad, arith="fp64")


# dims, bc, rng, K, g, kT, rough start (directors of lengths 0.5-2, spins in
# [-127, 127]), sweeps
SCREEN = [((32, 32, 16), "open", "philox", 1.0, 0.5, 0.2, False, 12),
          ((16, 16, 16), "periodic", "philox", 4.0, -1.5, 0.05, True, 12),
          ((12, 12, 8), "open", "stream", 1.5, 0.25, 0.7, True, 8),
          ((64, 64, 32), "open", "philox", 1.0, 0.5, 1.1, False, 6),
          ((16, 16, 8), "open", "philox", 1.0, 0.5, 0.0, False, 6)]


@pytest.mark.parametrize("dims,bc,rng,K,g,kT,rough,nsw", SCREEN,
                         ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}-{c[2]}-K{c[3]}-kT{c[5]}{'-rough' if c[6] else ''}"
                              for c in SCREEN])
def test_compat_screen_scales_agree(dims, bc, rng, K, g, kT, rough, nsw):
    """Reference arithmetic decides a trial from its binary64 energy
    difference where the binary32 roundings of the two site energies cannot
    change the outcome, and evaluates the rounded energies otherwise
    (compat_screen).  With the screen off (every trial rounds both energies),
    at its default scale 2^-21, and at 2^-11 and 2^-5 (a few per cent of the
    trials, then nearly all, take the rounded energies, so waves mix both
    paths) the lattice, cone step, counts and rows are the same bit for bit.
    Rough starts have directors of any length and spins up to +-127, which
    the reference's arithmetic accepts; kT = 0 is a quench, where the screen
    decides nothing."""
    rs = np.random.RandomState(sum(dims) + int(100 * kT))
    st = isotropic_state(*dims, 7000 + sum(dims))
    if rough:
        for f in ("nx", "ny", "nz"):
            st[f] = np.asfortranarray(st[f] * rs.uniform(0.5, 2.0, size=dims))
        st["s"] = np.asfortranarray(rs.randint(-127, 128, size=dims).astype(np.int32))
    print(f"compat screen dims={dims} bc={bc} rng={rng} K={K} g={g} kT={kT} rough={rough}", flush=True)
    res, held = [], []
    for scale in (0.0, 2.0 ** -21, 2.0 ** -11, 2.0 ** -5):
        with Lattice(*dims, bc=bc) as lat:
            lat.upload_state(st)
            lat.set_step(0.3)
            lat.set_screen(scale)
            out = lat.sweeps(1, nsw, 3, g=g, KK=K, kbt=kT, arith="compat", rng=rng, seed=0x5C4EE, run_index=1,
                             want_counts=True, adapt=True)
            res.append((lat.download(), lat.get_step(), out["counts"], out["rows"]))
            held.append(lat.held_trials())
    # with the screen on and kT > 0 the sweeps fuse, and the trials the screen
    # leaves open are held: at 2^-5 a large share of them
    print(f"held chirality trials per scale: {held}", flush=True)
    assert held[0] == 0
    if kT > 0:
        assert held[3] > 0
    else:
        assert held == [0, 0, 0, 0]
    base = res[0]
    for state, d, counts, rows in res[1:]:
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(state[f], base[0][f]), f
        assert d == base[1]
        assert np.array_equal(counts, base[2]) and np.array_equal(rows, base[3])
    c = np.asarray(base[2])
    assert c.sum() > 0 and c.min() < dims[0] * dims[1] * dims[2]  # decisions went both ways


FUSED_COMPAT = [((32, 32, 16), "open", "philox", 2.0 ** -21, 0.0),
                ((32, 32, 16), "open", "philox", 2.0 ** -8, 0.0),
                ((16, 16, 24), "periodic", "philox", 2.0 ** -8, 0.0),
                ((24, 12, 10), "open", "stream", 2.0 ** -8, 0.0),
                ((16, 16, 16), "periodic", "stream", 2.0 ** -21, 0.0),
                ((32, 16, 16), "open", "philox", 2.0 ** -8, 0.25)]


@pytest.mark.parametrize("dims,bc,rng,scale,kbt_end", FUSED_COMPAT,
                         ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}-{c[2]}-s{int(-np.log2(c[3]))}"
                              f"{'-ramp' if c[4] else ''}" for c in FUSED_COMPAT])
def test_compat_fused_equals_reference_order(dims, bc, rng, scale, kbt_end):
    """Reference arithmetic with colour 1's chirality trials in its director
    launch and held trials (fusion on) against the reference's order of four
    half-sweeps (fusion off): the same lattice, cone step, per-sweep counts
    and rows bit for bit, over sweeps split across calls, with measured
    sweeps inside and at the ends of calls, and with a kT ramp.  At scale
    2^-8 thousands of trials are held; the unfused lattice holds none."""
    st = isotropic_state(*dims, 4100 + sum(dims))
    kw = dict(g=0.5, KK=1.3, kbt=0.6, arith="compat", rng=rng, seed=0xF05ED, run_index=2, adapt=True)
    if kbt_end:
        kw["kbt_end"] = kbt_end
    splits = [(1, 5), (6, 6), (7, 16)] if not kbt_end else [(1, 16)]
    res = []
    for fuse in (True, False):
        print(f"compat fused dims={dims} bc={bc} rng={rng} scale={scale} kbt_end={kbt_end} fuse={fuse}",
              flush=True)
        with Lattice(*dims, bc=bc) as lat:
            lat.set_fusion(fuse)
            lat.set_screen(scale)
            lat.upload_state(st)
            lat.set_step(0.3)
            lat.set_timing(True)
            counts, rows = [], []
            for a, b in splits:
                o = lat.sweeps(a, b, 4, want_counts=True, **kw)
                counts.append(o["counts"])
                rows.append(o["rows"])
            t = lat.timing()
            res.append((lat.download(), lat.get_step(), np.concatenate(counts), np.concatenate(rows),
                        lat.held_trials(), t))
    on, off = res
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(on[0][f], off[0][f]), f
    assert on[1] == off[1]
    assert np.array_equal(on[2], off[2]) and np.array_equal(on[3], off[3])
    assert on[5]["fused"][1] == 16 and on[5]["held"][1] == 16 and off[5]["fused"][1] == 0
    assert off[4] == 0
    print(f"held trials: {on[4]}", flush=True)
    if scale > 2.0 ** -21:
        assert on[4] > 100


@pytest.mark.parametrize("scale", [2.0 ** -21, 2.0 ** -10], ids=["s21", "s10"])
def test_compat_fused_full_size(scale):
    """The L=256 workload (K=1, g=0.5, kT=0.7, Philox, nout=4) in reference
    arithmetic: 12 fused sweeps with held trials over two calls equal the
    reference's order of four half-sweeps bit for bit (lattice, cone step,
    counts, rows).  At the default scale about 90 trials a sweep are held; at
    2^-10 about 120,000, so the held-trial kernel runs dozens of passes over
    its queue."""
    L = 256
    st = isotropic_state(L, L, L, 20261203)
    kw = dict(g=0.5, KK=1.0, kbt=0.7, arith="compat", rng="philox", seed=20261203, adapt=True)
    res = []
    for fuse in (True, False):
        print(f"compat full size L={L} scale={scale} fuse={fuse}", flush=True)
        with Lattice(L, L, L) as lat:
            lat.set_fusion(fuse)
            lat.set_screen(scale)
            lat.upload_state(st)
            lat.set_step(0.3)
            o1 = lat.sweeps(1, 7, 4, want_counts=True, **kw)
            o2 = lat.sweeps(8, 12, 4, want_counts=True, **kw)
            res.append((lat.download(), lat.get_step(), np.concatenate([o1["counts"], o2["counts"]]),
                        np.concatenate([o1["rows"], o2["rows"]]), lat.held_trials()))
    on, off = res
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(on[0][f], off[0][f]), f
    assert on[1] == off[1]
    assert np.array_equal(on[2], off[2]) and np.array_equal(on[3], off[3])
    print(f"held trials: {on[4]} over 12 sweeps", flush=True)
    assert off[4] == 0 and on[4] > (12 * 500 if scale > 2.0 ** -21 else 12 * 20)


def test_compat_screen_scale_validation():
    with Lattice(8, 8, 4) as lat:
        for bad in (2.0 ** -22, -1.0, 2.0, float("nan"), float("inf")):
            with pytest.raises(_lib.ChllError, match="screen"):
                lat.set_screen(bad)
        for ok in (0.0, 2.0 ** -21, 1.0):
            lat.set_screen(ok)
