"""Batched replicas with parallel tempering (libchll_mod.so chll_ensemble_*).

Parity: each replica's sweeps use the same site-keyed Philox counters as the
single-lattice path (replica id = its index), so after every block of K sweeps
each replica must equal the oracle run of those K sweeps at the temperature
slot (kT, d) it held during the block -- bit-exact spins and counts, directors
within 1e-12.  Swap decisions are recomputed from the oracle's energies with
the parallel-tempering rule min(1, exp((1/kT_t - 1/kT_{t+1})(E_a - E_b))).

hipGraph: the captured block is replayed several times with synchronisations
and host work in between, and every replay is checked against eager launches.
"""
import math

import numpy as np
import pytest

from chllmodel_amd.ensemble import Ensemble, series_by_temperature
from chllmodel_amd.lattice import Lattice
from chllmodel_amd.states import isotropic_state
from oracle.oracle import ARITH_COMPAT, ARITH_FP64, BC_OPEN, RNG_PHILOX, Oracle, copy_state

pytestmark = pytest.mark.gpu


def small_replicas():
    # rungs close enough that about half of the swap attempts in five blocks are
    # accepted (6 of 12 for the test below), so both outcomes are exercised
    reps = []
    for kT in (0.6, 0.7, 0.8):
        reps.append({"ni": 4, "nj": 4, "nk": 4, "ladder": 0, "g": 0.5, "KK": 1.0, "kbt": kT, "d0": 0.3})
    for kT in (0.9, 1.0, 1.1, 1.25):
        reps.append({"ni": 6, "nj": 4, "nk": 5, "ladder": 1, "g": 1.0, "KK": 0.7, "kbt": kT, "d0": 0.2})
    reps.append({"ni": 8, "nj": 6, "nk": 4, "ladder": 2, "g": 0.0, "KK": 1.5, "kbt": 0.8, "d0": 0.25})
    return reps


def pair_replicas():
    # every replica with ni % 4 == 0: the batched kernels take two sites per
    # thread, and reference arithmetic fuses through held trials
    reps = []
    for kT in (0.6, 0.7, 0.8):
        reps.append({"ni": 4, "nj": 4, "nk": 4, "ladder": 0, "g": 0.5, "KK": 1.0, "kbt": kT, "d0": 0.3})
    for kT in (0.9, 1.0, 1.1, 1.25):
        reps.append({"ni": 8, "nj": 4, "nk": 3, "ladder": 1, "g": 1.0, "KK": 0.7, "kbt": kT, "d0": 0.2})
    reps.append({"ni": 8, "nj": 6, "nk": 4, "ladder": 2, "g": 0.0, "KK": 1.5, "kbt": 0.8, "d0": 0.25})
    return reps


@pytest.mark.parametrize("arith,repset,fuse", [("compat", "small", 1), ("fp64", "small", 1), ("compat", "pairs", 1),
                                               ("compat", "pairs", 0)])
@pytest.mark.parametrize("use_graph", [False, True])
def test_ensemble_matches_oracle_with_swaps(arith, repset, fuse, use_graph):
    """Every block of K sweeps bit-exact against oracle runs that follow the
    GPU's slot permutation, with the swap decisions recomputed.  The "pairs"
    replicas all take site pairs, so reference arithmetic runs fused with
    held trials (fuse 1) or as four half-sweeps (fuse 0)."""
    orc = Oracle()
    reps = small_replicas() if repset == "small" else pair_replicas()
    seed, run_index, K, nblocks = 0xC0FFEE1234, 2, 3, 5
    n = len(reps)
    states = [isotropic_state(r["ni"], r["nj"], r["nk"], 500 + i) for i, r in enumerate(reps)]
    slot_kbt = np.array([r["kbt"] for r in reps])
    slot_d = np.array([r["d0"] for r in reps])
    rep_slot = np.arange(n)
    lad_start = [0, 3, 7, 8]
    oarith = ARITH_COMPAT if arith == "compat" else ARITH_FP64
    print(f"ensemble parity arith={arith} replicas={repset} fuse={fuse} graph={use_graph}", flush=True)
    with Ensemble(reps, bc="open", arith=arith) as ens:
        ens.set_fusion(bool(fuse))
        for r in range(n):
            ens.upload(r, states[r])
        for b in range(nblocks):
            rows = ens.run(1, K, seed=seed, run_index=run_index, adapt=True, use_graph=use_graph)
            st = ens.state()
            # oracle: K sweeps per replica at the slot it held during the block
            E = np.zeros(n)
            for r in range(n):
                s = rep_slot[r]
                _, _, d_out, _ = orc.run(states[r], 1 + b * K, (b + 1) * K, 10 ** 6, reps[r]["g"], reps[r]["KK"],
                                         slot_d[s], slot_kbt[s], rng=RNG_PHILOX, seed=seed, run_index=run_index,
                                         replica=r, bc=BC_OPEN, arith=oarith)
                slot_d[s] = d_out
                E[r] = orc.energy_true(states[r], reps[r]["g"], reps[r]["KK"])
                got = ens.download(r)
                assert np.array_equal(got["s"], states[r]["s"]), (b, r)
                err = max(np.abs(got[f] - states[r][f]).max() for f in ("nx", "ny", "nz"))
                assert err <= 1e-12, (b, r, err)
                N = reps[r]["ni"] * reps[r]["nj"] * reps[r]["nk"]
                assert abs(rows[0, r, 2] - E[r] / N) < 1e-12
                assert rows[0, r, 13] == s and rows[0, r, 1] == slot_kbt[s]
            assert np.array_equal(st["slot_d"], slot_d)
            # swap rule, round q = (istep of the block's last sweep) // K
            rnd = ((b + 1) * K) // K
            slot_rep = np.argsort(rep_slot)
            for l in range(len(lad_start) - 1):
                t0, t1 = lad_start[l], lad_start[l + 1]
                for t in range(t0 + (rnd & 1), t1 - 1, 2):
                    a, c = slot_rep[t], slot_rep[t + 1]
                    delta = (1 / slot_kbt[t] - 1 / slot_kbt[t + 1]) * (E[a] - E[c])
                    w = orc.philox([t, rnd, run_index, 0x50540000], [seed & 0xFFFFFFFF, seed >> 32])
                    u = ((int(w[1]) << 32 | int(w[0])) >> 11) * 2.0 ** -53
                    if delta >= 0 or u < math.exp(delta):
                        slot_rep[t], slot_rep[t + 1] = c, a
            rep_slot = np.argsort(slot_rep)
            assert np.array_equal(st["rep_slot"], rep_slot), (b, st["rep_slot"], rep_slot)
        assert st["istep"] == nblocks * K + 1
        if repset == "small":
            assert st["swap_acc"].sum() > 0 and st["swap_acc"].sum() < st["swap_att"].sum()
        else:
            assert st["swap_att"].sum() > 0


def test_graph_replays_equal_eager():
    reps = small_replicas()
    outs = {}
    for use_graph in (False, True):
        with Ensemble(reps, arith="fp64") as ens:
            ens.init_isotropic(77)
            snaps = []
            for b in range(4):
                rows = ens.run(1, 4, seed=9, use_graph=use_graph)
                _ = np.linalg.norm(rows)  # host work between replays
                snaps.append((rows.copy(), ens.download(3), ens.state()["rep_slot"].copy()))
            outs[use_graph] = snaps
    for (ra, sa, pa), (rb, sb, pb) in zip(outs[False], outs[True]):
        assert np.array_equal(ra, rb)
        assert np.array_equal(pa, pb)
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(sa[f], sb[f])


@pytest.mark.parametrize("K", [1, 2, 5])
def test_fused_blocks_equal_unfused(K):
    """FP64 blocks with the fused launches (D1 + C1 in one launch, C0 of
    sweep t with D0 of sweep t + 1 in another) give the same replicas, rows,
    slots and swap counts as four separate half-sweeps per sweep."""
    reps = small_replicas()
    outs = []
    for fuse in (False, True):
        print(f"ensemble fusion K={K} fuse={fuse}", flush=True)
        with Ensemble(reps, arith="fp64") as ens:
            ens.set_fusion(fuse)
            ens.init_isotropic(21)
            rows = ens.run(4, K, seed=13, use_graph=True)
            st = ens.state()
            outs.append((rows, [ens.download(r) for r in range(len(reps))], st))
    (ra, sa, ta), (rb, sb, tb) = outs
    assert np.array_equal(ra, rb)
    for a, b in zip(sa, sb):
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(a[f], b[f])
    for k in ("rep_slot", "slot_d", "swap_att", "swap_acc"):
        assert np.array_equal(ta[k], tb[k]), k


@pytest.mark.parametrize("K", [1, 3])
@pytest.mark.parametrize("layout", ["small", "banded"])
def test_pair_kernels_equal_one_site(K, layout):
    """Replicas whose ni are all multiples of 4 run the fused FP64 launches with
    two sites per thread (reverse block order on the colour-0 launch; the
    32x64 planes at 64 threads per block also split into XCD row bands): same
    replicas, rows, slots and swap counts as one site per thread and as four
    separate half-sweeps."""
    ladders = [((8, 6, 4), 0.5, 1.0, (0.3, 0.5, 0.9), 0.3), ((12, 5, 7), 1.0, 0.7, (0.4, 0.7), 0.2),
               ((4, 4, 6), 0.0, 1.5, (0.8,), 0.25)]
    if layout == "banded":
        ladders.insert(0, ((32, 64, 6), 0.5, 1.0, (0.5, 0.8), 0.3))
    reps = [{"ni": d[0], "nj": d[1], "nk": d[2], "ladder": lad, "g": g, "KK": KK, "kbt": kT, "d0": d0}
            for lad, (d, g, KK, kTs, d0) in enumerate(ladders) for kT in kTs]
    outs = []
    for fuse, pairs in ((True, True), (True, False), (False, True)):
        print(f"ensemble pairs {layout} K={K} fuse={fuse} pairs={pairs}", flush=True)
        with Ensemble(reps, arith="fp64") as ens:
            if layout == "banded":
                ens.set_block(64)
            ens.set_fusion(fuse, pairs=pairs)
            ens.init_isotropic(5)
            rows = ens.run(4, K, seed=17, use_graph=True)
            st = ens.state()
            outs.append((rows, [ens.download(r) for r in range(len(reps))], st))
    ra, sa, ta = outs[0]
    for rb, sb, tb in outs[1:]:
        assert np.array_equal(ra, rb)
        for a, b in zip(sa, sb):
            for f in ("nx", "ny", "nz", "s"):
                assert np.array_equal(a[f], b[f])
        for k in ("rep_slot", "slot_d", "swap_att", "swap_acc"):
            assert np.array_equal(ta[k], tb[k]), k


def test_pair_measure_rows_match_oracle():
    """Replicas whose half (ni / 2) is even all measure with four sites of one
    row per thread, marching slabs of 8 planes (k_measure_ens_march): x wraps
    at every pair (ni = 4), odd nj and nk, a partial slab (nk = 5, 10), rows
    longer than a wave (ni = 260).  After a block of sweeps,
    every replica's row against the oracle's observables of its final state:
    E_true / N within 1e-12, eta exact, the Q means within 1e-14, the
    reference's binary32 E within one binary32 ulp."""
    orc = Oracle()
    reps = [{"ni": 4, "nj": 3, "nk": 5, "ladder": 0, "g": 0.5, "KK": 1.0, "kbt": 0.7, "d0": 0.3},
            {"ni": 8, "nj": 6, "nk": 4, "ladder": 1, "g": 0.5, "KK": 1.0, "kbt": 0.8, "d0": 0.3},
            {"ni": 12, "nj": 10, "nk": 8, "ladder": 2, "g": 1.0, "KK": 0.7, "kbt": 0.9, "d0": 0.2},
            {"ni": 260, "nj": 3, "nk": 2, "ladder": 3, "g": 0.25, "KK": 1.5, "kbt": 0.6, "d0": 0.25},
            {"ni": 16, "nj": 7, "nk": 10, "ladder": 4, "g": 0.5, "KK": 1.0, "kbt": 0.75, "d0": 0.3}]
    with Ensemble(reps, bc="open", arith="fp64") as ens:
        for r, rep in enumerate(reps):
            ens.upload(r, isotropic_state(rep["ni"], rep["nj"], rep["nk"], 900 + r))
        rows = ens.run(1, 3, seed=31)
        for r, rep in enumerate(reps):
            st = ens.download(r)
            N = rep["ni"] * rep["nj"] * rep["nk"]
            row = rows[0, r]
            assert abs(row[2] - orc.energy_true(st, rep["g"], rep["KK"]) / N) < 1e-12, r
            assert row[3] == st["s"].sum() / N, r
            Q = [st["nx"] ** 2, st["ny"] ** 2, st["nz"] ** 2, st["nx"] * st["ny"], st["nx"] * st["nz"],
                 st["ny"] * st["nz"]]
            for c, q in enumerate(Q):
                assert abs(row[4 + c] - q.mean()) < 1e-14, (r, c)
            e = orc.etot(st, rep["g"], rep["KK"])
            assert abs(row[14] - e) <= np.spacing(np.float32(abs(e))), r


@pytest.mark.parametrize("arith,fuse", [("fp64", True), ("fp64", False), ("compat", False), ("compat", True)])
def test_chunks_equal_one_launch(arith, fuse):
    """A block of sweeps run chunk by chunk (a few replicas at a time, all K
    sweeps each), on one stream or with the chunks taking turns on two or
    three forked streams, eager or replayed from a graph, gives the same replicas,
    rows, slots and swap counts as every replica in one launch per
    half-sweep."""
    ladders = [((16, 12, 10), 0.5, 1.0, (0.3, 0.6, 1.0), 0.3), ((8, 8, 8), 1.0, 0.7, (0.4, 0.7), 0.2),
               ((12, 4, 6), 0.0, 1.5, (0.5, 0.8, 1.2), 0.25)]
    reps = [{"ni": d[0], "nj": d[1], "nk": d[2], "ladder": lad, "g": g, "KK": KK, "kbt": kT, "d0": d0}
            for lad, (d, g, KK, kTs, d0) in enumerate(ladders) for kT in kTs]
    outs = []
    for chunk, branches, graph in ((0, 1, True), (60000, 1, True), (1, 1, True), (60000, 2, True), (1, 2, True),
                                   (1, 2, False), (1, 3, True)):
        print(f"ensemble chunks {arith} fuse={fuse} chunk={chunk} branches={branches} graph={graph}", flush=True)
        with Ensemble(reps, arith=arith) as ens:
            ens.set_fusion(fuse)
            ens.set_chunk(chunk)
            ens.set_branches(branches)
            n = ens.chunks()
            assert n == {0: 1, 1: len(reps)}.get(chunk, n) and (chunk != 60000 or 1 < n < len(reps))
            ens.init_isotropic(9)
            rows = ens.run(4, 3, seed=23, use_graph=graph)
            st = ens.state()
            outs.append((rows, [ens.download(r) for r in range(len(reps))], st))
    ra, sa, ta = outs[0]
    for rb, sb, tb in outs[1:]:
        assert np.array_equal(ra, rb)
        for a, b in zip(sa, sb):
            for f in ("nx", "ny", "nz", "s"):
                assert np.array_equal(a[f], b[f])
        for k in ("rep_slot", "slot_d", "swap_att", "swap_acc"):
            assert np.array_equal(ta[k], tb[k]), k


def test_single_replica_equals_lattice_path():
    r = {"ni": 16, "nj": 12, "nk": 10, "ladder": 0, "g": 0.4, "KK": 1.2, "kbt": 0.7, "d0": 0.3}
    st = isotropic_state(16, 12, 10, 5)
    with Ensemble([r], arith="compat") as ens:
        ens.upload(0, st)
        ens.run(3, 5, seed=31, run_index=1)
        a = ens.download(0)
    with Lattice(16, 12, 10) as lat:
        lat.upload_state(st)
        lat.set_step(0.3)
        lat.sweeps(1, 15, 0, g=0.4, KK=1.2, kbt=0.7, seed=31, run_index=1, replica=0)
        b = lat.download()
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(a[f], b[f])


def test_block_sizes_agree():
    reps = small_replicas()
    res = []
    for bs in (64, 256, 512):
        with Ensemble(reps, arith="compat") as ens:
            ens.init_isotropic(3)
            ens.set_block(bs)
            rows = ens.run(3, 3, seed=4)
            res.append((rows, ens.download(5)))
    for rows, st in res[1:]:
        assert np.array_equal(rows, res[0][0])
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(st[f], res[0][1][f])


def test_k0_ladder_follows_tanh():
    """K = 0 decouples spins from directors: <s> = tanh(g/kT) at every rung,
    through swaps (a check of the swap bookkeeping on the observables)."""
    reps = [{"ni": 16, "nj": 16, "nk": 16, "ladder": 0, "g": 0.3, "KK": 0.0, "kbt": kT, "d0": 0.3}
            for kT in (0.5, 0.7, 1.0, 1.4)]
    with Ensemble(reps, arith="fp64") as ens:
        ens.init_isotropic(11)
        rows = ens.run(400, 5, seed=21)
    ser = series_by_temperature(rows[50:], reps)
    N = 16 ** 3
    for s, o in ser.items():
        want = math.tanh(0.3 / o["kT"])
        sig = math.sqrt((1 - want ** 2) / N / len(o["eta"]))
        assert abs(o["eta"].mean() - want) < 5 * sig, (s, o["eta"].mean(), want)


def test_upload_checks_shapes_and_indices():
    """Host arrays must match the replica's lattice (the library reads ni nj nk
    elements of each) and the replica index must exist; a good upload reads
    back unchanged."""
    reps = [{"ni": 4, "nj": 4, "nk": 2, "ladder": 0, "g": 0.5, "KK": 1.0, "kbt": 0.7},
            {"ni": 6, "nj": 4, "nk": 2, "ladder": 1, "g": 0.5, "KK": 1.0, "kbt": 0.7}]
    ens = Ensemble(reps, arith="fp64")
    try:
        st = isotropic_state(4, 4, 2, 5)
        with pytest.raises(ValueError, match="field shape"):
            ens.upload(1, st)
        with pytest.raises(IndexError):
            ens.upload(2, st)
        with pytest.raises(IndexError):
            ens.download(-1)
        ens.upload(0, st)
        back = ens.download(0)
        for k in ("nx", "ny", "nz", "s"):
            assert np.array_equal(back[k], st[k]), k
    finally:
        ens.close()
