"""Two workload configurations at their full sizes, and the drop-in run()
called from two host threads at once.

* L=32 periodic (32,768 sites), K = 0.5, g = 0, kT = 1.0, d0 = 0.3, 1000
  sweeps with a row every 10, site-keyed Philox draws, in both arithmetic
  modes, against the oracle run of the same sweeps: every sweep's accept and
  flip counts and the final spins exact, directors within 1e-12, the 100 rows'
  eta / paccept / faccept / d exact and the reported energy within one
  binary32 ulp (it is a binary64 sum taken in another order before both round).
* The 64-replica batched scan (phase_scan_replicas: ladders g = 0, 0.5, 1, 2, 3
  on L = 32, 48, 64, 96, 128, kT geometric in [0.2, 2], 41.9 M sites), three
  blocks of 10 sweeps with parallel-tempering swaps, replayed from a hipGraph:
  - the 13 replicas of the L = 32 ladder bit-exact against oracle runs that
    follow the slot each replica held in each block (kT and adapted d of
    that slot), every swap decision recomputed from the oracle's energies;
  - the swap decisions of the other ladders recomputed from the rows' energies;
  - every replica's directors unit vectors, spins +-1, and the last row's
    Hamiltonian per site equal to the independent energy kernel's to 1e-12.
* run() from two host threads at once on different lattices (each call takes
  its own device lattice and its own run counter): each call's result equals
  the oracle run with one of the two run counters, and the two differ.
* The L=480 coarsening run at its full length (110.6 M sites, K = 1,
  g = 0.25, kT = 0.5, 5000 sweeps from a random state generated on the
  device, a row every 50 sweeps): no oracle finishes this, so the checks are
  the size-independent ones -- every row finite and in range, the adapted
  acceptance in the 0.4 - 0.6 band, the energy lowered by the coarsening,
  unit directors and +-1 spins at the end, and the last row's Hamiltonian
  equal to the independent energy kernel's on the final state to 1e-12.
* The 64-replica scan at its full length, 1000 blocks of 10 sweeps (10k
  sweeps of every replica, swaps every 10, one graph replay per block, chunks
  of 128 MiB on two graph branches as the bench runs it): every
  row finite, every temperature slot's C_v (jackknife) finite and positive
  and chi_c finite and non-negative (0 where a strong field saturates the
  spins), the sweep counter at 10001, and every replica's last row against
  the energy kernel on its final state to 1e-12, with unit directors and +-1
  spins.
* The L=128 anneal at its full length: kT from 1.5 to 0.3 linearly over
  20,000 sweeps, one call per sweep with that sweep's kT (the reference's
  first 50 sweeps are compared bit for bit in test_gpu_configs.py): a row
  every 100 sweeps finite, the energy lowered as kT falls, unit directors and
  +-1 spins at the end, and the last row against the energy kernel to 1e-12.
* The differentiable-energy workload at its full length: 16 lattices of
  128^3 with K in [0.5, 2] and g in [0, 1], a 500-step gradient quench
  (final energy below the initial one) and then 2000 exact Metropolis
  sweeps of every lattice at kT = 0.2, whose last rows must match the energy
  kernel on the final states to 1e-12.
"""
import ctypes as ct
import math
import os
import threading

import numpy as np
import pytest
import torch

from chllmodel_amd import _lib
from chllmodel_amd.autograd import energy_and_grad, quench_then_metropolis
from chllmodel_amd.ensemble import Ensemble, phase_scan_replicas, series_by_temperature
from chllmodel_amd.lattice import Lattice
from chllmodel_amd.observables import p2_from_q, summarize
from chllmodel_amd.states import isotropic_state
from oracle.oracle import ARITH_COMPAT, ARITH_FP64, BC_OPEN, BC_PERIODIC, RNG_PHILOX, Oracle, copy_state

pytestmark = pytest.mark.gpu


def host_threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def f32_ulp(x):
    return float(np.spacing(np.float32(abs(x))))


@pytest.mark.parametrize("arith", ["compat", "fp64"])
def test_l32_periodic_1000_sweeps_match_oracle(arith):
    L, K, g, kT, d0, seed = 32, 0.5, 0.0, 1.0, 0.3, 20261203
    nsw, nout = 1000, 10
    st0 = isotropic_state(L, L, L, seed)
    print(f"L=32 periodic, {arith}: 1000 GPU sweeps", flush=True)
    with Lattice(L, L, L, bc="periodic") as lat:
        lat.upload_state(st0)
        lat.set_step(d0)
        out = lat.sweeps(1, nsw, nout, g=g, KK=K, kbt=kT, rng="philox", arith=arith, seed=seed,
                         want_counts=True)
        got = lat.download()
        d_gpu = lat.get_step()
    print(f"L=32 periodic, {arith}: oracle", flush=True)
    st = copy_state(st0)
    ms, m, d_ref, _, counts = Oracle().run(st, 1, nsw, nout, g, K, d0, kT, rng=RNG_PHILOX, seed=seed,
                                           bc=BC_PERIODIC, arith=ARITH_COMPAT if arith == "compat" else ARITH_FP64,
                                           want_counts=True, threads=host_threads())
    rows = out["rows"]
    assert rows.shape[0] == nsw // nout == len(ms)
    assert np.array_equal(out["counts"], counts)
    assert np.array_equal(rows[:, 0].astype(np.int32), ms)
    for r in range(len(ms)):
        assert abs(rows[r, 1] - m[r, 0]) <= f32_ulp(m[r, 0]), r
    assert np.array_equal(rows[:, 2:6], m[:, 1:5])
    assert d_gpu == d_ref
    assert np.array_equal(got["s"], st["s"])
    err = max(np.abs(got[f] - st[f]).max() for f in ("nx", "ny", "nz"))
    assert err <= 1e-12, err
    # the run is not trivial: both kinds of trial are accepted and rejected
    pacc = counts[:, 0] / L ** 3
    facc = counts[:, 1] / L ** 3
    assert np.all((pacc > 0.05) & (pacc < 0.95)) and np.all((facc > 0.05) & (facc < 0.95))


def _swap_rounds(rep_slot, slot_kbt, energies, lad_start, rnd, seed, run_index, orc):
    """Slot permutation after one swap round (the parallel-tempering rule
    min(1, exp((1/kT_t - 1/kT_{t+1}) (E_a - E_b))), even or odd neighbour pairs
    of each ladder by round parity, uniforms from Philox counter
    (slot, round, run_index, 0x50540000))."""
    slot_rep = np.argsort(rep_slot)
    for lo, hi in zip(lad_start[:-1], lad_start[1:]):
        for t in range(lo + (rnd & 1), hi - 1, 2):
            a, c = slot_rep[t], slot_rep[t + 1]
            delta = (1 / slot_kbt[t] - 1 / slot_kbt[t + 1]) * (energies[a] - energies[c])
            w = orc.philox([t, rnd, run_index, 0x50540000], [seed & 0xFFFFFFFF, seed >> 32])
            u = ((int(w[1]) << 32 | int(w[0])) >> 11) * 2.0 ** -53
            if delta >= 0 or u < math.exp(delta):
                slot_rep[t], slot_rep[t + 1] = c, a
    return np.argsort(slot_rep)


@pytest.mark.timeout(900)
def test_phase_scan_64_replicas():
    reps = phase_scan_replicas()
    n = len(reps)
    assert n == 64 and sum(r["ni"] * r["nj"] * r["nk"] for r in reps) == 41938944
    lad_start = [0]
    for i in range(1, n):
        if reps[i]["ladder"] != reps[i - 1]["ladder"]:
            lad_start.append(i)
    lad_start.append(n)
    small = range(lad_start[0], lad_start[1])  # the L = 32 ladder
    assert all(reps[r]["ni"] == 32 for r in small) and len(small) == 13
    orc = Oracle()
    seed, run_index, K, nblocks = 20261203, 0, 10, 3
    print("phase scan: initial states", flush=True)
    states = {r: orc.init_state(32, 32, 32, kind=0, seed=seed, replica=r, p_flip=0.5) for r in small}
    slot_kbt = np.array([r["kbt"] for r in reps])
    slot_d = np.array([r["d0"] for r in reps])
    rep_slot = np.arange(n)
    with Ensemble(reps, bc="open", arith="fp64") as ens:
        ens.init_state("isotropic", seed=seed, p_flip=0.5)  # on the device, keyed by (seed, replica r, site)
        for r in small:
            got = ens.download(r)
            assert np.array_equal(got["s"], states[r]["s"])
            assert max(np.abs(got[f] - states[r][f]).max() for f in ("nx", "ny", "nz")) <= 1e-15
            states[r] = got  # the oracle continues from the device's state
        for b in range(nblocks):
            print(f"phase scan: block {b}", flush=True)
            rows = ens.run(1, K, seed=seed, run_index=run_index, adapt=True, use_graph=True)
            st = ens.state()
            E = np.array([rows[0, r, 2] * reps[r]["ni"] * reps[r]["nj"] * reps[r]["nk"] for r in range(n)])
            for r in small:
                s = rep_slot[r]
                _, _, d_out, _ = orc.run(states[r], 1 + b * K, (b + 1) * K, 10 ** 6, reps[r]["g"], reps[r]["KK"],
                                         slot_d[s], slot_kbt[s], rng=RNG_PHILOX, seed=seed, run_index=run_index,
                                         replica=r, bc=BC_OPEN, arith=ARITH_FP64, threads=host_threads())
                slot_d[s] = d_out
                got = ens.download(r)
                assert np.array_equal(got["s"], states[r]["s"]), (b, r)
                err = max(np.abs(got[f] - states[r][f]).max() for f in ("nx", "ny", "nz"))
                assert err <= 1e-12, (b, r, err)
                E[r] = orc.energy_true(states[r], reps[r]["g"], reps[r]["KK"])
                assert abs(rows[0, r, 2] - E[r] / 32 ** 3) < 1e-12, (b, r)
                assert rows[0, r, 13] == s and rows[0, r, 1] == slot_kbt[s]
            assert np.array_equal(st["slot_d"][list(small)], slot_d[list(small)])
            rep_slot = _swap_rounds(rep_slot, slot_kbt, E, lad_start, b + 1, seed, run_index, orc)
            assert np.array_equal(st["rep_slot"], rep_slot), (b, st["rep_slot"], rep_slot)
        assert st["istep"] == nblocks * K + 1
        print("phase scan: invariants and energies of all 64 replicas", flush=True)
        for r in range(n):
            got = ens.download(r)
            N = reps[r]["ni"] * reps[r]["nj"] * reps[r]["nk"]
            assert np.all(np.isin(got["s"], (-1, 1))), r
            norm = np.sqrt(got["nx"] ** 2 + got["ny"] ** 2 + got["nz"] ** 2)
            assert np.abs(norm - 1.0).max() < 1e-12, r
            nt = torch.from_numpy(np.stack([got[f].transpose(2, 1, 0) for f in ("nx", "ny", "nz")])[None]).cuda()
            sg = torch.from_numpy(got["s"].transpose(2, 1, 0).astype(np.int8)[None]).cuda()
            Ek, *_ = energy_and_grad(nt, sg, reps[r]["KK"], reps[r]["g"], 1.0, want_grad=False)
            e_kernel = float(Ek[0]) / N
            assert abs(rows[0, r, 2] - e_kernel) <= 1e-12 * max(1.0, abs(e_kernel)), (r, rows[0, r, 2], e_kernel)


def _dropin(st, nstart, nstop, nout, g, KK, d, kbt):
    ni, nj, nk = st["nx"].shape
    nsub = ni * nj * nk // 2
    rows = (1 + nstop - nstart) // nout
    msteps = np.zeros(max(rows, 1), dtype=np.int32)
    m = np.zeros((max(rows, 1), 5), dtype=np.float64, order="F")
    sl1 = np.zeros(nsub, np.int32)
    sl2 = np.zeros(nsub, np.int32)
    dd = ct.c_double(d)
    _lib.lib().run(ct.c_int(nstart), ct.c_int(nstop), ct.c_int(nout), st["nx"], st["ny"], st["nz"], st["s"],
                   st["dope"], sl1, sl2, ct.c_double(g), ct.c_double(KK), dd, ct.c_double(kbt),
                   ct.c_int(ni), ct.c_int(nj), ct.c_int(nk), ct.c_int(0), msteps, m)
    return msteps[:rows], m[:rows], dd.value, _lib.last_error()


def test_dropin_run_two_threads():
    L = _lib.lib()
    cases = [((16, 12, 10), 31), ((8, 8, 8), 32)]
    st0 = [isotropic_state(*dims, sd) for dims, sd in cases]
    work = [copy_state(s) for s in st0]
    res = [None, None]
    _lib.check(L.chll_configure(0, 1, 0, 0))
    _lib.check(L.chll_set_seed(4242, 0))
    try:
        def call(i):
            res[i] = _dropin(work[i], 1, 40, 10, 0.5, 1.0, 0.3, 0.7)

        print("drop-in run() from two threads", flush=True)
        th = [threading.Thread(target=call, args=(i,)) for i in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    finally:
        _lib.check(L.chll_configure(0, 0, 0, 1))
    orc = Oracle()
    used = []
    for i in range(2):
        ms, m, d_out, err = res[i]
        assert err == "", err
        match = None
        for run_index in (0, 1):
            b = copy_state(st0[i])
            oms, om, od, _ = orc.run(b, 1, 40, 10, 0.5, 1.0, 0.3, 0.7, rng=RNG_PHILOX, seed=4242,
                                     run_index=run_index, arith=ARITH_FP64)
            if np.array_equal(b["s"], work[i]["s"]):
                match = run_index
                assert np.array_equal(ms, oms) and d_out == od
                assert np.array_equal(m[:, 1:], om[:, 1:])
                for f in ("nx", "ny", "nz"):
                    assert np.abs(b[f] - work[i][f]).max() <= 1e-12
                break
        assert match is not None, i
        used.append(match)
    assert sorted(used) == [0, 1]


def _kernel_energy_per_site(st, K, g):
    n = torch.from_numpy(np.stack([st[f].transpose(2, 1, 0) for f in ("nx", "ny", "nz")])[None]).cuda()
    sg = torch.from_numpy(st["s"].transpose(2, 1, 0).astype(np.int8)[None]).cuda()
    E, *_ = energy_and_grad(n, sg, K, g, 1.0, want_grad=False)
    return float(E[0]) / st["s"].size


@pytest.mark.timeout(900)
def test_l480_coarsening_5000_sweeps():
    L, K, g, kT, d0 = 480, 1.0, 0.25, 0.5, 0.3
    N = L ** 3
    rows = []
    with Lattice(L, L, L) as lat:
        print("L=480 coarsening: initial state on the device", flush=True)
        lat.init_state("isotropic", seed=4800, p_flip=0.5)
        lat.set_step(d0)
        for c in range(10):  # ten calls of 500 sweeps, a row every 50
            out = lat.sweeps(1 + 500 * c, 500 * (c + 1), 50, g=g, KK=K, kbt=kT, rng="philox", arith="fp64", seed=48)
            rows.append(out["rows"])
            print(f"L=480 coarsening: sweep {500 * (c + 1)}, E/N {out['rows'][-1][6]:.6f}", flush=True)
        st = lat.download()
    rows = np.concatenate(rows)
    assert rows.shape[0] == 100 and np.array_equal(rows[:, 0], 50.0 * np.arange(1, 101))
    assert np.all(np.isfinite(rows))
    e, eta, pacc, facc = rows[:, 6], rows[:, 7], rows[:, 3], rows[:, 4]
    # the cone step adapts each sweep (istep < 10000): acceptance in the band after the first rows
    assert np.all((pacc[4:] > 0.38) & (pacc[4:] < 0.62)), pacc
    assert np.all((facc >= 0.0) & (facc <= 1.0)) and np.all(np.abs(eta) <= 1.0)
    p2 = np.array([p2_from_q(r[8:14]) for r in rows])
    assert np.all((p2 >= 0.0) & (p2 <= 1.0))
    # coarsening from a random state at kT = 0.5 lowers the energy
    assert e[-10:].mean() < e[:10].mean() - 1e-3, (e[:10].mean(), e[-10:].mean())
    assert np.all(np.isin(st["s"], (-1, 1)))
    norm = np.sqrt(st["nx"] ** 2 + st["ny"] ** 2 + st["nz"] ** 2)
    assert np.abs(norm - 1.0).max() < 1e-12
    print("L=480 coarsening: energy kernel on the final state", flush=True)
    e_kernel = _kernel_energy_per_site(st, K, g)
    assert abs(e[-1] - e_kernel) <= 1e-12 * max(1.0, abs(e_kernel)), (e[-1], e_kernel)


@pytest.mark.timeout(900)
def test_quench_16_l128_full_length():
    B, L = 16, 128
    gen = torch.Generator().manual_seed(1616)
    n = torch.randn(B, 3, L, L, L, generator=gen, dtype=torch.float64)
    n /= torch.linalg.vector_norm(n, dim=1, keepdim=True)
    s = torch.where(torch.rand(B, L, L, L, generator=gen) < 0.5, -1, 1).to(torch.int8)
    n, s = n.cuda(), s.cuda()
    K = torch.linspace(0.5, 2.0, B, dtype=torch.float64, device="cuda")
    g = torch.linspace(0.0, 1.0, B, dtype=torch.float64, device="cuda")
    print("16 x 128^3: 500-step quench, then 2000 sweeps per lattice at kT = 0.2", flush=True)
    hist, rows, finals = quench_then_metropolis(n, s, K, g, steps=500, lr=0.05, sweeps=2000, kT=0.2, nout=100)
    e = hist.cpu().numpy()  # (steps, B)
    assert e.shape == (500, B) and np.all(np.isfinite(e))
    assert np.all(e[-1] < e[0])
    for b in range(B):
        r = rows[b]
        assert r.shape[0] == 20 and np.all(np.isfinite(r)), b
        st = finals[b]
        assert np.all(np.isin(st["s"], (-1, 1))), b
        norm = np.sqrt(st["nx"] ** 2 + st["ny"] ** 2 + st["nz"] ** 2)
        assert np.abs(norm - 1.0).max() < 1e-12, b
        e_kernel = _kernel_energy_per_site(st, float(K[b]), float(g[b]))
        assert abs(r[-1][6] - e_kernel) <= 1e-12 * max(1.0, abs(e_kernel)), (b, r[-1][6], e_kernel)
    print("16 x 128^3: done", flush=True)


@pytest.mark.timeout(900)
def test_l128_anneal_20000_sweeps():
    L, K, g, d0, nsw = 128, 1.0, 0.5, 0.3, 20000
    with Lattice(L, L, L) as lat:
        lat.upload_state(isotropic_state(L, L, L, 20261203))
        lat.set_step(d0)
        # one call: kT ramps linearly from 1.5 at sweep 1 to 0.3 at sweep 20000
        # (chll_params.kbt_end; test_anneal_ramp pins it against one call per
        # sweep and the oracle), a row every 100 sweeps
        print("L=128 anneal: 20000 sweeps in one call", flush=True)
        rows = lat.sweeps(1, nsw, 100, g=g, KK=K, kbt=1.5, kbt_end=0.3, rng="philox", arith="fp64", seed=128)["rows"]
        st = lat.download()
    rows = np.array(rows)
    assert rows.shape[0] == nsw // 100 and np.all(np.isfinite(rows))
    assert np.array_equal(rows[:, 0], 100.0 * np.arange(1, nsw // 100 + 1))
    e = rows[:, 6]
    assert e[-10:].mean() < e[:10].mean() - 0.1, (e[:10].mean(), e[-10:].mean())
    assert np.all(np.isin(st["s"], (-1, 1)))
    norm = np.sqrt(st["nx"] ** 2 + st["ny"] ** 2 + st["nz"] ** 2)
    assert np.abs(norm - 1.0).max() < 1e-12
    e_kernel = _kernel_energy_per_site(st, K, g)
    assert abs(e[-1] - e_kernel) <= 1e-12 * max(1.0, abs(e_kernel)), (e[-1], e_kernel)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("arith", ["fp64", "compat"])
def test_phase_scan_10000_sweeps(arith):
    """The scan workload's full 10k sweeps; in reference arithmetic its blocks
    run fused with held chirality trials (every replica takes site pairs)."""
    reps = phase_scan_replicas()
    n = len(reps)
    print(f"phase scan, 10k sweeps ({arith}): 1000 blocks of 10", flush=True)
    with Ensemble(reps, bc="open", arith=arith) as ens:
        # the configuration the bench's autotuner picks on MI355X: chunks of
        # 128 MiB taking turns on two graph branches
        ens.set_chunk(128 << 20)
        ens.set_branches(2)
        assert ens.chunks() > 2
        ens.init_state("isotropic", seed=10000, p_flip=0.5)
        rows = ens.run(1000, 10, seed=10000, run_index=2, adapt=True, use_graph=True)
        st = ens.state()
        assert st["istep"] == 10001
        assert rows.shape == (1000, n, rows.shape[2]) and np.all(np.isfinite(rows))
        ser = series_by_temperature(rows, reps)
        for sl in range(n):
            N = reps[sl]["ni"] * reps[sl]["nj"] * reps[sl]["nk"]
            sm = summarize(ser[sl], reps[sl]["kbt"], N)
            assert all(np.isfinite(v) for v in sm.values()), (sl, sm)
            # chi_c is 0 where a strong field has saturated every spin
            assert sm["Cv"] > 0 and sm["chi"] >= 0, (sl, sm)
        print("phase scan, 10k sweeps: final states against the energy kernel", flush=True)
        for r in range(n):
            got = ens.download(r)
            N = reps[r]["ni"] * reps[r]["nj"] * reps[r]["nk"]
            assert np.all(np.isin(got["s"], (-1, 1))), r
            norm = np.sqrt(got["nx"] ** 2 + got["ny"] ** 2 + got["nz"] ** 2)
            assert np.abs(norm - 1.0).max() < 1e-12, r
            e_kernel = _kernel_energy_per_site(got, reps[r]["KK"], reps[r]["g"])
            assert abs(rows[-1, r, 2] - e_kernel) <= 1e-12 * max(1.0, abs(e_kernel)), (r, rows[-1, r, 2], e_kernel)
