"""Pins of the CPU oracle (oracle/chll_oracle.c) before it is trusted as the
checker of the HIP library.

* Philox4x32-10 against the published known-answer vectors.
* Site energy (chll/chll_mod.f90:363-412) against closed forms and exact
  symmetries, each chosen so that a dropped term, a flipped sign or a swapped
  cross-product component fails.
* The whole sweep (chll/chll_mod.f90:40-245) and etot (:414-456) against the
  golden vectors written by the reference's own compiled source
  (tests/golden/make_golden.py), bit for bit, and against that build live when
  it is present.
* Statistical laws that fix the sweep's acceptance rule: with d = 0 the
  directors are frozen and every spin is independent in a local field, so
  <s_a> = tanh((K sum_b sgn_b c_b q_b + 2 g) / (2 kT)) exactly.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle.oracle import (ARITH_COMPAT, ARITH_FP64, BC_OPEN, BC_PERIODIC, RNG_BUFFER, RNG_PHILOX,
                           RNG_STREAM, REF_STREAM_SO, Oracle, Reference, copy_state, isotropic_state)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


# --------------------------------------------------------------------------
# Philox4x32-10 known-answer vectors (Random123 kat_vectors, philox4x32_10).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_kat(orc, ctr, key, want):
    assert tuple(int(x) for x in orc.philox(ctr, key)) == want


def test_stream_uniform_range_and_moments(orc):
    u = orc.stream(7, 0, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5 * math.sqrt(1 / 12 / u.size)
    # consecutive draws share a Philox block but are distinct words
    assert np.all(u[0::2] != u[1::2])
    # the stream is a pure function of (seed, index)
    assert np.array_equal(orc.stream(7, 1000, 50), u[1000:1050])


# --------------------------------------------------------------------------
def _state(ni, nj, nk, seed=1):
    return isotropic_state(ni, nj, nk, seed)


def test_energy_uniform_alignment(orc):
    """All directors equal: every bond has c = 1, q = 0, so E_site = -g*sigma."""
    st = _state(5, 4, 3)
    v = np.array([0.3, -0.5, 0.8])
    v /= np.linalg.norm(v)
    for a, c in zip(("nx", "ny", "nz"), v):
        st[a][...] = c
    for arith in (ARITH_COMPAT, ARITH_FP64):
        for (i, j, k) in [(0, 0, 0), (2, 1, 1), (4, 3, 2)]:
            for s in (1, -1):
                e = orc.site_energy(st, i, j, k, v, s, g=0.37, KK=1.7, arith=arith)
                # 1 - c*c with c = v.v rounds to within a few ulp of 0 per bond
                assert abs(e - (-0.37 * s)) < 1e-6


@pytest.mark.parametrize("K", [0.0, 0.5, 1.0, 2.5, -1.3])
def test_energy_two_site_twist(orc, K):
    """2x1x1 lattice, directors twisted by D about the bond axis x:
    e = sin^2 D - K cos D sin D f - g s for both sites (f = 1 for s = +1)."""
    st = _state(2, 1, 1)
    st["s"][...] = 1
    for D in np.linspace(-1.4, 1.4, 15):
        ta = 0.3
        tb = ta + D
        st["nx"][0, 0, 0], st["ny"][0, 0, 0], st["nz"][0, 0, 0] = 0.0, math.cos(ta), math.sin(ta)
        st["nx"][1, 0, 0], st["ny"][1, 0, 0], st["nz"][1, 0, 0] = 0.0, math.cos(tb), math.sin(tb)
        want = math.sin(D) ** 2 - K * math.cos(D) * math.sin(D) - 0.2
        for site in (0, 1):
            n = (st["nx"][site, 0, 0], st["ny"][site, 0, 0], st["nz"][site, 0, 0])
            e64 = orc.site_energy(st, site, 0, 0, n, 1, g=0.2, KK=K, arith=ARITH_FP64)
            e32 = orc.site_energy(st, site, 0, 0, n, 1, g=0.2, KK=K, arith=ARITH_COMPAT)
            assert abs(e64 - want) < 1e-13
            assert abs(e32 - want) < 3e-7
    # minimum of the pair energy at tan(2D) = K, value (1 - sqrt(1+K^2))/2
    D = 0.5 * math.atan(K)
    emin = math.sin(D) ** 2 - K * math.cos(D) * math.sin(D)
    assert abs(emin - (1 - math.sqrt(1 + K * K)) / 2) < 1e-14


def test_energy_symmetries_bit_exact(orc):
    st = _state(6, 5, 4, seed=3)
    rs = np.random.RandomState(5)
    for _ in range(30):
        i, j, k = rs.randint(6), rs.randint(5), rs.randint(4)
        n = (st["nx"][i, j, k], st["ny"][i, j, k], st["nz"][i, j, k])
        s = int(st["s"][i, j, k])
        for arith in (ARITH_COMPAT, ARITH_FP64):
            for bc in (BC_OPEN,):
                e = orc.site_energy(st, i, j, k, n, s, 0.4, 1.3, bc=bc, arith=arith)
                # n -> -n at the site: c, q change sign together
                e_neg = orc.site_energy(st, i, j, k, tuple(-x for x in n), s, 0.4, 1.3, bc=bc, arith=arith)
                assert e_neg == e
                # all spins, K and g negated together
                st2 = copy_state(st)
                st2["s"] = -st2["s"]
                e_flip = orc.site_energy(st2, i, j, k, n, -s, -0.4, -1.3, bc=bc, arith=arith)
                assert e_flip == e


def test_energy_open_boundary_locality(orc):
    """Open boundaries: a corner site sees exactly its three neighbours, so its
    energy equals that of the same corner in a 2x2x2 copy; periodic differs."""
    st = _state(6, 6, 6, seed=4)
    sub = {k: np.asfortranarray(v[:2, :2, :2].copy()) for k, v in st.items()}
    n = (st["nx"][0, 0, 0], st["ny"][0, 0, 0], st["nz"][0, 0, 0])
    for arith in (ARITH_COMPAT, ARITH_FP64):
        e_big = orc.site_energy(st, 0, 0, 0, n, 1, 0.3, 1.1, arith=arith)
        e_small = orc.site_energy(sub, 0, 0, 0, n, 1, 0.3, 1.1, arith=arith)
        assert e_big == e_small
        e_per = orc.site_energy(st, 0, 0, 0, n, 1, 0.3, 1.1, bc=BC_PERIODIC, arith=arith)
        assert e_per != e_big


@pytest.mark.parametrize("bc", [BC_OPEN, BC_PERIODIC])
def test_site_energies_sum_to_hamiltonian(orc, bc):
    """Sum over sites of E_site counts every bond twice and the field once:
    (sum_a E_a + g sum s) / 2 = E_H + g sum s, with E_H = oracle_energy_true."""
    st = _state(6, 4, 4, seed=6)
    g, K = 0.45, 1.6
    tot = 0.0
    for k in range(4):
        for j in range(4):
            for i in range(6):
                n = (st["nx"][i, j, k], st["ny"][i, j, k], st["nz"][i, j, k])
                tot += orc.site_energy(st, i, j, k, n, int(st["s"][i, j, k]), g, K, bc=bc, arith=ARITH_FP64)
    S = float(st["s"].sum())
    eh = orc.energy_true(st, g, K, bc=bc)
    assert abs((tot + g * S) / 2 - (eh + g * S)) < 1e-10


def test_trial_director_is_cone(orc):
    rs = np.random.RandomState(9)
    for _ in range(200):
        v = rs.normal(size=3)
        v /= np.linalg.norm(v)
        if abs(v[2]) > 0.999:
            continue
        d = rs.uniform(0.01, 1.0)
        t = orc.trial_director(v, d, rs.random_sample())
        assert abs(np.linalg.norm(t) - 1) < 1e-15
        assert abs(float(np.dot(t, v)) - 1 / math.sqrt(1 + d * d)) < 1e-14
    # d = 0 returns the director itself
    v = np.array([0.6, 0.0, 0.8])
    assert np.allclose(orc.trial_director(v, 0.0, 0.3), v, atol=1e-16)


def test_init_lists(orc):
    for dims in [(2, 2, 2), (4, 3, 5), (6, 1, 1), (3, 2, 2)]:
        sl1, sl2 = orc.init(*dims)
        ni, nj, nk = dims
        idx = np.arange(1, ni * nj * nk + 1)
        i = (idx - 1) % ni + 1
        j = ((idx - 1) // ni) % nj + 1
        k = (idx - 1) // (ni * nj) + 1
        odd = (i + j + k) % 2 == 1
        assert np.array_equal(sl1[: odd.sum()], idx[odd])
        assert np.array_equal(sl2[: (~odd).sum()], idx[~odd])


def test_rank_is_half_the_linear_index(orc):
    """The uniform index of a site is its position in its sublattice list
    (chll/chll_mod.f90:131-134, 146): the inverse of the lists equals the
    closed form l >> 1 for every shape with N even, odd ni and nj included,
    exhaustively on small shapes."""
    for ni in range(1, 8):
        for nj in range(1, 6):
            for nk in range(1, 5):
                if (ni * nj * nk) % 2:
                    continue
                sl1, sl2 = orc.init(ni, nj, nk)
                half = ni * nj * nk // 2
                for lst in (sl1[:half], sl2[:half]):
                    l = lst.astype(np.int64) - 1
                    i, j, k = l % ni, (l // ni) % nj, l // (ni * nj)
                    want = np.arange(half)
                    assert np.array_equal(l >> 1, want), (ni, nj, nk)
                    for q in range(0, half, max(1, half // 7)):
                        assert orc.rank(ni, nj, int(i[q]), int(j[q]), int(k[q])) == want[q]


def test_init_state_generators(orc):
    """Initial states keyed by (seed, replica, site): the aligned states
    exactly, unit directors, spin fractions p_flip = 0 / 1 exactly, and for
    isotropic directors the moments of the uniform distribution on the
    sphere (<n> = 0, <n_a n_b> = delta_ab / 3, so P2 -> 0) and <s> = 1 - 2 p
    within 5 standard errors; the same key gives the same state, another
    replica another."""
    from chllmodel_amd.observables import p2_from_q

    ni, nj, nk = 64, 48, 40
    N = ni * nj * nk
    ax = orc.init_state(4, 3, 2, kind=1, seed=5, p_flip=0.0)
    assert np.all(ax["nx"] == 1.0) and np.all(ax["ny"] == 0.0) and np.all(ax["nz"] == 0.0)
    assert np.all(ax["s"] == 1)
    axz = orc.init_state(4, 3, 2, kind=2, seed=5, p_flip=1.0)
    assert np.all(axz["nx"] == 2 ** -0.5) and np.all(axz["ny"] == 0.0) and np.all(axz["nz"] == 2 ** -0.5)
    assert np.all(axz["s"] == -1)
    st = orc.init_state(ni, nj, nk, kind=0, seed=20261203, replica=3, p_flip=0.3)
    n = np.stack([st["nx"].ravel(), st["ny"].ravel(), st["nz"].ravel()])
    assert np.abs(np.sqrt((n * n).sum(0)) - 1.0).max() < 1e-15
    se1, se2 = np.sqrt(1 / (3 * N)), np.sqrt(4 / (45 * N))  # sd of n_a, of n_a^2, over sqrt(N)
    for a in range(3):
        assert abs(n[a].mean()) < 5 * se1
        assert abs((n[a] ** 2).mean() - 1 / 3) < 5 * se2
        assert abs((n[a] * n[(a + 1) % 3]).mean()) < 5 * np.sqrt(1 / (15 * N))
    q = [(n[0] ** 2).mean(), (n[1] ** 2).mean(), (n[2] ** 2).mean(), (n[0] * n[1]).mean(), (n[0] * n[2]).mean(),
         (n[1] * n[2]).mean()]
    assert p2_from_q(q) < 0.01
    sbar = st["s"].mean()
    assert abs(sbar - (1 - 2 * 0.3)) < 5 * 2 * np.sqrt(0.3 * 0.7 / N)
    again = orc.init_state(ni, nj, nk, kind=0, seed=20261203, replica=3, p_flip=0.3)
    other = orc.init_state(ni, nj, nk, kind=0, seed=20261203, replica=4, p_flip=0.3)
    assert all(np.array_equal(st[f], again[f]) for f in ("nx", "ny", "nz", "s"))
    assert not np.array_equal(st["nz"], other["nz"]) and not np.array_equal(st["s"], other["s"])
    # the draw of site l: Philox counter (l, replica, "INIT", 0) gives costh, phi
    w = orc.philox([7, 3, 0x494E4954, 0], [20261203 & 0xFFFFFFFF, 20261203 >> 32])
    u1 = ((int(w[0]) << 32 | int(w[1])) >> 11) * 2.0 ** -53
    assert st["nz"].ravel(order="F")[7] == 2.0 * (u1 - 0.5)


# --------------------------------------------------------------------------
GOLDEN_FILES = sorted(glob.glob(os.path.join(GOLDEN, "ref_*.npz")))


def load_golden(path):
    z = np.load(path, allow_pickle=False)
    st0 = {k: np.asfortranarray(z[k + "0"]) for k in ("nx", "ny", "nz", "s")}
    st0["dope"] = np.zeros_like(st0["s"])
    p = z["params"]
    nstart, nstop, nout = int(p[0]), int(p[1]), int(p[2])
    args = dict(g=p[3], KK=p[4], d=p[5], kbt=p[6], seed=int(p[7]))
    return z, st0, (nstart, nstop, nout), args


@pytest.mark.parametrize("path", GOLDEN_FILES, ids=[os.path.basename(p)[:-4] for p in GOLDEN_FILES])
def test_oracle_matches_reference_golden(orc, path):
    z, st0, (nstart, nstop, nout), a = load_golden(path)
    st = copy_state(st0)
    if str(z["mode"]) == "buffer":
        ms, m, d_out, pos = orc.run(st, nstart, nstop, nout, a["g"], a["KK"], a["d"], a["kbt"],
                                    rng=RNG_BUFFER, buf=z["buffer"])
    else:
        ms, m, d_out, pos = orc.run(st, nstart, nstop, nout, a["g"], a["KK"], a["d"], a["kbt"],
                                    rng=RNG_STREAM, seed=a["seed"])
    assert np.array_equal(ms, z["msteps"])
    assert np.array_equal(m, z["m"])
    assert d_out == float(z["d_out"])
    assert pos == int(z["stream_pos_out"])
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(st[f], z[f]), f


@pytest.mark.skipif(not os.path.exists(REF_STREAM_SO), reason="reference build absent (needs the reference source)")
def test_oracle_matches_reference_live(orc):
    ref = Reference(stream=True)
    for dims, seed in [((4, 6, 4), 31), ((12, 10, 8), 32)]:
        st0 = isotropic_state(*dims, seed)
        a, b = copy_state(st0), copy_state(st0)
        r1 = orc.run(a, 1, 15, 3, 0.7, 1.2, 0.25, 0.6, seed=seed)
        r2 = ref.run(b, 1, 15, 3, 0.7, 1.2, 0.25, 0.6, seed=seed)
        assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
        assert r1[2] == r2[2] and r1[3] == r2[3]
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(a[f], b[f])


@pytest.mark.skipif(not os.path.exists(REF_STREAM_SO), reason="reference build absent (needs the reference source)")
@pytest.mark.parametrize("dims,nstart,nstop,nout", [((6, 4, 5), 1, 7, 20), ((6, 4, 5), 5, 2, 4), ((2, 1, 1), 1, 6, 2),
                                                    ((2, 1, 3), 1, 9, 3), ((4, 3, 1), 3, 11, 4)])
def test_oracle_matches_reference_live_edge_cases(orc, dims, nstart, nstop, nout):
    """Calls at the edges of the reference's interface: nout longer than the
    run (sweeps, no rows), nstop < nstart (no sweep), one site per colour
    (2x1x1) and flat lattices, bit for bit against the reference build."""
    ref = Reference(stream=True)
    st0 = isotropic_state(*dims, 77 + sum(dims))
    a, b = copy_state(st0), copy_state(st0)
    r1 = orc.run(a, nstart, nstop, nout, 0.5, 1.0, 0.3, 0.7, seed=91)
    r2 = ref.run(b, nstart, nstop, nout, 0.5, 1.0, 0.3, 0.7, seed=91)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
    assert r1[2] == r2[2] and r1[3] == r2[3]
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(a[f], b[f])
    if nstop < nstart:
        assert r1[3] == 0 and r1[2] == 0.3


# --------------------------------------------------------------------------
def test_zero_step_accepts_everything(orc):
    st = isotropic_state(4, 4, 4, 8)
    st0 = copy_state(st)
    ms, m, d_out, _ = orc.run(st, 1, 6, 1, 0.5, 1.0, 0.0, 0.7, seed=3)
    assert np.all(m[:, 2] == 1.0)
    assert d_out == 0.0
    for f in ("nx", "ny", "nz"):
        assert np.abs(st[f] - st0[f]).max() < 1e-15


def _local_field_mean_spin(st, g, K, kT):
    """<s_a> = tanh((K sum_b sgn_b c_b q_b + 2 g)/(2 kT)) under frozen directors
    (open boundaries); from the Hamiltonian, independently of the sweep code."""
    nx, ny, nz = st["nx"], st["ny"], st["nz"]
    ni, nj, nk = nx.shape
    out = np.zeros(nx.shape)
    for k in range(nk):
        for j in range(nj):
            for i in range(ni):
                a = np.array([nx[i, j, k], ny[i, j, k], nz[i, j, k]])
                h = 0.0
                for axis in range(3):
                    for sgn in (+1, -1):
                        q = [i, j, k]
                        q[axis] += sgn
                        if not (0 <= q[0] < ni and 0 <= q[1] < nj and 0 <= q[2] < nk):
                            continue
                        b = np.array([nx[tuple(q)], ny[tuple(q)], nz[tuple(q)]])
                        c = float(a @ b)
                        cr = np.cross(a, b)[axis]
                        h += sgn * c * cr
                out[i, j, k] = math.tanh((K * h + 2 * g) / (2 * kT))
    return out


def test_frozen_directors_local_field_law(orc):
    st = isotropic_state(4, 4, 2, 21)
    g, K, kT = 0.3, 1.5, 0.9
    want = _local_field_mean_spin(st, g, K, kT)
    nsw = 20000
    acc = np.zeros(st["s"].shape)
    cur = copy_state(st)
    for blk in range(20):
        orc.run(cur, 1 + blk * (nsw // 20), (blk + 1) * (nsw // 20), nsw, g, K, 0.0, kT,
                rng=RNG_PHILOX, seed=77, run_index=0)
        acc += cur["s"]
    # sampled once per block of 1000 sweeps: effectively independent draws
    mean = acc / 20
    # tolerance from the binomial variance of 20 samples per site
    sig = np.sqrt((1 - want ** 2) / 20)
    assert np.all(np.abs(mean - want) <= 4.5 * sig + 1e-9)


def test_k0_spins_follow_tanh(orc):
    """K = 0: spins decouple from directors, <s> = tanh(g/kT) (the reference's own
    check of eta, notebooks/paper_plots2.ipynb)."""
    st = isotropic_state(8, 8, 8, 22)
    g, kT = 0.4, 0.8
    ms, m, _, _ = orc.run(st, 1, 600, 1, g, 0.0, 0.3, kT, rng=RNG_PHILOX, seed=5)
    eta = m[100:, 1]
    want = math.tanh(g / kT)
    # per-row std of eta = sqrt((1-want^2)/N); rows are nearly independent at K=0
    sig = math.sqrt((1 - want ** 2) / 512) / math.sqrt(len(eta))
    assert abs(eta.mean() - want) < 5 * sig


@pytest.mark.parametrize("arith", [ARITH_FP64, ARITH_COMPAT])
def test_slab_with_site_offset_matches_whole_lattice(orc, arith):
    """A slab of planes k0.. run with site_offset = ni nj k0 / 2 draws the same
    Philox uniforms as those planes of the whole lattice.  Its open bottom face
    misleads the planes next to it, by at most four planes per sweep (D0, D1,
    C0 and C1 each reach one plane further), so planes k0 + 4 t and above agree
    bit for bit after t sweeps.  Sweeps at istep >= 10000 keep
    the cone step fixed (the adaptation, chll/chll_mod.f90:83-90, would see the
    slab's own accept counts)."""
    ni, nj, nk, k0 = 8, 6, 20, 8
    whole = isotropic_state(ni, nj, nk, 31)
    slab = {f: np.array(whole[f][:, :, k0:], order="F") for f in whole}
    for st, off in ((whole, 0), (slab, ni * nj * k0 // 2)):
        orc.run(st, 20001, 20002, 1, 0.5, 1.0, 0.3, 0.7, rng=RNG_PHILOX, seed=9, run_index=2,
                arith=arith, site_offset=off)
    keep = 8
    for f in ("nx", "ny", "nz", "s"):
        assert np.array_equal(slab[f][:, :, keep:], whole[f][:, :, k0 + keep:]), f
    # the bottom planes do differ, so the comparison above is not vacuous
    assert not np.array_equal(slab["nx"][:, :, :2], whole["nx"][:, :, k0:k0 + 2])
    # and without the offset the draws are another site's
    other = {f: np.array(whole[f][:, :, k0:], order="F") for f in whole}
    orc.run(other, 20003, 20003, 1, 0.5, 1.0, 0.3, 0.7, rng=RNG_PHILOX, seed=9, run_index=2, arith=arith)
    ref = {f: np.array(whole[f][:, :, k0:], order="F") for f in whole}
    orc.run(ref, 20003, 20003, 1, 0.5, 1.0, 0.3, 0.7, rng=RNG_PHILOX, seed=9, run_index=2, arith=arith,
            site_offset=ni * nj * k0 // 2)
    assert not np.array_equal(other["nx"], ref["nx"])


@pytest.mark.parametrize("arith", [ARITH_COMPAT, ARITH_FP64])
@pytest.mark.parametrize("rng,bc", [(RNG_STREAM, BC_OPEN), (RNG_PHILOX, BC_OPEN), (RNG_PHILOX, BC_PERIODIC)])
def test_threaded_and_hostfilled_runs_equal_the_serial_run(orc, arith, rng, bc):
    """The threaded restatement (the CPU baseline and the checker of the
    longer GPU runs) and the host-filled stream (every sweep's 10 nsub
    uniforms drawn first, as the reference's random_number calls do,
    chll/chll_mod.f90:66-80) leave the lattice, rows, counts and stream
    position of the serial run bit for bit: sites of one colour are
    independent within a half-sweep (chll/chll_mod.f90:129, 213)."""
    st0 = isotropic_state(12, 10, 8, 77)
    kw = dict(rng=rng, seed=99, bc=bc, arith=arith, want_counts=True)
    ref = copy_state(st0)
    want = orc.run(ref, 1, 30, 5, 0.5, 1.0, 0.3, 0.7, **kw)
    variants = [dict(threads=4), dict(threads=1, hostfill=True), dict(threads=3, hostfill=True)]
    for v in variants:
        got_st = copy_state(st0)
        got = orc.run(got_st, 1, 30, 5, 0.5, 1.0, 0.3, 0.7, **kw, **v)
        for f in ("nx", "ny", "nz", "s"):
            assert np.array_equal(got_st[f], ref[f]), (v, f)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), v
        assert got[2] == want[2] and got[3] == want[3], v
        assert np.array_equal(got[4], want[4]), v
