"""Pins of the analysis oracle (oracle/analysis.py) and the host-side pieces of
the analysis API, without a GPU.

* The vectorised correlation sums against the reference's six nested loops
  written out literally (notebooks/paper_plots2.ipynb, cell 0), on small random
  lattices, and against closed forms: unit directors that all agree give 1 in
  every shell, uniform spins give 0, the checkerboard (-1)^{i+j+k} gives
  sum_d (-1)^{di+dj+dk} c(d) / c(shell) - <s>^2, and the pair counts are
  sum_d prod_c (n_c - |d_c|) (shell 1, the 26 cube neighbours, factorises).
* The vectorised structure factor against the literal get_sc loop
  (notebooks/.ipynb_checkpoints/hero_stripe-checkpoint.py:10-17), the Dirichlet
  kernel of a uniform field along (1, 0, 0), S(0) = <s>^2, and the peak of a
  stripe pattern.
* The library's host-only pair counts (chll_corr_counts) against the oracle.
* The VTK writer: a read-back of both layouts, and the reference's array order
  (viz/np2vti.py:13-22) for layout="reference".
"""
import math

import numpy as np
import pytest

from oracle import analysis as ora


def rand_state(shape, seed):
    rng = np.random.default_rng(seed)
    n = rng.normal(size=(3,) + shape)
    n /= np.linalg.norm(n, axis=0)
    s = np.where(rng.random(shape) < 0.5, -1, 1).astype(np.int32)
    return n[0], n[1], n[2], s


@pytest.mark.parametrize("shape,ns", [((4, 5, 3), 3), ((3, 3, 6), 2), ((5, 2, 4), 4), ((2, 2, 2), 1)])
def test_vectorised_correlations_equal_loops(shape, ns):
    nx, ny, nz, s = rand_state(shape, sum(shape) + ns)
    gl, cl = ora.calc_n_corr_loops(nx, ny, nz, ns)
    gv, cv = ora.calc_n_corr(nx, ny, nz, ns)
    assert np.array_equal(cl, cv)
    assert np.allclose(gv, gl, rtol=1e-13, atol=0)
    sl, csl = ora.calc_s_corr_loops(s, ns)
    sv, csv = ora.calc_s_corr(s, ns)
    assert np.array_equal(csl, csv)
    assert np.allclose(sv, sl, rtol=0, atol=1e-13)
    assert np.array_equal(ora.corr_counts(shape, ns).astype(float), cv)


def test_correlation_closed_forms():
    shape, ns = (6, 5, 7), 4
    u = np.array([0.6, 0.0, 0.8])
    g, _ = ora.calc_n_corr(*(np.full(shape, x) for x in u), ns)
    assert np.allclose(g, 1.0, rtol=0, atol=1e-15)
    g, _ = ora.calc_s_corr(np.ones(shape, dtype=np.int32), ns)
    assert np.array_equal(g, np.zeros(ns))
    i, j, k = np.indices(shape)
    cb = np.where((i + j + k) % 2 == 0, 1, -1).astype(np.int32)
    g, c = ora.calc_s_corr(cb, ns)
    want = np.zeros(ns)
    for di, dj, dk, sh in ora._offsets(ns):
        a = [n - abs(x) for n, x in zip(shape, (di, dj, dk))]
        if min(a) > 0:
            want[sh] += (-1) ** ((di + dj + dk) % 2) * a[0] * a[1] * a[2]
    assert np.allclose(g, want / c - np.mean(cb) ** 2, rtol=0, atol=1e-15)
    # shell 0 is d = 0 alone; shell 1 (1 <= |d|^2 < 4) the 26 neighbours of the
    # 3x3x3 cube, whose counts factorise: prod_c (n_c + 2 (n_c - 1)) - N
    N = np.prod(shape)
    assert c[0] == N
    assert c[1] == np.prod([3 * n - 2 for n in shape]) - N


def test_library_counts_match_oracle():
    from chllmodel_amd.analysis import corr_counts

    for shape, ns in (((6, 5, 7), 4), ((1, 9, 3), 6), ((33, 17, 9), 12)):
        assert np.array_equal(corr_counts(shape, ns), ora.corr_counts(shape, ns))


def test_structure_factor_oracle():
    rng = np.random.default_rng(4)
    s = np.where(rng.random((3, 4, 5)) < 0.5, -1, 1).astype(np.int32)
    for plane in ([1, 0, 0], [1, -1, 1], [0, 0, 1]):
        for kv in (0.0, 0.13, 0.5, 2.9):
            kvec = kv * np.array(plane, dtype=float)
            a, b = ora.get_sc_loops(s, kvec), ora.get_sc(s, kvec)
            assert abs(a - b) <= 1e-12 * max(1.0, abs(a))
    # S(0) = <s>^2 for every plane
    sf, _ = ora.structure_factor(s, 3, 0.0, 0.5, [1, 1, -1])
    assert sf[0] == pytest.approx(np.mean(s) ** 2, abs=1e-15)
    # uniform field along (1, 0, 0): the Dirichlet kernel
    ni = 12
    sf, kvs = ora.structure_factor(np.ones((ni, 3, 4), dtype=np.int32), 25, 0.01, 0.5, [1, 0, 0])
    want = np.sin(kvs * ni / 2) ** 2 / (ni ** 2 * np.sin(kvs / 2) ** 2)
    assert np.allclose(sf, want, rtol=1e-12, atol=1e-15)
    # stripes of period 8 along j: the peak sits at 2 pi / 8
    j = np.arange(1, 33)
    st = np.broadcast_to(np.where(np.cos(2 * math.pi * j / 8 + 0.3) >= 0, 1, -1)[None, :, None], (4, 32, 4))
    sf, kvs = ora.structure_factor(np.ascontiguousarray(st), 200, 0.0, 1.5, [0, 1, 0])
    assert abs(kvs[np.argmax(sf)] - 2 * math.pi / 8) < 1.5 / 199


@pytest.mark.parametrize("layout", ["reference", "points"])
def test_vti_roundtrip(tmp_path, layout):
    from chllmodel_amd.analysis import read_vti, write_vti

    shape = (4, 3, 5)
    nx, ny, nz, s = rand_state(shape, 9)
    p = str(tmp_path / "x.vti")
    write_vti(p, nx, ny, nz, s, layout=layout)
    r = read_vti(p)
    assert r["dims"] == shape
    if layout == "reference":
        assert np.array_equal(r["s"], s.flatten())
        assert np.array_equal(r["n"], np.stack((nx, ny, nz)).reshape(-1, 3).astype(np.float32))
    else:
        # VTK point id x + y ni + z ni nj
        pid = lambda i, j, k: i + shape[0] * (j + shape[1] * k)  # noqa: E731
        for (i, j, k) in ((0, 0, 0), (3, 1, 2), (1, 2, 4)):
            assert r["s"][pid(i, j, k)] == s[i, j, k]
            assert np.array_equal(r["n"][pid(i, j, k)], np.float32([nx[i, j, k], ny[i, j, k], nz[i, j, k]]))
    head = open(p, "rb").read(600).decode(errors="replace")
    assert 'WholeExtent="0 3 0 2 0 4"' in head and 'Name="n" NumberOfComponents="3"' in head


def test_convert_npz_tree(tmp_path):
    from chllmodel_amd.analysis import main, read_vti

    nx, ny, nz, s = rand_state((3, 4, 2), 1)
    d = tmp_path / "data" / "run1"
    d.mkdir(parents=True)
    np.savez(d / "run1_data.npz", nx=nx, ny=ny, nz=nz, s=s, dope=np.zeros_like(s))
    assert main([str(tmp_path / "data")]) == 0
    r = read_vti(str(d / "run1_data.vti"))
    assert r["dims"] == (3, 4, 2) and np.array_equal(r["s"], s.flatten())
    assert main([str(tmp_path / "nothing.txt")]) == 1
