"""Analysis kernels on the GPU (chll_shell_corr, chll_structure_factor) against
the oracle (oracle/analysis.py, pinned by tests/test_analysis_oracle.py).

Tolerances: the chiral shell sums are exact integers on both sides and the
normalisation is the same two roundings, so the chiral correlation must agree
to 1e-15; the nematic sums are float64 sums in a different order (relative
1e-12); the structure factor sums integer histograms times float64 phases kv t
that the oracle forms as kv p . (r + 1) (a few ulp of the phase apart), so
|dS| <= 1e-10 S + 1e-15.
"""
import numpy as np
import pytest
import torch

from chllmodel_amd.analysis import (calc_n_corr, calc_s_corr, shell_correlations, structure_factor,
                                    structure_factors)
from oracle import analysis as ora

pytestmark = pytest.mark.gpu

PLANES = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1]]


def rand_state(shape, seed):
    rng = np.random.default_rng(seed)
    n = rng.normal(size=(3,) + shape)
    n /= np.linalg.norm(n, axis=0)
    s = np.where(rng.random(shape) < 0.5, -1, 1).astype(np.int32)
    return n[0], n[1], n[2], s


@pytest.mark.parametrize("shape,ns", [((7, 5, 6), 3), ((16, 16, 16), 5), ((1, 9, 5), 4), ((33, 4, 3), 2),
                                      ((5, 5, 5), 1), ((12, 10, 11), 7)])
def test_correlations_match_oracle(shape, ns):
    print(f"shell correlations {shape} ns={ns}", flush=True)
    nx, ny, nz, s = rand_state(shape, sum(shape) * 7 + ns)
    gn, gs, c = shell_correlations(nx, ny, nz, s, nshells=ns)
    rn, rc = ora.calc_n_corr(nx, ny, nz, ns)
    rs, _ = ora.calc_s_corr(s, ns)
    assert np.array_equal(c, rc)
    assert np.allclose(gn, rn, rtol=1e-12, atol=0)
    assert np.allclose(gs, rs, rtol=0, atol=1e-15)
    # the reference-named entry points return the same (g, c)
    g1, c1 = calc_n_corr(nx, ny, nz, ns)
    g2, c2 = calc_s_corr(s, ns)
    assert np.array_equal(g1, gn) and np.array_equal(g2, gs) and np.array_equal(c1, c) and np.array_equal(c2, c)


def test_correlations_large_and_invariant():
    """64^3, 5 shells: the oracle, and exact invariances (n -> -n on any subset
    of sites, s -> -s everywhere) that hold bit for bit."""
    shape, ns = (64, 64, 64), 5
    print(f"shell correlations {shape} ns={ns}", flush=True)
    nx, ny, nz, s = rand_state(shape, 5)
    # a correlated state: smooth the spins along i so the shells differ
    s = np.where(np.cumsum(s, axis=0) % 3 == 0, 1, -1).astype(np.int32)
    gn, gs, _ = shell_correlations(nx, ny, nz, s, nshells=ns)
    rn, _ = ora.calc_n_corr(nx, ny, nz, ns)
    rs, _ = ora.calc_s_corr(s, ns)
    assert np.allclose(gn, rn, rtol=1e-12, atol=0)
    assert np.allclose(gs, rs, rtol=0, atol=1e-15)
    flip = np.where(np.random.default_rng(2).random(shape) < 0.5, -1.0, 1.0)
    gn2, gs2, _ = shell_correlations(nx * flip, ny * flip, nz * flip, -s, nshells=ns)
    assert np.array_equal(gn2, gn) and np.array_equal(gs2, gs)
    # torch tensors on the device give the same bits as numpy input
    t = [torch.from_numpy(a).cuda() for a in (nx, ny, nz, s)]
    gn3, gs3, _ = shell_correlations(*t, nshells=ns)
    assert np.array_equal(gn3, gn) and np.array_equal(gs3, gs)


def test_correlation_errors():
    nx, ny, nz, s = rand_state((1, 1, 2), 0)
    with pytest.raises(RuntimeError, match="no pairs"):
        shell_correlations(nx, ny, nz, s, nshells=3)
    with pytest.raises(ValueError):
        shell_correlations(nx, ny, nz, s, nshells=0)
    with pytest.raises(ValueError):
        shell_correlations(s=np.full((2, 2, 2), 300), nshells=1)


@pytest.mark.parametrize("shape", [(9, 6, 7), (1, 5, 8), (20, 3, 2)])
def test_structure_factor_matches_oracle(shape):
    print(f"structure factor {shape}", flush=True)
    _, _, _, s = rand_state(shape, sum(shape))
    kvs = np.concatenate([np.linspace(0.0, 0.5, 40), [1.3, 2.9, 6.0]])
    sf = structure_factors(s, kvs, PLANES)
    for p, plane in enumerate(PLANES):
        ref = np.array([abs(ora.get_sc(s, kv * np.array(plane, dtype=float)) / s.size) ** 2 for kv in kvs])
        assert np.all(np.abs(sf[p] - ref) <= 1e-10 * ref + 1e-15), plane
    got, kv = structure_factor(s, 25, 0.0, 0.5, [1, -1, 1])
    assert np.array_equal(kv, np.linspace(0.0, 0.5, 25))
    ref, _ = ora.structure_factor(s, 25, 0.0, 0.5, [1, -1, 1])
    assert np.all(np.abs(got - ref) <= 1e-10 * ref + 1e-15)


def test_structure_factor_large():
    """128^3 stripes along (1, 1, 0)-ish plus noise: the oracle at a few kv,
    the Dirichlet kernel of a uniform field, and s -> -s invariance."""
    L = 128
    print(f"structure factor {L}^3", flush=True)
    i, j, k = np.indices((L, L, L))
    rng = np.random.default_rng(3)
    s = np.where(np.cos(2 * np.pi * (i + j) / 16.0) + 0.8 * rng.standard_normal((L, L, L)) >= 0, 1, -1)
    s = s.astype(np.int32)
    kvs = np.array([0.0, 0.1, 2 * np.pi / 16, 0.45])
    sf = structure_factors(s, kvs, PLANES)
    for p, plane in enumerate(PLANES):
        ref = np.array([abs(ora.get_sc(s, kv * np.array(plane, dtype=float)) / s.size) ** 2 for kv in kvs])
        assert np.all(np.abs(sf[p] - ref) <= 1e-10 * ref + 1e-15), plane
    sfm = structure_factors(-s, kvs, PLANES)
    assert np.array_equal(sfm, sf)
    u = np.ones((L, 8, 8), dtype=np.int32)
    kv = np.linspace(0.01, 0.5, 30)
    got = structure_factors(u, kv, [[1, 0, 0]])[0]
    want = np.sin(kv * L / 2) ** 2 / (L ** 2 * np.sin(kv / 2) ** 2)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-15)


def test_structure_factor_plane_groups_and_limits():
    _, _, _, s = rand_state((8, 8, 8), 11)
    rng = np.random.default_rng(5)
    planes = rng.integers(-40, 41, size=(17, 3))
    kvs = np.linspace(0.0, 0.2, 7)
    together = structure_factors(s, kvs, planes)  # more entries than one call holds: split
    for p, plane in enumerate(planes):
        alone = structure_factors(s, kvs, [plane])[0]
        assert np.array_equal(together[p], alone)
    with pytest.raises(RuntimeError, match="histogram entries"):
        structure_factors(np.ones((200, 2, 2), dtype=np.int32), kvs, [[100, 0, 0]])


def test_structure_factor_general_planes():
    """Planes whose rows' level sets run along general directions (|a|, |b| > 1),
    p0 = 0 with both p1, p2 set, p0 = 2, and a row of 1100 sites (past the
    line-walk limit of 1024: the per-site histogram kernel) against the oracle."""
    _, _, _, s = rand_state((12, 9, 7), 31)
    planes = [[2, 3, -1], [1, 0, 2], [0, 2, 1], [3, 1, 1], [-1, 2, 5], [2, 0, 0], [0, 0, -3], [1, -4, 2]]
    kvs = np.array([0.0, 0.07, 0.21, 0.5, 1.7])
    sf = structure_factors(s, kvs, planes)
    for p, plane in enumerate(planes):
        ref = np.array([abs(ora.get_sc(s, kv * np.array(plane, dtype=float)) / s.size) ** 2 for kv in kvs])
        assert np.all(np.abs(sf[p] - ref) <= 1e-10 * ref + 1e-15), plane
    _, _, _, w = rand_state((1100, 3, 2), 32)
    for plane in ([1, 0, 0], [1, 1, -1], [0, 1, 1]):
        got = structure_factors(w, kvs, [plane])[0]
        ref = np.array([abs(ora.get_sc(w, kv * np.array(plane, dtype=float)) / w.size) ** 2 for kv in kvs])
        assert np.all(np.abs(got - ref) <= 1e-10 * ref + 1e-15), plane
