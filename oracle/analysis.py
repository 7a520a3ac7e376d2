"""CPU restatement of the reference's analysis helpers (numpy).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` (and the CPU-baseline leg of
``bench.py``) as the checker for ``chll_shell_corr`` and
``chll_structure_factor``.  The product library never imports this module.

* ``calc_n_corr_loops`` / ``calc_s_corr_loops``: the six nested loops of
  calc_n_corr / calc_s_corr (notebooks/paper_plots2.ipynb, cell 0) written out
  literally in Python, same visiting order and accumulation; tiny lattices only.
* ``calc_n_corr`` / ``calc_s_corr``: the same sums vectorised per offset d over
  the overlap box of the lattice and its shift (every ordered pair once, as the
  loops visit them, but summed in numpy order).
* ``get_sc_loops`` / ``get_sc`` / ``structure_factor``: the structure factor of
  notebooks/.ipynb_checkpoints/hero_stripe-checkpoint.py:10-29, the sum over
  every site of s exp(-1j kvec . (i+1, j+1, k+1)), literally and vectorised.

Pins (tests/test_analysis_oracle.py): the vectorised forms against the loops
on small random lattices; closed forms (uniform directors give 1 in every
shell; uniform spins give 0; the checkerboard gives (-1)^{di+dj+dk} per
offset; the counts equal sum_d prod_c (n_c - |d_c|); a uniform field along
(1, 0, 0) gives the Dirichlet kernel; S(0) = <s>^2).
"""
from __future__ import annotations

import math

import numpy as np


def _offsets(nshells):
    """Every offset with |d|^2 < nshells^2 and its shell floor(|d|)."""
    m = nshells - 1
    out = []
    for di in range(-m, m + 1):
        for dj in range(-m, m + 1):
            for dk in range(-m, m + 1):
                r2 = di * di + dj * dj + dk * dk
                if r2 < nshells * nshells:
                    out.append((di, dj, dk, int(math.floor(math.sqrt(r2)))))
    return out


def _overlap(shape, d):
    """Slices (a, b) with b = a shifted by d, both inside the lattice."""
    sa, sb = [], []
    for n, dc in zip(shape, d):
        if abs(dc) >= n:
            return None
        if dc >= 0:
            sa.append(slice(0, n - dc)); sb.append(slice(dc, n))
        else:
            sa.append(slice(-dc, n)); sb.append(slice(0, n + dc))
    return tuple(sa), tuple(sb)


def corr_counts(shape, nshells):
    c = np.zeros(nshells, dtype=np.int64)
    for di, dj, dk, sh in _offsets(nshells):
        a = [n - abs(x) for n, x in zip(shape, (di, dj, dk))]
        if min(a) > 0:
            c[sh] += a[0] * a[1] * a[2]
    return c


def calc_n_corr(nx, ny, nz, nshells):
    """Vectorised calc_n_corr: (g, c) with g[sh] = mean of (n . n')^2."""
    shape = nx.shape
    g = np.zeros(nshells)
    c = np.zeros(nshells)
    for di, dj, dk, sh in _offsets(nshells):
        ov = _overlap(shape, (di, dj, dk))
        if ov is None:
            continue
        a, b = ov
        dot = nx[a] * nx[b] + ny[a] * ny[b] + nz[a] * nz[b]
        g[sh] += np.sum(dot * dot)
        c[sh] += dot.size
    return g / c, c


def calc_s_corr(s, nshells):
    """Vectorised calc_s_corr: (g, c) with g[sh] = mean of s s' - <s>^2
    (integer pair sums, exact)."""
    shape = s.shape
    si = np.asarray(s, dtype=np.int64)
    m2 = np.mean(s) ** 2
    S = np.zeros(nshells, dtype=np.int64)
    c = np.zeros(nshells)
    for di, dj, dk, sh in _offsets(nshells):
        ov = _overlap(shape, (di, dj, dk))
        if ov is None:
            continue
        a, b = ov
        S[sh] += int(np.sum(si[a] * si[b]))
        c[sh] += si[a].size
    return S / c - m2, c


def calc_n_corr_loops(nx, ny, nz, nshells):
    """The reference's loops verbatim in structure (small lattices only)."""
    g = np.zeros(nshells)
    c = np.zeros(nshells)
    ni, nj, nk = nx.shape
    for i in range(ni):
        for j in range(nj):
            for k in range(nk):
                for di in range(-nshells, nshells + 1):
                    for dj in range(-nshells, nshells + 1):
                        for dk in range(-nshells, nshells + 1):
                            r2 = di * di + dj * dj + dk * dk
                            if r2 < nshells * nshells:
                                ii, jj, kk = i + di, j + dj, k + dk
                                if 0 <= ii < ni and 0 <= jj < nj and 0 <= kk < nk:
                                    sh = int(math.floor(math.sqrt(r2)))
                                    v = nx[i, j, k] * nx[ii, jj, kk] + ny[i, j, k] * ny[ii, jj, kk] + \
                                        nz[i, j, k] * nz[ii, jj, kk]
                                    g[sh] = g[sh] + v ** 2
                                    c[sh] += 1
    return g / c, c


def calc_s_corr_loops(s, nshells):
    g = np.zeros(nshells)
    c = np.zeros(nshells)
    ni, nj, nk = s.shape
    s_avg2 = np.mean(s) ** 2
    for i in range(ni):
        for j in range(nj):
            for k in range(nk):
                for di in range(-nshells, nshells + 1):
                    for dj in range(-nshells, nshells + 1):
                        for dk in range(-nshells, nshells + 1):
                            r2 = di * di + dj * dj + dk * dk
                            if r2 < nshells * nshells:
                                ii, jj, kk = i + di, j + dj, k + dk
                                if 0 <= ii < ni and 0 <= jj < nj and 0 <= kk < nk:
                                    sh = int(math.floor(math.sqrt(r2)))
                                    g[sh] = g[sh] + float(s[i, j, k]) * float(s[ii, jj, kk]) - s_avg2
                                    c[sh] += 1
    return g / c, c


def get_sc_loops(s, kvec):
    sc = 0.0 + 0.0j
    for i in range(s.shape[0]):
        for j in range(s.shape[1]):
            for k in range(s.shape[2]):
                sc += s[i, j, k] * np.exp(-1j * np.dot(kvec, np.array([float(i + 1), float(j + 1), float(k + 1)])))
    return sc


def get_sc(s, kvec):
    ni, nj, nk = s.shape
    ph = (kvec[0] * np.arange(1, ni + 1, dtype=np.float64)[:, None, None]
          + kvec[1] * np.arange(1, nj + 1, dtype=np.float64)[None, :, None]
          + kvec[2] * np.arange(1, nk + 1, dtype=np.float64)[None, None, :])
    return np.sum(np.asarray(s, dtype=np.float64) * np.exp(-1j * ph))


def structure_factor(data, bins, k_min, k_max, plane):
    N3 = data.shape[0] * data.shape[1] * data.shape[2]
    kvs = np.linspace(k_min, k_max, bins)
    sc = np.zeros(len(kvs), dtype=complex)
    for idx, kv in enumerate(kvs):
        sc[idx] = get_sc(data, kv * np.array(plane, dtype=np.float64)) / N3
    sf = sc.real * sc.real + sc.imag * sc.imag
    return sf, kvs
