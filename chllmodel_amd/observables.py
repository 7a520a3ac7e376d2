"""Observables and statistics of chiral Lebwohl-Lasher runs.

* ``p2_from_q``: nematic order P2, the largest eigenvalue of
  Q = <(3 n n - I)/2> built from the device's <n_a n_b> sums.  (The reference
  computes no order parameter; its analysis notebooks use the energy and the
  enantiomeric excess only.)
* ``cv_notebook`` / ``chi_notebook``: the reference notebooks' definitions,
  C_v = (<E^2> - <E>^2)/kT over the last 100 samples of the reported energy
  per site (notebooks/paper_plots2.ipynb, calc_specific_heat) and
  chi_c = N (<eta^2> - <eta>^2)/kT over the last 500 samples of eta
  (notebooks/.ipynb_checkpoints/paper_plots-checkpoint.ipynb:440-452).
* ``cv_per_site`` / ``chi_per_site``: the canonical fluctuation formulas
  C_v = N var(e)/kT^2 and chi = N var(eta)/kT with e = E/N, with jackknife
  errors over blocks.
* ``blocking_error``: Flyvbjerg-Petersen blocking estimate of the standard
  error of a correlated time series' mean.
"""
from __future__ import annotations

import numpy as np


def q_tensor(qsums) -> np.ndarray:
    """Q = (3 <n n> - I)/2 from (xx, yy, zz, xy, xz, yz) means."""
    xx, yy, zz, xy, xz, yz = qsums
    M = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]], dtype=np.float64)
    return 1.5 * M - 0.5 * np.eye(3)


def p2_from_q(qsums) -> float:
    return float(np.linalg.eigvalsh(q_tensor(qsums))[-1])


def p2_series(rows: np.ndarray, first_col: int) -> np.ndarray:
    """P2 per row from the six <n_a n_b> columns starting at first_col."""
    return np.array([p2_from_q(r[first_col:first_col + 6]) for r in rows])


def blocking_error(x) -> float:
    """Flyvbjerg & Petersen (J. Chem. Phys. 91, 461 (1989)): halve the series
    by pairwise averaging until the error estimate stops growing; return the
    largest estimate over the levels with at least 16 blocks."""
    x = np.asarray(x, dtype=np.float64)
    best = 0.0
    while x.size >= 16:
        n = x.size
        err = np.sqrt(np.var(x) / (n - 1))
        best = max(best, err)
        x = 0.5 * (x[: n // 2 * 2 : 2] + x[1 : n // 2 * 2 : 2])
    if best == 0.0 and x.size > 1:
        best = float(np.sqrt(np.var(x) / (x.size - 1)))
    return float(best)


def _jackknife(x, stat, nblocks=16):
    x = np.asarray(x, dtype=np.float64)
    nb = min(nblocks, x.size)
    if nb < 2:
        return float(stat(x)), float("nan")
    blocks = np.array_split(x, nb)
    full = stat(x)
    loo = np.array([stat(np.concatenate(blocks[:i] + blocks[i + 1:])) for i in range(nb)])
    err = np.sqrt((nb - 1) / nb * np.sum((loo - loo.mean()) ** 2))
    return float(full), float(err)


def cv_per_site(e_per_site, kT, N, nblocks=16):
    """C_v/N = N var(e)/kT^2, e = E/N; returns (value, jackknife error)."""
    return _jackknife(e_per_site, lambda y: N * np.var(y) / kT ** 2, nblocks)


def chi_per_site(eta, kT, N, nblocks=16):
    """chi = N var(eta)/kT; returns (value, jackknife error)."""
    return _jackknife(eta, lambda y: N * np.var(y) / kT, nblocks)


def cv_notebook(E_per_site, kT, last=100) -> float:
    E = np.asarray(E_per_site, dtype=np.float64)[-last:]
    return float((np.mean(E ** 2) - np.mean(E) ** 2) / kT)


def chi_notebook(eta, kT, N, last=500) -> float:
    x = np.asarray(eta, dtype=np.float64)[-last:]
    return float(N * (np.mean(x ** 2) - np.mean(x) ** 2) / kT)


def summarize(series: dict, kT: float, N: int, discard: float = 0.25) -> dict:
    """Means with blocking errors, C_v and chi with jackknife errors, from a
    dict of time series {'E': E/N, 'eta': mean spin, 'Q': (T,6) or absent}."""
    out = {}
    n0 = int(len(series["E"]) * discard)
    e = np.asarray(series["E"])[n0:]
    eta = np.asarray(series["eta"])[n0:]
    out["E"], out["E_err"] = float(e.mean()), blocking_error(e)
    out["eta"], out["eta_err"] = float(eta.mean()), blocking_error(eta)
    out["Cv"], out["Cv_err"] = cv_per_site(e, kT, N)
    out["chi"], out["chi_err"] = chi_per_site(eta, kT, N)
    if "Q" in series:
        q = np.asarray(series["Q"])[n0:]
        p2 = np.array([p2_from_q(r) for r in q])
        out["P2"], out["P2_err"] = float(p2.mean()), blocking_error(p2)
    return out
