"""Initial lattice states, vectorised in the reference scripts' draw order.

The reference builds them with Python triple loops over (i, j, k), i
outermost, drawing from numpy's global legacy generator; a C-ordered numpy
draw of the same shape consumes the identical sequence, so these functions
return bit-identical arrays for the same generator state.
"""
from __future__ import annotations

import numpy as np


def _rng(seed):
    return np.random.RandomState(seed) if seed is not None else np.random.mtrand._rand


def isotropic_directors(ni, nj, nk, seed=None):
    """Isotropic unit directors, scripts/pdanneal.py:5-16 (init_random):
    per site costh = 2(u1 - 0.5), sinth = sqrt(1 - costh^2), phi = 2 pi u2."""
    u = _rng(seed).random_sample((ni, nj, nk, 2))
    costh = 2.0 * (u[..., 0] - 0.5)
    sinth = np.sqrt(1.0 - costh * costh)
    phi = u[..., 1] * 2.0 * np.pi
    return (np.asfortranarray(sinth * np.cos(phi)), np.asfortranarray(sinth * np.sin(phi)),
            np.asfortranarray(costh))


def random_spins(ni, nj, nk, seed=None, p=0.5):
    """s = +1, then -1 where u < p, chll/chll.py:85-90 (init_rand_s)."""
    u = _rng(seed).random_sample((ni, nj, nk))
    return np.asfortranarray(np.where(u < p, -1, 1).astype(np.int32))


def aligned_directors(ni, nj, nk, kind="x"):
    """'x': n = (1,0,0) (chll/chll.py:91-98); 'xz': n = (1/sqrt2, 0, 1/sqrt2)
    (scripts/hero480.py:18-21, init_aligned)."""
    if kind == "x":
        v = (1.0, 0.0, 0.0)
    elif kind == "xz":
        v = (1.0 / np.sqrt(2), 0.0, 1.0 / np.sqrt(2))
    else:
        raise ValueError(kind)
    return tuple(np.full((ni, nj, nk), c, dtype=np.float64, order="F") for c in v)


def isotropic_state(ni, nj, nk, seed, p=0.5):
    """Directors then spins from one seeded generator (the order a reference
    script draws them: init_random(sim) then sim.init_rand_s())."""
    rs = np.random.RandomState(seed)
    u = rs.random_sample((ni, nj, nk, 2))
    costh = 2.0 * (u[..., 0] - 0.5)
    sinth = np.sqrt(1.0 - costh * costh)
    phi = u[..., 1] * 2.0 * np.pi
    us = rs.random_sample((ni, nj, nk))
    return {
        "nx": np.asfortranarray(sinth * np.cos(phi)),
        "ny": np.asfortranarray(sinth * np.sin(phi)),
        "nz": np.asfortranarray(costh),
        "s": np.asfortranarray(np.where(us < p, -1, 1).astype(np.int32)),
        "dope": np.zeros((ni, nj, nk), dtype=np.int32, order="F"),
    }
