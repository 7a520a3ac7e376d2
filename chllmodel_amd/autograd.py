"""Differentiable chiral Lebwohl-Lasher lattice energy on the GPU.

``ChllEnergy`` is a ``torch.autograd.Function`` over a batch of lattices:

    E_b = sum_bonds [eps_b (1 - c^2) - K_b c q f] - g_b sum_a s_a

with c = n_a . n_b, q = (n_a x n_b)_axis, f = (s_a + s_b)/2 over the +x, +y,
+z bonds of each site (the Hamiltonian of chll/chll_mod.f90:369-411, each
bond once; the reference fixes eps = 1).  Forward and backward run in one HIP
kernel (``chll_energy_grad`` in libchll_mod.so); the director gradient is
returned projected on each director's tangent plane, the torque a unit-length
director feels.  Spins stay integer and are not differentiated.

Tensors: n (B, 3, nk, nj, ni) float64 on the GPU (the Fortran (ni, nj, nk)
order of the reference's arrays, component planes first), s (B, nk, nj, ni)
int8, eps / K / g (B,) float64.

``gradient_quench`` relaxes directors by projected gradient descent; the result
can be refined by exact Metropolis sweeps (``quench_then_metropolis``).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from ._lib import BC, lib

_vp = ct.c_void_p


def _declare(L):
    if getattr(L, "_grad_declared", False):
        return L
    L.chll_energy_last_error.restype = ct.c_char_p
    L.chll_energy_workspace_bytes.argtypes = [ct.c_int32] * 4
    L.chll_energy_workspace_bytes.restype = ct.c_size_t
    L.chll_energy_grad.argtypes = [_vp, _vp, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32,
                                   _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ct.c_size_t, _vp]
    L.chll_energy_grad.restype = ct.c_int
    L._grad_declared = True
    return L


def _ptr(t):
    return _vp(t.data_ptr()) if t is not None else None


def energy_and_grad(n, s, K, g, eps, bc="open", want_grad=True):
    """Returns (E, dE/dn tangent-projected or None, dE/dK, dE/dg, dE/deps)."""
    L = _declare(lib())
    if n.dim() != 5 or n.shape[1] != 3:
        raise ValueError("n must have shape (B, 3, nk, nj, ni)")
    if n.dtype != torch.float64 or not n.is_cuda or s.dtype != torch.int8 or not s.is_cuda:
        raise TypeError("n must be float64 and s int8, both on the GPU")
    n = n.contiguous()
    s = s.contiguous()
    B, _, nk, nj, ni = n.shape
    if tuple(s.shape) != (B, nk, nj, ni):
        raise ValueError("s must have shape (B, nk, nj, ni)")
    dev = n.device
    K, g, eps = (torch.as_tensor(x, dtype=torch.float64, device=dev).expand(B).contiguous() for x in (K, g, eps))
    E = torch.empty(B, dtype=torch.float64, device=dev)
    dK, dg, de = torch.empty_like(E), torch.empty_like(E), torch.empty_like(E)
    dn = torch.empty_like(n) if want_grad else None
    wsb = L.chll_energy_workspace_bytes(B, ni, nj, nk)
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = L.chll_energy_grad(_ptr(n), _ptr(s), B, ni, nj, nk, BC[bc], _ptr(K), _ptr(g), _ptr(eps), _ptr(E), _ptr(dn),
                            _ptr(dK), _ptr(dg), _ptr(de), _ptr(ws), wsb, _vp(stream))
    if rc != 0:
        raise RuntimeError(L.chll_energy_last_error().decode())
    return E, dn, dK, dg, de


class ChllEnergy(torch.autograd.Function):
    """E = ChllEnergy.apply(n, K, g, eps, s, bc) with n, K, g, eps
    differentiable; bc is 'open' or 'periodic'."""

    @staticmethod
    def forward(ctx, n, K, g, eps, s, bc="open"):
        want = any(ctx.needs_input_grad[:4])
        E, dn, dK, dg, de = energy_and_grad(n, s, K, g, eps, bc=bc, want_grad=ctx.needs_input_grad[0])
        ctx.save_for_backward(dn if dn is not None else torch.empty(0), dK, dg, de)
        # parameters may be scalars, (1,) or (B,) tensors (broadcast over the
        # batch) or plain numbers; their gradients are summed back to that shape
        ctx.pshapes = tuple((x.shape, x.dtype) if torch.is_tensor(x) else None for x in (K, g, eps))
        ctx.want = want
        return E

    @staticmethod
    def backward(ctx, gE):
        dn, dK, dg, de = ctx.saved_tensors
        gn = gK = gg = ge = None
        if ctx.needs_input_grad[0]:
            gn = gE.view(-1, 1, 1, 1, 1) * dn
        grads = []
        for i, d in enumerate((dK, dg, de)):
            ps = ctx.pshapes[i]
            grads.append((gE * d).sum_to_size(ps[0]).to(ps[1]) if ctx.needs_input_grad[1 + i] and ps else None)
        gK, gg, ge = grads
        return gn, gK, gg, ge, None, None


def chll_energy(n, K, g, eps, s, bc="open"):
    return ChllEnergy.apply(n, K, g, eps, s, bc)


def gradient_quench(n, s, K, g, eps=1.0, steps=500, lr=0.05, bc="open"):
    """Projected gradient descent on the directors, renormalised each step.
    Returns (n, E per step)."""
    n = n.clone()
    hist = []
    for _ in range(steps):
        E, dn, *_ = energy_and_grad(n, s, K, g, eps, bc=bc)
        hist.append(E)
        n.add_(dn, alpha=-lr)
        n.div_(torch.linalg.vector_norm(n, dim=1, keepdim=True))
    return n, torch.stack(hist)


def to_fortran_state(n, s, b=0):
    """Lattice b of a (B, 3, nk, nj, ni) batch as the reference's Fortran-ordered
    (ni, nj, nk) arrays."""
    a = n[b].detach().cpu().numpy()
    sb = s[b].detach().cpu().numpy().astype(np.int32)
    return {"nx": np.asfortranarray(a[0].T), "ny": np.asfortranarray(a[1].T), "nz": np.asfortranarray(a[2].T),
            "s": np.asfortranarray(sb.T)}


def from_fortran_state(st, device="cuda"):
    n = np.stack([st["nx"].T, st["ny"].T, st["nz"].T])[None]
    s = np.asarray(st["s"]).T[None].astype(np.int8)
    return (torch.from_numpy(np.ascontiguousarray(n)).to(device),
            torch.from_numpy(np.ascontiguousarray(s)).to(device))


def quench_then_metropolis(n, s, K, g, *, steps=500, lr=0.05, sweeps=2000, kT=0.2, d0=0.3, bc="open",
                           arith="fp64", seed=1, nout=100):
    """Gradient quench of every lattice of the batch, then exact Metropolis
    sweeps of each at kT; returns (quench energies (steps, B), list of row
    arrays, final Fortran-ordered states)."""
    from .lattice import Lattice

    nq, hist = gradient_quench(n, s, K, g, steps=steps, lr=lr, bc=bc)
    B = n.shape[0]
    Kt = torch.as_tensor(K, dtype=torch.float64).expand(B)
    gt = torch.as_tensor(g, dtype=torch.float64).expand(B)
    rows, finals = [], []
    for b in range(B):
        st = to_fortran_state(nq, s, b)
        ni, nj, nk = st["nx"].shape
        with Lattice(ni, nj, nk, bc=bc) as lat:
            lat.upload_state(st)
            lat.set_step(d0)
            out = lat.sweeps(1, sweeps, nout, g=float(gt[b]), KK=float(Kt[b]), kbt=kT, arith=arith, seed=seed,
                             replica=b)
            rows.append(out["rows"])
            finals.append(lat.download())
    return hist, rows, finals
