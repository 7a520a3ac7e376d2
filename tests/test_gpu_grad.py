"""Differentiable lattice energy (chll_energy_grad, ChllEnergy autograd.Function).

Checks, each independent of the HIP kernel's own formulas:
* a CPU float64 reference built from torch tensor ops (rolls / slices, no
  per-bond derivative written by hand), differentiated by torch autograd;
* central finite differences at h = 1e-6 along random tangent fields, on small
  lattices and on the 16 x L=128 batch;
* E is linear in K, g and eps, so its parameter derivatives equal exact
  difference quotients;
* a gradient quench lowers the energy and hands over to exact Metropolis sweeps.
"""
import numpy as np
import pytest
import torch

from chllmodel_amd.autograd import ChllEnergy, energy_and_grad, gradient_quench, quench_then_metropolis

pytestmark = pytest.mark.gpu


def random_batch(B, nk, nj, ni, seed, device="cuda"):
    g = torch.Generator().manual_seed(seed)
    n = torch.randn(B, 3, nk, nj, ni, generator=g, dtype=torch.float64)
    n /= torch.linalg.vector_norm(n, dim=1, keepdim=True)
    s = torch.where(torch.rand(B, nk, nj, ni, generator=g) < 0.5, -1, 1).to(torch.int8)
    return n.to(device), s.to(device)


def ref_energy(n, s, K, g, eps, bc):
    """n (B,3,nk,nj,ni), s (B,nk,nj,ni) on the CPU; the Hamiltonian from tensor ops."""
    sf = s.to(torch.float64)
    E = -g * sf.sum(dim=(1, 2, 3))
    for axis, dim in ((0, 4), (1, 3), (2, 2)):
        if bc == "periodic":
            a, b = n, torch.roll(n, -1, dims=dim)
            sa, sb = sf, torch.roll(sf, -1, dims=dim - 1)
        else:
            size = n.shape[dim]
            a, b = n.narrow(dim, 0, size - 1), n.narrow(dim, 1, size - 1)
            sa, sb = sf.narrow(dim - 1, 0, size - 1), sf.narrow(dim - 1, 1, size - 1)
        c = (a * b).sum(dim=1)
        cr = torch.cross(a, b, dim=1)[:, axis]
        f = 0.5 * (sa + sb)
        t = eps.view(-1, 1, 1, 1) * (1 - c * c) - K.view(-1, 1, 1, 1) * c * cr * f
        E = E + t.sum(dim=(1, 2, 3))
    return E


@pytest.mark.parametrize("bc", ["open", "periodic"])
# (nk, nj, ni); (21, 9, 96) marches 3 slabs (the last partial) of 4 tiles that
# split rows (48 pairs per row, 256 per tile); (64, 1, 2) takes the one-pass
# two-site kernel (the march's blocks would exceed the workspace)
@pytest.mark.parametrize("shape", [(2, 3, 4), (6, 5, 4), (8, 8, 8), (1, 4, 6), (3, 4, 5), (2, 3, 2), (4, 1, 7),
                                   (21, 9, 96), (64, 1, 2)])
def test_against_cpu_autograd(bc, shape):
    B = 3
    n, s = random_batch(B, *shape, seed=sum(shape))
    K = torch.tensor([0.5, 1.0, 2.0], dtype=torch.float64)
    g = torch.tensor([0.0, 0.3, -1.0], dtype=torch.float64)
    eps = torch.tensor([1.0, 0.7, 1.3], dtype=torch.float64)
    E, dn, dK, dg, de = energy_and_grad(n, s, K.cuda(), g.cuda(), eps.cuda(), bc=bc)
    nc = n.cpu().requires_grad_(True)
    Kc, gc, ec = (x.clone().requires_grad_(True) for x in (K, g, eps))
    Er = ref_energy(nc, s.cpu(), Kc, gc, ec, bc)
    Er.sum().backward()
    G = nc.grad
    Gp = G - (G * nc.detach()).sum(dim=1, keepdim=True) * nc.detach()
    assert torch.allclose(E.cpu(), Er.detach(), rtol=1e-13, atol=1e-11)
    assert (dn.cpu() - Gp).abs().max() < 1e-12
    assert torch.allclose(dK.cpu(), Kc.grad, rtol=1e-13, atol=1e-11)
    assert torch.allclose(dg.cpu(), gc.grad, rtol=1e-13, atol=1e-11)
    assert torch.allclose(de.cpu(), ec.grad, rtol=1e-13, atol=1e-11)


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_whole_lattices_per_xcd_match_cpu_autograd(bc):
    """A batch of 8 (a multiple of the XCD count) takes the march kernel's
    one-lattice-per-XCD block order; 8 lattices of (nk, nj, ni) = (19, 6, 40):
    3 slabs (the last partial) of one tile each, against CPU autograd."""
    B = 8
    n, s = random_batch(B, 19, 6, 40, seed=88)
    K = torch.linspace(0.5, 2.0, B, dtype=torch.float64)
    g = torch.linspace(0.0, 1.0, B, dtype=torch.float64)
    eps = torch.linspace(0.8, 1.2, B, dtype=torch.float64)
    E, dn, dK, dg, de = energy_and_grad(n, s, K.cuda(), g.cuda(), eps.cuda(), bc=bc)
    nc = n.cpu().requires_grad_(True)
    Kc, gc, ec = (x.clone().requires_grad_(True) for x in (K, g, eps))
    Er = ref_energy(nc, s.cpu(), Kc, gc, ec, bc)
    Er.sum().backward()
    G = nc.grad
    Gp = G - (G * nc.detach()).sum(dim=1, keepdim=True) * nc.detach()
    assert torch.allclose(E.cpu(), Er.detach(), rtol=1e-13, atol=1e-11)
    assert (dn.cpu() - Gp).abs().max() < 1e-12
    assert torch.allclose(dK.cpu(), Kc.grad, rtol=1e-13, atol=1e-11)
    assert torch.allclose(dg.cpu(), gc.grad, rtol=1e-13, atol=1e-11)
    assert torch.allclose(de.cpu(), ec.grad, rtol=1e-13, atol=1e-11)


def test_block_orders_give_the_same_gradient():
    """The march kernel's two block orders (a batch of 8: whole lattices per
    XCD; a batch of 3: slabs of every lattice across the XCDs) give every
    site's gradient bit for bit (each site's bonds are summed in one fixed
    order by one thread), and E to the rounding of the partial sums."""
    n, s = random_batch(8, 24, 16, 32, seed=5)
    K = torch.linspace(0.5, 2.0, 8, dtype=torch.float64, device="cuda")
    g = torch.linspace(0.0, 1.0, 8, dtype=torch.float64, device="cuda")
    eps = torch.ones(8, dtype=torch.float64, device="cuda")
    E8, dn8, *_ = energy_and_grad(n, s, K, g, eps)
    E3, dn3, *_ = energy_and_grad(n[:3].contiguous(), s[:3].contiguous(), K[:3], g[:3], eps[:3])
    assert torch.equal(dn8[:3], dn3)
    assert torch.allclose(E8[:3], E3, rtol=1e-14, atol=0)


def _tangent_field(n, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    t = torch.randn(n.shape, generator=g, dtype=torch.float64).to(n.device)
    return t - (t * n).sum(dim=1, keepdim=True) * n


@pytest.mark.parametrize("shape,B", [((6, 6, 6), 2), ((128, 128, 128), 16)])
def test_central_differences(shape, B):
    """(E(n + h t) - E(n - h t)) / 2h = <dE/dn, t> for tangent t, h = 1e-6."""
    print(f"central differences B={B} shape={shape}", flush=True)
    n, s = random_batch(B, *shape, seed=7)
    K = torch.linspace(0.5, 2.0, B, dtype=torch.float64, device="cuda")
    g = torch.linspace(0.0, 1.0, B, dtype=torch.float64, device="cuda")
    eps = torch.ones(B, dtype=torch.float64, device="cuda")
    E, dn, dK, dg, de = energy_and_grad(n, s, K, g, eps)
    h = 1e-6
    for trial in range(2):
        t = _tangent_field(n, 100 + trial)
        Ep, *_ = energy_and_grad(n + h * t, s, K, g, eps, want_grad=False)
        Em, *_ = energy_and_grad(n - h * t, s, K, g, eps, want_grad=False)
        fd = (Ep - Em) / (2 * h)
        an = (dn * t).sum(dim=(1, 2, 3, 4))
        scale = an.abs() + (dn.abs() * t.abs()).sum(dim=(1, 2, 3, 4)) * 1e-3
        rel = ((fd - an).abs() / scale).max().item()
        # FD truncation O(h^2) plus cancellation ~ eps_mach |E| / h
        assert rel < 1e-5, rel
    # E is linear in K, g, eps: exact difference quotients
    for which, dd in ((0, dK), (1, dg), (2, de)):
        pars = [K.clone(), g.clone(), eps.clone()]
        pars[which] = pars[which] + 0.25
        E2, *_ = energy_and_grad(n, s, *pars, want_grad=False)
        assert torch.allclose((E2 - E) / 0.25, dd, rtol=1e-9, atol=1e-9 * float(E.abs().max()))


def test_autograd_function_chain():
    n, s = random_batch(2, 5, 6, 4, seed=3)
    n.requires_grad_(True)
    K = torch.tensor([1.0, 1.5], dtype=torch.float64, device="cuda", requires_grad=True)
    g = torch.tensor([0.2, 0.4], dtype=torch.float64, device="cuda", requires_grad=True)
    eps = torch.tensor([1.0, 1.0], dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor([2.0, -3.0], dtype=torch.float64, device="cuda")
    E = ChllEnergy.apply(n, K, g, eps, s, "open")
    (E * w).sum().backward()
    E0, dn, dK, dg, de = energy_and_grad(n.detach(), s, K.detach(), g.detach(), eps.detach())
    assert torch.allclose(n.grad, w.view(-1, 1, 1, 1, 1) * dn)
    assert torch.allclose(K.grad, w * dK) and torch.allclose(g.grad, w * dg) and torch.allclose(eps.grad, w * de)


def test_autograd_broadcast_parameters():
    """Parameters shared by the batch (a (1,) tensor, a 0-d tensor, a float32
    tensor, a plain number) get gradients of their own shape and dtype, the
    batch's per-lattice derivatives summed."""
    n, s = random_batch(3, 4, 5, 6, seed=8)
    K = torch.tensor([1.25], dtype=torch.float64, device="cuda", requires_grad=True)
    g = torch.tensor(0.3, dtype=torch.float64, device="cuda", requires_grad=True)
    eps = torch.tensor([1.0], dtype=torch.float32, device="cuda", requires_grad=True)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    (ChllEnergy.apply(n, K, g, eps, s, "periodic") * w).sum().backward()
    _, _, dK, dg, de = energy_and_grad(n, s, K.detach(), g.detach(), eps.detach(), bc="periodic")
    assert K.grad.shape == (1,) and g.grad.shape == () and eps.grad.shape == (1,)
    assert eps.grad.dtype == torch.float32
    assert torch.allclose(K.grad, (w * dK).sum().view(1))
    assert torch.allclose(g.grad, (w * dg).sum())
    assert torch.allclose(eps.grad, (w * de).sum().view(1).float())
    E = ChllEnergy.apply(n, 1.25, 0.3, 1.0, s, "periodic")  # plain numbers: no gradient, same energy
    E0, *_ = energy_and_grad(n, s, 1.25, 0.3, 1.0, bc="periodic")
    assert torch.equal(E, E0)


def test_quench_lowers_energy_then_metropolis():
    n, s = random_batch(2, 16, 16, 16, seed=11)
    nq, hist = gradient_quench(n, s, K=1.0, g=0.5, steps=100, lr=0.05)
    e = hist.cpu().numpy()
    assert np.all(np.diff(e, axis=0) <= 1e-9 * np.abs(e[:-1]))  # monotone for a small step
    assert torch.allclose(torch.linalg.vector_norm(nq, dim=1), torch.ones(1, dtype=torch.float64, device="cuda"))
    hist, rows, finals = quench_then_metropolis(n, s, 1.0, 0.5, steps=50, lr=0.05, sweeps=200, kT=0.2, nout=50)
    assert len(rows) == 2 and rows[0].shape[0] == 4
    assert np.all(np.isfinite(rows[0][:, 6]))
