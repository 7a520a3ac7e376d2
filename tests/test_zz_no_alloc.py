"""chll_energy_grad_noalloc: the energy gradient with its workspace in the
library's static per-GPU scratch buffer instead of a caller argument.

Called through the C ABI and compared with chll_energy_grad (caller
workspace) on the same inputs: the same kernels run, so E, the tangent
gradient and the parameter derivatives must be bit-identical.  Two calls in a
row, the second on a smaller input (and one-site rows, odd ni), so the second
reads none of the first call's partial sums; then a size past the buffer is
refused with nothing launched.
"""
import ctypes as ct

import pytest
import torch

from chllmodel_amd import _lib

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 8 << 20


def _p(t):
    return ct.c_void_p(t.data_ptr()) if t is not None else ct.c_void_p(0)


def _lib_noalloc():
    L = _lib.lib()
    vp = ct.c_void_p
    L.chll_energy_workspace_bytes.argtypes = [ct.c_int32] * 4
    L.chll_energy_workspace_bytes.restype = ct.c_size_t
    L.chll_energy_grad.argtypes = [vp, vp, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32,
                                   vp, vp, vp, vp, vp, vp, vp, vp, vp, ct.c_size_t, vp]
    L.chll_energy_grad.restype = ct.c_int
    L.chll_energy_grad_noalloc.argtypes = [vp, vp, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32,
                                           vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.chll_energy_grad_noalloc.restype = ct.c_int
    L.chll_energy_last_error.restype = ct.c_char_p
    return L


def _inputs(B, ni, nj, nk, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    n = torch.randn(B, 3, nk, nj, ni, device="cuda", dtype=torch.float64, generator=gen)
    n = (n / n.norm(dim=1, keepdim=True)).contiguous()
    s = (torch.randint(0, 2, (B, nk, nj, ni), device="cuda", generator=gen) * 2 - 1).to(torch.int8).contiguous()
    K = torch.linspace(0.5, 2.0, B, device="cuda", dtype=torch.float64)
    g = torch.linspace(0.0, 1.0, B, device="cuda", dtype=torch.float64)
    eps = torch.ones(B, device="cuda", dtype=torch.float64)
    return n, s, K, g, eps


def _outputs(n):
    B = n.shape[0]
    return [torch.full((B,), float("nan"), device="cuda", dtype=torch.float64) for _ in range(4)] + \
        [torch.full_like(n, float("nan"))]


@pytest.mark.parametrize("bc", [0, 1])
def test_noalloc_matches_workspace_entry(bc):
    L = _lib_noalloc()
    stream = torch.cuda.current_stream().cuda_stream
    # second call smaller; periodic needs even dims, open takes odd ni (one site per thread)
    shapes = [(6, 64, 48, 40, 1), (3, 16, 12, 10, 2)] if bc else [(6, 64, 48, 40, 1), (3, 15, 12, 10, 2)]
    for B, ni, nj, nk, seed in shapes:
        print(f"noalloc bc={bc} B={B} {ni}x{nj}x{nk}", flush=True)
        assert L.chll_energy_workspace_bytes(B, ni, nj, nk) <= SCRATCH_BYTES
        n, s, K, g, eps = _inputs(B, ni, nj, nk, seed)
        E1, dK1, dg1, de1, dn1 = _outputs(n)
        ws = torch.empty(max(L.chll_energy_workspace_bytes(B, ni, nj, nk), 8), dtype=torch.uint8, device="cuda")
        rc = L.chll_energy_grad(_p(n), _p(s), B, ni, nj, nk, bc, _p(K), _p(g), _p(eps), _p(E1), _p(dn1), _p(dK1),
                                _p(dg1), _p(de1), _p(ws), ws.numel(), ct.c_void_p(stream))
        assert rc == 0, L.chll_energy_last_error()
        E2, dK2, dg2, de2, dn2 = _outputs(n)
        rc = L.chll_energy_grad_noalloc(_p(n), _p(s), B, ni, nj, nk, bc, _p(K), _p(g), _p(eps), _p(E2), _p(dn2),
                                        _p(dK2), _p(dg2), _p(de2), ct.c_void_p(stream))
        assert rc == 0, L.chll_energy_last_error()
        torch.cuda.synchronize()
        for a, b in ((E1, E2), (dK1, dK2), (dg1, dg2), (de1, de2), (dn1, dn2)):
            assert not torch.isnan(a).any()
            assert torch.equal(a, b)


def test_noalloc_rejects_sizes_past_the_buffer():
    L = _lib_noalloc()
    B, ni, nj, nk = 1, 600, 600, 600
    assert L.chll_energy_workspace_bytes(B, ni, nj, nk) > SCRATCH_BYTES
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    sp = torch.zeros(8, dtype=torch.int8, device="cuda")
    rc = L.chll_energy_grad_noalloc(_p(t), _p(sp), B, ni, nj, nk, 0, _p(t), _p(t), _p(t), _p(t), None, None, None,
                                    None, ct.c_void_p(0))
    assert rc == -1
    assert b"static buffer" in L.chll_energy_last_error()
    torch.cuda.synchronize()
