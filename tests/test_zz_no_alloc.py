"""ck_cg_fused_noalloc (the allocation-free variant of ck_cg_fused, exchange buffer in a
file-scope device array) against ck_cg_fused with a caller workspace, through the C ABI:
identical x, r and control block on two calls in a row, the second smaller; sizes beyond the
array are refused with CK_ERR_WORKSPACE."""
import ctypes

import numpy as np
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd.linalg._engine import lower

pytestmark = pytest.mark.gpu


def _system(n, B, seed):
    g = torch.Generator().manual_seed(seed)
    M = torch.randn(n, n // 2, generator=g)
    A = M @ M.T / (n // 2) + 0.05 * torch.eye(n)
    return A, torch.randn(n, B, generator=g)


def _solve(dev, A, b, noalloc, max_iters=500):
    n, B = b.shape
    aop = lower(ca.Dense(A.to(dev)))
    bd = b.to(dev).contiguous()
    x = torch.full((n, B), float("nan"), device=dev)
    r = torch.full((n, B), float("nan"), device=dev)
    state = torch.zeros(N.q("ck_cg_state_bytes", B, max_iters), dtype=torch.uint8, device=dev)
    s = N.stream_ptr(dev)
    if noalloc:
        N.call("ck_cg_fused_noalloc", aop[0], None, B, N.ptr(bd), bd.stride(0), bd.stride(1), None, 0, 0, N.ptr(x),
               x.stride(0), x.stride(1), N.ptr(r), r.stride(0), r.stride(1), N.ptr(state), 1e-6, max_iters, 0, 0, 0, 0, s)
    else:
        wb = N.q("ck_cg_fused_workspace_bytes", n, B)
        work = torch.empty((wb + 7) // 8, dtype=torch.float64, device=dev)
        N.call("ck_cg_fused", aop[0], None, B, N.ptr(bd), bd.stride(0), bd.stride(1), None, 0, 0, N.ptr(x), x.stride(0),
               x.stride(1), N.ptr(r), r.stride(0), r.stride(1), N.ptr(state), 1e-6, max_iters, N.ptr(work),
               work.numel() * 8, 0, 0, 0, 0, s)
    torch.cuda.synchronize()
    return x.cpu(), r.cpu(), state.cpu()


def test_noalloc_matches_workspace_entry_point_twice(dev):
    # first call 1000 x 2, then a smaller 300 x 1 system: the shared array keeps the first call's
    # words, which the reset kernel clears
    for n, B, seed in ((1000, 2, 1), (300, 1, 2)):
        A, b = _system(n, B, seed)
        x1, r1, s1 = _solve(dev, A, b, noalloc=False)
        x2, r2, s2 = _solve(dev, A, b, noalloc=True)
        k = int(s1[:4].view(torch.int32)[0])
        print(n, B, "steps", k, flush=True)
        assert k > 0 and int(s2[:4].view(torch.int32)[0]) == k
        assert torch.equal(x1, x2) and torch.equal(r1, r2)
        assert torch.equal(s1, s2)
        xd = torch.linalg.solve(A.double(), b.double())
        assert float((x2.double() - xd).norm() / xd.norm()) < 2e-5


def test_noalloc_refuses_sizes_beyond_its_array(dev):
    cap = N.q("ck_cg_fused_noalloc_max_bytes")
    assert cap == 64 * 1024
    # a size the workspace entry point supports but whose exchange buffer exceeds the array
    n, B = 2500, 2
    assert N.q("ck_cg_fused_workspace_bytes", n, B) > cap
    A, b = _system(n, B, 3)
    aop = lower(ca.Dense(A.to(dev)))
    assert N.q("ck_cg_fused_supported", aop[0], None, B) == N.CK_OK
    bd = b.to(dev).contiguous()
    x = torch.empty((n, B), device=dev)
    state = torch.zeros(N.q("ck_cg_state_bytes", B, 10), dtype=torch.uint8, device=dev)
    with pytest.raises(N.KrylovError) as e:
        N.call("ck_cg_fused_noalloc", aop[0], None, B, N.ptr(bd), bd.stride(0), bd.stride(1), None, 0, 0, N.ptr(x),
               x.stride(0), x.stride(1), None, 0, 0, N.ptr(state), 1e-6, 10, 0, 0, 0, 0, N.stream_ptr(dev))
    assert e.value.status == N.CK_ERR_WORKSPACE


_ = (ctypes, np)
