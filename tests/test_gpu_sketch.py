"""BF16 MFMA sketch products (ck_sketch_bf16, ck_round_bf16): the product against an FP64
product of the same bf16-rounded operands, the rounding against torch's bfloat16 cast, and
the learned and Nystrom preconditioners built from it against their FP32 constructions."""
import math

import numpy as np
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd._blocks import ld_of
from cola_amd.linalg import learned as L
from cola_amd.linalg import NystromPrecond
from cola_amd.linalg.cg import cg

pytestmark = pytest.mark.gpu


def test_round_bf16_matches_torch(dev):
    """Round to nearest even, bit for bit with torch's cast, ties and extremes included."""
    g = torch.Generator().manual_seed(0)
    x = torch.randn(5, 1000, generator=g) * torch.logspace(-30, 30, 1000)[None, :]
    ties = torch.tensor([1.0 + 2**-8, 1.0 + 3 * 2**-8, -(1.0 + 2**-8), 3.0e38, -3.0e38, 1e-40, -1e-40, 0.0, -0.0])
    x[0, :ties.numel()] = ties
    x = x.to(dev)
    n, rows, ld = 1000, 5, 1024
    X = torch.zeros((rows, ld), device=dev)
    X[:, :n] = x
    Xb = torch.zeros((rows, ld), dtype=torch.int16, device=dev)
    Xr = torch.zeros((rows, ld), device=dev)
    N.call("ck_round_bf16", n, rows, N.ptr(X), ld, N.ptr(Xb), ld, N.ptr(Xr), ld, N.stream_ptr(dev))
    ref = x.to(torch.bfloat16)
    assert torch.equal(Xb[:, :n], ref.view(torch.int16))
    assert torch.equal(Xr[:, :n], ref.float())


@pytest.mark.parametrize("n,k,scale,shift,sym,split", [(256, 32, 1.0, 0.0, False, True),
                                                       (4096, 64, 0.5, 0.25, False, True),
                                                       (8192, 64, 1.0, 0.0, True, True),
                                                       (1280, 32, -2.0, 1.5, False, False),
                                                       (4096, 64, 1.0, 0.0, False, False)])
def test_sketch_bf16_against_fp64(dev, n, k, scale, shift, sym, split):
    """Yt = scale A G + shift G for a non-symmetric A (so a transposed operand would show) against
    an FP64 product of the rounded operands (A as the bf16 pair hi + lo with split, else one
    bf16 value): within FP32 accumulation error, measured against sum_l |A_il| |G_lc|."""
    g = torch.Generator(device=dev).manual_seed(n + k)
    A = torch.randn(n, n, generator=g, device=dev)
    if sym:
        A = 0.5 * (A + A.T)
    ld = ld_of(n)
    G = torch.zeros((k, ld), device=dev)
    G[:, :n] = torch.randn(k, n, generator=g, device=dev)
    op = ca.Dense(A) if scale == 1.0 else scale * ca.Dense(A)
    if shift != 0.0:
        op = op + shift * ca.I_like(ca.Dense(A))
    low = op.hip_op()
    assert low is not None and low[0].kind == N.CK_OP_DENSE_F32
    Y, Gr = L.dense_sketch_bf16(low, G, n, split=split)
    Ahi = A.to(torch.bfloat16)
    A16 = Ahi.double() + ((A - Ahi.float()).to(torch.bfloat16).double() if split else 0.0)
    G16 = G[:, :n].to(torch.bfloat16).double()
    ref = scale * (G16 @ A16.T) + shift * G16
    bound = abs(scale) * (G16.abs() @ A16.abs().T) + abs(shift) * G16.abs()
    err = (Y[:, :n].double() - ref).abs()
    full = (Y[:, :n].double() - (scale * (G16 @ A.double().T) + shift * G16)).abs()
    print(n, k, "split" if split else "bf16", "max err / bound", float((err / bound).max()),
          "vs fp32 A:", float((full / bound).max()), flush=True)
    assert float((err / bound).max()) < 2e-6 * math.sqrt(n / 256)
    if split:  # the pair keeps 16 significant bits of A
        assert float((full / bound).max()) < 2e-5
    assert torch.equal(Gr[:, :n], G[:, :n].to(torch.bfloat16).float())
    assert float(Y[:, n:].abs().max()) == 0.0 if ld > n else True


def test_learned_preconditioner_bf16_sketch(dev):
    """The learned preconditioner from the BF16 MFMA sketch (K as a bf16 pair) against the FP32
    construction on a 16384-point RBF system: captured trace within 1e-3 relative, the factors'
    products F F^T close in norm, and the PCG step counts within 5%."""
    n, B = 16384, 2
    X, ell, noise = L.sample_instance(100003, n)
    Xd = X.to(dev)
    K = L.rbf_matrix(Xd, ell)
    net = L.PointMLP(4, 64, device=dev, seed=0)
    P16 = L.LearnedPreconditioner(K, Xd, net, noise)
    P32 = L.LearnedPreconditioner(K, Xd, net, noise, bf16_sketch=False)
    assert P16.sketch == "bf16_mfma" and P32.sketch == "fp32"
    rel = abs(P16.captured_trace - P32.captured_trace) / P32.captured_trace
    A = ca.PSD(ca.Dense(K) + noise * ca.I_like(ca.Dense(K)))
    b = torch.randn(n, B, generator=torch.Generator().manual_seed(1)).to(dev)
    _, i16 = cg(A, b, P=P16, tol=1e-5, max_iters=2000)
    _, i32 = cg(A, b, P=P32, tol=1e-5, max_iters=2000)
    V = torch.randn(n, 4, generator=torch.Generator().manual_seed(2)).to(dev)
    d = float((P16.F @ (P16.F.T @ V) - P32.F @ (P32.F.T @ V)).norm() / (P32.F @ (P32.F.T @ V)).norm())
    print("captured trace rel diff", rel, "FF^T V rel diff", d, "steps bf16", i16["iterations"], "fp32",
          i32["iterations"], flush=True)
    assert rel < 1e-3 and d < 2e-2
    assert abs(i16["iterations"] - i32["iterations"]) <= max(2, 0.05 * i32["iterations"])


def test_nystrom_bf16_sketch(dev):
    """NystromPrecond(bf16_sketch=True) on a 4096-point RBF matrix (K as a bf16 pair): captured
    eigenvalues within 1e-2 relative of the FP32 construction from the same key, and the same
    PCG step count within 5%."""
    n = 4096
    Xp = torch.rand(n, 3, generator=torch.Generator().manual_seed(3))
    K = torch.exp(-torch.cdist(Xp, Xp)**2 / (2 * 0.3**2)).to(dev)
    P16 = NystromPrecond(ca.Dense(K), rank=64, mu=1e-2, adjust_mu=False, bf16_sketch=True)
    P32 = NystromPrecond(ca.Dense(K), rank=64, mu=1e-2, adjust_mu=False)
    l16, l32 = P16.Lambda.double().cpu().numpy(), P32.Lambda.double().cpu().numpy()
    print("top eigenvalues", l16[:3], l32[:3], flush=True)
    assert np.abs(l16[:32] - l32[:32]).max() / l32[0] < 1e-2
    A = ca.PSD(ca.Dense(K) + 1e-2 * ca.I_like(ca.Dense(K)))
    b = torch.randn(n, 2, generator=torch.Generator().manual_seed(4)).to(dev)
    _, i16 = cg(A, b, P=P16, tol=1e-5, max_iters=2000)
    _, i32 = cg(A, b, P=P32, tol=1e-5, max_iters=2000)
    print("steps", i16["iterations"], i32["iterations"], flush=True)
    assert abs(i16["iterations"] - i32["iterations"]) <= max(2, 0.05 * i32["iterations"])
