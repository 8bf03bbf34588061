"""Symmetric dense product (CK_OP_SYMMETRIC: a self-adjoint annotated dense operator whose order
is a multiple of 128 reads only the 128-column strips on and above the diagonal) against the
oracle's dense matmat in FP64 and against the full product, and inside CG, whose p.Ap partials
then come from the symmetric product's reduction pass."""
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from oracle import operators as oop

pytestmark = pytest.mark.gpu


@pytest.fixture
def sym_everywhere(monkeypatch):
    # the lowering sets the flag from 11776 on, where the product pays; the kernels take it at
    # any order that is a multiple of 128
    import cola_amd.ops as O
    monkeypatch.setattr(O, "SYMMETRIC_MIN_ORDER", 0)


def rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).norm() / max(b.norm(), 1e-30))


def sym(n, seed, device="cpu"):
    g = torch.Generator(device=device).manual_seed(seed)
    M = torch.randn(n, n, generator=g, device=device)
    return (M + M.T) * 0.5  # exactly symmetric: the two sums are the same float


@pytest.mark.parametrize("n,B", [(128, 1), (256, 8), (384, 3), (512, 3), (640, 2), (1024, 5), (1024, 8),
                                 (1280, 16), (2304, 4), (4096, 8), (4096, 11)])
def test_sym_dense_matmat(dev, sym_everywhere, n, B):
    print(f"sym dense n={n} B={B}", flush=True)
    A = sym(n, n * 7 + B)
    X = torch.randn(n, B, generator=torch.Generator().manual_seed(B))
    op = ca.SelfAdjoint(ca.Dense(A.to(dev)))
    assert op.hip_op()[0].flags & N.CK_OP_SYMMETRIC
    out = op @ X.to(dev)
    assert rel(out, oop.dense_matmat(A.double(), X.double())) < 2e-6
    assert rel(out, ca.Dense(A.to(dev)) @ X.to(dev)) < 2e-6


def test_sym_dense_shift_scale(dev, sym_everywhere):
    n = 768
    A = sym(n, 5)
    X = torch.randn(n, 6, generator=torch.Generator().manual_seed(6))
    Ad = ca.Dense(A.to(dev))
    op = ca.PSD(2.5 * Ad + 0.1 * ca.I_like(Ad))
    got = op.hip_op()[0]
    assert got.kind == N.CK_OP_DENSE_F32 and got.flags & N.CK_OP_SYMMETRIC
    ref = oop.shifted(lambda V: 2.5 * (A.double() @ V), 0.1)(X.double())
    assert rel(op @ X.to(dev), ref) < 2e-6


def test_sym_dense_orders_off_the_strip_use_full_product(dev, sym_everywhere):
    # 1000 is not a multiple of 128: the flag is kept and the full product runs
    A = sym(1000, 9)
    X = torch.randn(1000, 3, generator=torch.Generator().manual_seed(1))
    out = ca.SelfAdjoint(ca.Dense(A.to(dev))) @ X.to(dev)
    assert rel(out, oop.dense_matmat(A.double(), X.double())) < 2e-6


def test_sym_dense_full_size(dev):
    # the learned workload's order: 16384, 8 rhs; rows checked in FP64 on a sample
    n, B = 16384, 8
    print(f"sym dense n={n} B={B}", flush=True)
    A = sym(n, 11, device=dev)
    X = torch.randn(n, B, device=dev, generator=torch.Generator(device=dev).manual_seed(12))
    assert ca.SelfAdjoint(ca.Dense(A)).hip_op()[0].flags & N.CK_OP_SYMMETRIC  # the default threshold
    sym_out = ca.SelfAdjoint(ca.Dense(A)) @ X
    full_out = ca.Dense(A) @ X
    assert rel(sym_out, full_out) < 2e-6
    rows = torch.arange(0, n, 97, device=dev)
    ref = A[rows].double() @ X.double()
    assert rel(sym_out[rows], ref) < 2e-6
    # a second call reuses the same scratch layout and gives the same bits
    assert torch.equal(ca.SelfAdjoint(ca.Dense(A)) @ X, sym_out)


@pytest.mark.parametrize("B", [1, 4, 8])
def test_sym_dense_cg(dev, sym_everywhere, B):
    # CG on an SPD matrix with and without the annotation: same iteration count (within
    # rounding), both solutions at the FP32 tolerance of the FP64 solve
    n = 1024
    g = torch.Generator().manual_seed(40 + B)
    M = torch.randn(n, n, generator=g) / n ** 0.5
    A = M @ M.T + 0.5 * torch.eye(n)
    A = (A + A.T) * 0.5
    b = torch.randn(n, B, generator=g)
    x64 = torch.linalg.solve(A.double(), b.double())
    Ad = A.to(dev)
    xs, info_s = ca.linalg.cg(ca.PSD(ca.Dense(Ad)), b.to(dev), tol=1e-6, max_iters=400, fused=False)
    xf, info_f = ca.linalg.cg(ca.Dense(Ad), b.to(dev), tol=1e-6, max_iters=400, fused=False)
    assert abs(info_s["iterations"] - info_f["iterations"]) <= 2
    assert rel(xs, x64) < 2e-5 and rel(xf, x64) < 2e-5


def test_sym_flag_threshold(dev):
    A = sym(1024, 3).to(dev)
    assert not ca.SelfAdjoint(ca.Dense(A)).hip_op()[0].flags & N.CK_OP_SYMMETRIC
    assert not ca.Dense(A).hip_op()[0].flags & N.CK_OP_SYMMETRIC
