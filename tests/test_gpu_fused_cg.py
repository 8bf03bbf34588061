"""The persistent fused CG (ck_cg_fused: one cooperative launch per solve) against the
oracle's restatement of cola's run_batched_cg (cg.py:94-178), against the chained-launch
engine path on the same inputs, and its status codes and bounded grid-wide waits.
Tolerances as in test_gpu_cg: FP32 iterates agree with the FP32 oracle to rounding, step
counts within +-2 (reduction order)."""
import numpy as np
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd.linalg import HipCG, inv, run_batched_cg
from cola_amd.linalg.cg import fused_solve
from oracle import cg as ocg

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).norm() / max(b.norm(), 1e-30))


def system(n, rank, shift, B, seed):
    g = torch.Generator().manual_seed(seed)
    M = torch.randn(n, rank, generator=g)
    A = M @ M.T / rank + shift * torch.eye(n)
    b = torch.randn(n, B, generator=g)
    return A, b


def test_fused_config1_one_launch(dev):
    """A = X X^T/256 + 0.1 I, N = 1024, one rhs (seed 0): the persistent solve matches the
    oracle's step count, iterate and residual history, and is the path inv() takes."""
    g = torch.Generator().manual_seed(0)
    X = torch.randn(1024, 256, generator=g)
    b = torch.randn(1024, generator=g)
    A = X @ X.T / 256 + 0.1 * torch.eye(1024)
    xo, io = ocg.cg(lambda V: A @ V, b[:, None], tol=1e-6, max_iters=1000)
    solver = inv(ca.PSD(ca.Dense(A.to(dev))), HipCG(tol=1e-6, max_iters=1000))
    x = solver @ b.to(dev)
    info = solver.info
    print("fused iters", info["iterations"], "oracle", io["iterations"], info["status"], flush=True)
    assert info["fused"] and info["status"] == "ok"
    assert abs(info["iterations"] - io["iterations"]) <= 2
    assert rel(x, xo[:, 0]) < 1e-5
    res = (A.double() @ x.double().cpu() - b.double()).norm() / b.double().norm()
    assert res < 1e-5
    n = min(len(info["errors"]), len(io["errors"]))
    ge, oe = np.log10(info["errors"][:n]), np.log10(io["errors"][:n])
    above = oe > -4.5
    assert np.allclose(ge[above], oe[above], atol=0.02)


# (n, B, nblocks): LDS-resident rows (small n) and rows streamed from global memory (n = 4096),
# n not a multiple of 4, tiny n, one workgroup
@pytest.mark.parametrize("n,B,nblocks", [(1024, 1, 0), (1024, 1, 64), (1024, 1, 256), (1000, 3, 0), (513, 4, 0),
                                         (4096, 1, 0), (2048, 2, 0), (7, 2, 0), (64, 1, 1)])
def test_fused_matches_oracle_and_chained(dev, n, B, nblocks):
    A, b = system(n, max(2, n // 2), 0.05, B, n + B)
    xo, ro, ko, io = ocg.run_batched_cg(lambda V: A @ V, b, None, 600, 1e-6, None)
    # the same restatement in FP64: these systems end with the residual creeping along the threshold,
    # where FP32 summation order alone moves the FP32 oracle's stopping step by up to ~5 (it differs
    # between host BLAS builds); the engine's FP64-accumulated reductions track the FP64 count
    A64, b64 = A.double(), b.double()
    _, _, k64, _ = ocg.run_batched_cg(lambda V: A64 @ V, b64, None, 600, 1e-6, None)
    out = fused_solve(ca.Dense(A.to(dev)), b.to(dev), None, 600, 1e-6, None, nblocks=nblocks)
    assert out is not None
    x, r, k, info = out
    print(n, B, nblocks, "fused", k, "oracle", ko, "fp64 oracle", k64, info["status"], flush=True)
    assert abs(k - ko) <= 2 or abs(k - k64) <= 2
    assert rel(x, xo) < 1e-5
    # r is the un-normalised recursive residual (cg.py:119): below the stopping threshold
    # (tol*||r0|| + tol on the normalised system) times ||b|| in every column
    assert torch.all(r.cpu().norm(dim=0) <= 2.0e-6 * b.norm(dim=0) * 1.01)
    xc, rc, kc, ic = run_batched_cg(ca.Dense(A.to(dev)), b.to(dev), None, 600, 1e-6, None, fused=False)
    assert not ic["fused"]
    assert abs(k - kc) <= max(2, k // 10) and rel(x, xc) < 1e-5


def test_fused_scale_shift_diag_preconditioner_warm_start(dev):
    """y = 2 A x + 0.3 x with a diagonal preconditioner and an x0: the descriptor's scale and
    shift, z = P r and r0 = b/||b|| - A x0 (cg.py:122-130) against the oracle."""
    n, B = 700, 4
    A, b = system(n, 300, 0.0, B, 11)
    d = torch.linspace(0.5, 2.0, n)
    A = A * d[:, None] * d[None, :]
    Aeff = 2.0 * A + 0.3 * torch.eye(n)
    g = torch.Generator().manual_seed(12)
    x0 = 0.01 * torch.randn(n, B, generator=g)
    Pd = 1.0 / torch.diag(Aeff)
    xo, ro, ko, io = ocg.run_batched_cg(lambda V: Aeff @ V, b, x0, 800, 1e-6, lambda V: Pd[:, None] * V)
    op = ca.Dense(A.to(dev)) * 2.0 + 0.3 * ca.I_like(ca.Dense(A.to(dev)))
    x, r, k, info = run_batched_cg(op, b.to(dev), x0.to(dev), 800, 1e-6, ca.Diagonal(Pd.to(dev)))
    print("fused pcg", k, "oracle", ko, flush=True)
    assert info["fused"]
    assert abs(k - ko) <= max(2, ko // 50)
    assert rel(x, xo) < 1e-5


def test_fused_wait_times_out_and_recovers(dev):
    """Every grid-wide wait is bounded: a host-set generation offset that no workgroup ever
    publishes makes each wait give up after 1000 polling rounds; the kernel ends, the
    status is CK_ERR_TIMEOUT, and the next solve on the device is correct."""
    A, b = system(512, 256, 0.1, 2, 5)
    op = ca.Dense(A.to(dev))
    with pytest.raises(N.KrylovError) as e:
        fused_solve(op, b.to(dev), None, 100, 1e-6, None, spin_limit=1000, wait_offset=1 << 30)
    assert e.value.status == N.CK_ERR_TIMEOUT
    x, r, k, info = fused_solve(op, b.to(dev), None, 500, 1e-6, None)
    assert info["status"] == "ok"
    assert rel(x, torch.linalg.solve(A.double(), b.double())) < 2e-5


@pytest.mark.parametrize("fused", [True, False])
def test_status_codes(dev, fused):
    """ck_cg_ctrl.status on both engine paths: zero rhs -> ok (x = 0), step cap -> maxiter,
    NaN in the rhs -> nan (the loop stops as in the reference), an indefinite operator ->
    breakdown (the reference's safe division continues; the status records it)."""
    A, b = system(200, 100, 0.1, 2, 21)
    op = ca.Dense(A.to(dev))
    z = torch.zeros(200, 2, device=dev)
    x, r, k, info = run_batched_cg(op, z, None, 50, 1e-6, None, fused=fused)
    assert info["fused"] == fused
    assert info["status"] == "ok" and float(x.abs().max()) == 0.0
    x, r, k, info = run_batched_cg(op, b.to(dev), None, 1, 1e-6, None, fused=fused)
    assert info["status"] == "maxiter" and k == 1 and info["iterations"] == 2
    bn = b.clone()
    bn[3, 1] = float("nan")
    x, r, k, info = run_batched_cg(op, bn.to(dev), None, 50, 1e-6, None, fused=fused)
    assert info["status"] == "nan"
    ev = torch.linspace(-2.0, -0.5, 200)  # negative definite: p.Ap < 0 from the first step
    rs = np.random.RandomState(3)
    Q = torch.from_numpy(np.linalg.qr(rs.normal(size=(200, 200)))[0]).float()
    Ai = Q @ torch.diag(ev) @ Q.T
    x, r, k, info = run_batched_cg(ca.Dense(Ai.to(dev)), b.to(dev), None, 60, 1e-6, None, fused=fused)
    assert info["status"] == "breakdown"


def test_fused_cooperative_launch_matches_plain(dev):
    """The same solve launched cooperatively (co-residency guaranteed by the runtime) and plainly
    (after the same occupancy check): identical iterates, bit for bit (same arithmetic order)."""
    A, b = system(1024, 512, 0.05, 2, 77)
    op = ca.Dense(A.to(dev))
    x1, r1, k1, i1 = fused_solve(op, b.to(dev), None, 600, 1e-6, None, cooperative=True)
    x2, r2, k2, i2 = fused_solve(op, b.to(dev), None, 600, 1e-6, None, cooperative=False)
    assert k1 == k2 and torch.equal(x1, x2) and torch.equal(r1, r2)
