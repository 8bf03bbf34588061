"""Backward rules of the iterative solvers on the engine: cg_bwd (cola/linalg/inverse/cg.py:72-86)
and slq_bwd (cola/linalg/tbd/slq.py:10-31) through iterative_autograd
(cola/utils/custom_autodiff.py:4-86).  Pinned by the reference's known answer
(tests/algorithms/test_cg.py:20-62: d/dtheta sum(diag(theta)^-1 1) = -1/theta^2), by FP64
torch autograd of a dense solve, and, for the stochastic log-determinant gradient, by the
same estimator evaluated in FP64 on the host with the same probes plus a case where the
estimator is exact."""
import numpy as np
import pytest
import torch

import cola_amd as ca
from cola_amd.linalg import HipCG, HipSLQ, inv, slogdet, stochastic_lanczos_quad
from cola_amd.linalg.cg import run_cg
from cola_amd.linalg.keys import PRNGKey
from oracle import rng as orng

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).norm() / max(b.norm(), 1e-30))


@pytest.mark.parametrize("form,fused", [("diagonal", False), ("dense", True), ("dense", False)])
def test_cg_vjp_diag_known_answer(dev, form, fused):
    """test_cg.py:20-62: A = diag(theta), theta = (3, 4, 5), one rhs of ones, 5 steps at tol 1e-6:
    the gradient of sum(A^-1 1) is -1/theta^2 (A as a Diagonal operator, and stored dense so
    that both the persistent and the chained CG run the forward and backward solves)."""
    theta = torch.tensor([3., 4., 5.], device=dev, requires_grad=True)
    A = ca.Diagonal(theta) if form == "diagonal" else ca.Dense(torch.diag(theta))
    ones = torch.ones(3, 1, device=dev)
    x0 = torch.zeros_like(ones)
    soln, *_ = run_cg(A, ones, x0, 5, 1e-6, None, 16, fused)
    soln.sum().backward()
    expect = -1.0 / torch.tensor([3., 4., 5.], dtype=torch.float64) ** 2
    assert rel(theta.grad, expect) < 1e-6


def test_inv_dense_gradient_matches_fp64_autograd(dev):
    """x = inv(PSD(Dense(M))) @ b through HipCG, loss = <w, x>: the gradient with respect to M
    (-A^-1 w x^T) and to a ScalarMul shift c (-(A^-1 w)^T x) against FP64 autograd of a dense solve."""
    g = torch.Generator().manual_seed(3)
    n = 300
    Q = torch.randn(n, n // 2, generator=g)
    M0 = Q @ Q.T / n
    b = torch.randn(n, 2, generator=g)
    w = torch.randn(n, 2, generator=g)
    M = M0.clone().to(dev).requires_grad_(True)
    c = torch.tensor(0.2, device=dev, requires_grad=True)
    A = ca.PSD(ca.Dense(M) + ca.ScalarMul(c, (n, n), device=dev))
    x = inv(A, HipCG(tol=1e-7, max_iters=2000)) @ b.to(dev)
    (x * w.to(dev)).sum().backward()
    M64 = M0.double().requires_grad_(True)
    c64 = torch.tensor(0.2, dtype=torch.float64, requires_grad=True)
    x64 = torch.linalg.solve(M64 + c64 * torch.eye(n, dtype=torch.float64), b.double())
    (x64 * w.double()).sum().backward()
    print("dM rel", rel(M.grad, M64.grad), "dc rel", rel(c.grad, c64.grad), flush=True)
    assert rel(x.detach(), x64.detach()) < 1e-5
    assert rel(M.grad, M64.grad) < 1e-4
    assert rel(c.grad, c64.grad) < 1e-4


def test_no_grad_path_is_plain_solve(dev):
    """Without tensors that require grad, cg() runs the solver directly (no autograd node)."""
    A = ca.PSD(ca.Diagonal(torch.tensor([2., 4.], device=dev)))
    x = inv(A, HipCG(tol=1e-7, max_iters=10)) @ torch.ones(2, device=dev)
    assert not x.requires_grad
    assert rel(x, torch.tensor([0.5, 0.25])) < 1e-6


def test_slq_logdet_gradient_exact_for_diagonal(dev):
    """HipSLQ with Rademacher probes on a diagonal matrix stored dense: A^-1 z = z / theta and
    z_i^2 = 1, so (1/m) sum_b (A^-1 z_b)^T dA z_b is exactly 1/theta_i, the gradient of the
    log-determinant."""
    theta = torch.linspace(0.5, 3.0, 200, device=dev).requires_grad_(True)
    A = ca.PSD(ca.Dense(torch.diag(theta)))
    _, ld = slogdet(A, HipSLQ(num_probes=16, max_iters=30))
    ld.backward()
    t = torch.linspace(0.5, 3.0, 200, dtype=torch.float64)
    assert abs(float(ld) - float(torch.log(t).sum())) / float(torch.log(t).sum().abs()) < 1e-4
    assert rel(theta.grad, 1.0 / t) < 1e-5


def test_stochastic_lanczos_quad_backward_matches_host_estimator(dev):
    """slq_bwd (slq.py:10-31): m Gaussian probes from PRNGKey(0), CG solves at tol 1e-6 (100 steps),
    dA = (g/m) sum_j (A^-1 z_j) z_j^T; against the same formula in FP64 on the host with the
    same probes (the engine's counter stream restated in the oracle), and within the estimator's
    spread of the exact gradient A^-1."""
    g = torch.Generator().manual_seed(5)
    n = 256
    Q = torch.randn(n, n, generator=g)
    M0 = Q @ Q.T / n + 0.5 * torch.eye(n)
    M = M0.clone().to(dev).requires_grad_(True)
    vtol = 0.125  # int(1/vtol^2) = 64 probes
    est = stochastic_lanczos_quad(ca.PSD(ca.Dense(M)), "log", max_iters=60, tol=1e-6, vtol=vtol)
    est.backward()
    m = int(1 / vtol**2)
    Z = orng.gaussian(PRNGKey(0), n, m, torch.float64).T  # (n, m)
    A64 = M0.double()
    host = torch.linalg.solve(A64, Z) @ Z.T / m
    print("slq grad rel to host estimator", rel(M.grad, host), flush=True)
    assert rel(M.grad, host) < 1e-4
    # the matrix estimate is noisy entrywise (relative spread ~ sqrt(n/m)); its trace estimates
    # tr(A^-1) = d logdet(A + t I)/dt with relative spread ~ sqrt(2/m) ||A^-1||_F / tr(A^-1)
    exact = torch.linalg.inv(A64)
    spread = np.sqrt(2.0 / m) * float(exact.norm()) / float(exact.trace())
    assert abs(float(M.grad.double().trace().cpu()) - float(exact.trace())) / float(exact.trace()) < 4 * spread
    ld64 = float(torch.linalg.slogdet(A64)[1])
    assert abs(float(est) - ld64) / abs(ld64) < 0.1
