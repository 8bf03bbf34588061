"""HipGMRES (cola/linalg/inverse/gmres.py:11-124) on the engine's Arnoldi and Givens kernels,
against the oracle's restatement (normal equations of the square Hessenberg block with the
reference's row padding), FP64 solves, the reference's known answers (tests/algorithms/
test_gmres.py) and a dense FP64 evaluation of the padded normal equations for the small
solve alone."""
import ctypes

import numpy as np
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd.linalg import GMRES, Auto, HipGMRES, gmres, inv
from cola_amd.linalg.gmres import gmres_fwd
from oracle import gmres as og

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).norm() / max(b.norm(), 1e-30))


def nonsym(n, seed, shift=3.0):
    rs = np.random.RandomState(seed)
    return torch.from_numpy(rs.normal(size=(n, n)) / np.sqrt(n) + shift * np.eye(n)).float()


def test_gmres_easy_known_answer(dev):
    """test_gmres.py:72-88: diag(3, 4, 5), rhs of ones, 3 steps, tol 1e-8 -> (1/3, 1/4, 1/5) within 1e-6;
    the same through inv(A, GMRES); inv(A, Auto) takes GMRES for an operator that is not PSD and
    has more than 1e6 entries (a dense LU below that, inv.py:73-92)."""
    A = ca.Dense(torch.diag(torch.tensor([3., 4., 5.])).to(dev))
    rhs = torch.ones(3, 1, device=dev)
    x, info = gmres(A, rhs, torch.zeros_like(rhs), 3, 1e-8)
    soln = torch.tensor([[1 / 3], [1 / 4], [1 / 5]])
    assert rel(x, soln) < 1e-6
    assert rel(inv(A, GMRES(max_iters=3, tol=1e-8)) @ rhs, soln) < 1e-6
    assert rel(inv(A, Auto()) @ rhs, soln) < 1e-6
    big = ca.Dense(nonsym(1001, 0, 4.0).to(dev))
    assert isinstance(inv(big, Auto()).alg, HipGMRES)


def test_gmres_random_spectrum_known_answer(dev):
    """test_gmres.py:56-69: spectrum 1 - x^0.5 on 25 points (PD from a seeded orthogonal basis),
    max_iters = n - 5, tol 1e-8, three rhs of ones: relative error < 5e-4."""
    x_ = np.linspace(0, 1, num=26)[:-1]
    diag = (1 - x_**0.5).astype(np.float32)
    rs = np.random.RandomState(7)
    Q, _ = np.linalg.qr(rs.normal(size=(25, 25)))
    A = torch.from_numpy(Q.T @ np.diag(diag) @ Q).float()
    rhs = torch.ones(25, 3)
    soln = torch.linalg.solve(A.double(), rhs.double())
    approx = inv(ca.Dense(A.to(dev)), GMRES(max_iters=20, tol=1e-8)) @ rhs.to(dev)
    assert rel(approx, soln) < 5e-4


@pytest.mark.parametrize("n,m,nb,mgs", [(300, 12, 1, True), (300, 12, 3, True), (64, 64, 1, True)])
def test_gmres_matches_oracle(dev, n, m, nb, mgs):
    """A non-symmetric, well-conditioned system with the reference's MGS Arnoldi, in the range of steps
    where FP32 MGS keeps the basis orthogonal: the engine against the oracle's FP32 restatement and the
    FP64 solve.  (Past that range any two FP32 evaluation orders of MGS drift apart in the late
    Hessenberg columns, the reference's own included.)"""
    A = nonsym(n, n + m)
    g = torch.Generator().manual_seed(nb)
    b = torch.randn(n, nb, generator=g)
    xo, io = og.gmres(lambda V: A @ V, n, b, None, m, 1e-7)
    x, info = gmres(ca.Dense(A.to(dev)), b.to(dev), None, m, 1e-7, mgs=mgs)
    xd = torch.linalg.solve(A.double(), b.double())
    res = float((A.double() @ x.double().cpu() - b.double()).norm() / b.double().norm())
    res_o = float((A.double() @ xo.double() - b.double()).norm() / b.double().norm())
    print(n, m, nb, mgs, "rel to oracle", rel(x, xo), "res", res, "oracle res", res_o, flush=True)
    assert rel(x, xo) < 1e-4
    assert res < 2 * res_o + 1e-6
    assert rel(x, xd) < 10 * rel(xo, xd) + 1e-5


@pytest.mark.parametrize("n,m,nb", [(1000, 60, 2), (300, 40, 3)])
def test_gmres_cgs2_matches_fp64_oracle(dev, n, m, nb):
    """Longer runs with the engine's CGS2 Arnoldi (two classical passes keep the FP32 basis orthogonal):
    against the same restatement evaluated in FP64, which is the FOM solution the reference's formula
    defines in exact arithmetic."""
    A = nonsym(n, n + m)
    g = torch.Generator().manual_seed(nb)
    b = torch.randn(n, nb, generator=g)
    A64 = A.double()
    x64, _ = og.gmres(lambda V: A64 @ V, n, b.double(), None, m, 1e-7, torch.float64)
    x, info = gmres(ca.Dense(A.to(dev)), b.to(dev), None, m, 1e-7, mgs=False)
    res = float((A64 @ x.double().cpu() - b.double()).norm() / b.double().norm())
    res64 = float((A64 @ x64 - b.double()).norm() / b.double().norm())
    print(n, m, nb, "rel to fp64 oracle", rel(x, x64), "res", res, "fp64 oracle res", res64, flush=True)
    assert rel(x, x64) < 1e-4
    assert res < 10 * res64 + 1e-6


def test_gmres_warm_start(dev):
    """gmres.py:88-124: r0 = b - A x0 starts Arnoldi and x = x0 + Q y."""
    n, m = 200, 12
    A = nonsym(n, 5)
    g = torch.Generator().manual_seed(9)
    b = torch.randn(n, 2, generator=g)
    x0 = 0.1 * torch.randn(n, 2, generator=g)
    xo, _ = og.gmres(lambda V: A @ V, n, b, x0, m, 1e-7)
    x, _ = gmres(ca.Dense(A.to(dev)), b.to(dev), x0.to(dev), m, 1e-7)
    assert rel(x, xo) < 1e-4


def _arnoldi_hessenberg(n, m, seed):
    """H (m+1 x m) of an FP64 Arnoldi run with full re-orthogonalisation on 3 I + N(0,1)/sqrt(n)."""
    rs = np.random.RandomState(seed)
    A = rs.normal(size=(n, n)) / np.sqrt(n) + 3 * np.eye(n)
    Q = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    v = rs.normal(size=n)
    Q[:, 0] = v / np.linalg.norm(v)
    for j in range(m):
        w = A @ Q[:, j]
        for _ in range(2):
            h = Q[:, :j + 1].T @ w
            w -= Q[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        Q[:, j + 1] = w / H[j + 1, j]
    return H


@pytest.mark.parametrize("m,cols,padded_rows", [(30, 30, []), (30, 12, []), (40, 40, [7, 19]), (200, 200, [])])
def test_hessenberg_lstsq_equals_padded_normal_equations(dev, m, cols, padded_rows):
    """ck_hessenberg_lstsq_f64 against the reference's padded normal equations (gmres.py:108-121), here
    solved as the equivalent augmented least-squares problem by LAPACK in FP64: full runs, a run that
    stopped at step 12 (tail rows and columns zero, the last norm tiny), and interior rows small enough
    to be padded (their y entries zero)."""
    H = _arnoldi_hessenberg(300, m, m + cols).astype(np.float32)
    for i in padded_rows:
        H[i, :] *= 1e-9
    if cols < m:
        H[cols, cols - 1] = 1e-9  # the last norm of a run that stopped early
        H[cols + 1:, :] = 0
        H[:, cols:] = 0
    rows = m if cols == m else cols + 1
    beta = np.float32(1.7)
    tol = 1e-7
    Hd = torch.from_numpy(H).to(dev).contiguous()
    bd = torch.tensor([float(beta)], dtype=torch.float32, device=dev)
    y = torch.zeros((1, cols), dtype=torch.float64, device=dev)
    wb = N.q("ck_hessenberg_lstsq_workspace_bytes", 1, rows, cols)
    work = torch.empty(wb // 8 + 1, dtype=torch.float64, device=dev)
    N.call("ck_hessenberg_lstsq_f64", 1, m, N.ptr(Hd), (m + 1) * m, rows, cols, N.ptr(bd), 10 * tol, N.ptr(work),
           work.numel() * 8, N.ptr(y), cols, N.stream_ptr(dev))
    # the reference's system: square block, padded rows where the row max < 10 tol max|H|
    Hs = H[:m, :m].astype(np.float64)
    pad = np.abs(Hs).max(axis=1) < 10 * tol * np.abs(Hs).max()
    aug = np.vstack([Hs, np.eye(m)[pad]])
    rhs = np.zeros(aug.shape[0])
    rhs[0] = float(beta)
    ref = np.linalg.lstsq(aug, rhs, rcond=None)[0]
    ref[pad] = 0.0
    got = y[0].cpu().numpy()
    assert np.linalg.norm(got - ref[:cols]) / np.linalg.norm(ref) < 1e-10
    for i in padded_rows:
        assert got[i] == 0.0
    # and the normal equations as the reference writes them agree to their own conditioning
    G = Hs.T @ Hs + np.diag(pad.astype(np.float64))
    ne = np.linalg.solve(G, Hs.T[:, 0] * float(beta))
    ne[pad] = 0.0
    assert np.linalg.norm(got - ne[:cols]) / np.linalg.norm(ne) < 1e-7


def test_gmres_vjp_known_answer(dev):
    """test_gmres.py:11-53: A = diag(theta), theta = (3, 4, 5), rhs of ones, 5 steps at tol 1e-6:
    the gradient of sum(A^-1 1) is -1/theta^2 (gmres_bwd, gmres.py:72-82)."""
    theta = torch.tensor([3., 4., 5.], device=dev, requires_grad=True)
    A = ca.Diagonal(theta)
    rhs = torch.ones(3, 1, device=dev)
    soln, _ = gmres_fwd(A, rhs, torch.zeros_like(rhs), 5, 1e-6, None)
    soln.sum().backward()
    assert rel(theta.grad, -1.0 / torch.tensor([3., 4., 5.], dtype=torch.float64)**2) < 1e-6


_ = ctypes
