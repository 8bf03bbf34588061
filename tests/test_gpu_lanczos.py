"""HipLanczos / HipArnoldi against the oracle's restatements of cola's lanczos_fact
(lanczos.py:235-296) and arnoldi_fact (arnoldi.py:289-335), the reference's
hand-computed cases, and closed forms."""
import numpy as np
import pytest
import torch

import cola_amd as ca
from cola_amd.linalg import HipArnoldi, HipLanczos, eig, lanczos
from cola_amd.linalg.arnoldi import run_arnoldi
from cola_amd.linalg.lanczos import run_lanczos
from oracle import arnoldi as oar
from oracle import lanczos as olz
from oracle import operators as oop

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / max(b.norm(), 1e-30))


def spd(diag, seed):
    rs = np.random.RandomState(seed)
    Q, _ = np.linalg.qr(rs.normal(size=(len(diag), len(diag))))
    return torch.from_numpy(Q.T @ np.diag(diag) @ Q).float()


@pytest.mark.parametrize("case", ["early", "3", "2"])
def test_lanczos_known_cases(dev, case):
    """tests/algorithms/test_lanczos.py:268-300: exact alpha, beta and iteration counts."""
    if case == "early":
        A, rhs, beta_s, alpha_s, idx_s = [[4., 0, 0], [0, 2, 0], [0, 0, 1]], [1., 0, 0], [4.], [], 1
    elif case == "3":
        A, rhs, beta_s, alpha_s, idx_s = [[7., 0, 1.], [0, 1, .1], [1., .1, 3]], [0., 1, 0], [1., 3, 7], [.1, 1.], 3
    else:
        A, rhs, beta_s, alpha_s, idx_s = [[1., .1, 0], [.1, 2, .1], [0, .1, 4]], [1., 0, 0], [1., 2, 4], [.1, .1], 3
    A = torch.tensor(A)
    op = ca.SelfAdjoint(ca.Dense(A.to(dev)))
    run = run_lanczos(op, torch.tensor(rhs, device=dev)[:, None], 3, 1e-7)
    assert run.info["iterations"] - 1 == idx_s
    assert rel(run.beta()[0], torch.tensor(beta_s)) < 1e-6
    if alpha_s:
        assert rel(run.alpha()[0], torch.tensor(alpha_s)) < 1e-6
    else:
        assert run.alpha().numel() == 0


def test_lanczos_max_eig(dev):
    """tests/algorithms/test_lanczos.py:195-217: diag(4,2,1,6), 7 start vectors of ones."""
    A = torch.diag(torch.tensor([4., 2., 1., 6.]))
    op = ca.SelfAdjoint(ca.Dense(A.to(dev)))
    Q, T, info = lanczos(op, torch.ones(4, 7, device=dev), max_iters=4, tol=1e-7)
    ev = torch.linalg.eigvalsh(ca.Tridiagonal(T.alpha[0], T.beta[0], T.alpha[0]).to_dense().cpu().double())
    assert abs(float(ev[-1]) - 6.0) < 1e-5


@pytest.mark.parametrize("n,m,nb", [(10, 10, 1), (200, 40, 1), (300, 25, 4)])
def test_lanczos_matches_oracle(dev, n, m, nb):
    diag = 1 - np.linspace(0, 1, n + 1)[:-1]**0.5 + 1e-3
    A = spd(diag, n)
    g = torch.Generator().manual_seed(n)
    S = torch.randn(n, nb, generator=g)
    Qo, ao, bo, io = olz.lanczos(lambda V: A @ V, n, S, m, 1e-7)
    op = ca.SelfAdjoint(ca.Dense(A.to(dev)))
    run = run_lanczos(op, S.to(dev), m, 1e-7)
    assert run.info["iterations"] == io["iterations"]
    k = run.iters
    assert rel(run.beta(), bo) < 5e-5
    assert rel(run.alpha(), ao) < 5e-5
    for b in range(nb):
        Q = run.basis(b).cpu().double()
        assert rel(Q, Qo[b]) < 5e-4
        assert torch.abs(Q.T @ Q - torch.eye(k, dtype=torch.float64)).max() < 1e-5


def test_lanczos_generic_operator(dev):
    A = spd(np.linspace(0.1, 3, 150), 7).to(dev)
    op = ca.SelfAdjoint(ca.LinearOperator(torch.float32, (150, 150), matmat=lambda V: A @ V, device=dev))
    S = torch.randn(150, 2)
    Qo, ao, bo, io = olz.lanczos(lambda V: A.cpu() @ V, 150, S, 30, 1e-7)
    run = run_lanczos(op, S.to(dev), 30, 1e-7)
    assert run.iters == bo.shape[1]
    assert rel(run.beta(), bo) < 5e-5 and rel(run.alpha(), ao) < 5e-5


def test_lanczos_laplacian_ritz(dev):
    """Top Ritz values of a 32x32-grid Laplacian: engine (FP64 tridiagonal solve) vs the
    oracle's FP64 Lanczos from the same start vector; error vs the closed form stated."""
    m = 32
    ip, ix, v = oop.laplacian_2d_csr(m, 0.0)
    A = ca.Sparse.from_csr(torch.from_numpy(ip).to(dev), torch.from_numpy(ix).to(dev), torch.from_numpy(v).to(dev),
                           (m * m, m * m))
    op = ca.SelfAdjoint(A + 1e-3 * ca.I_like(A))
    g = torch.Generator().manual_seed(42)
    s = torch.randn(m * m, generator=g)
    vals, vecs = eig(op, 8, "LM", HipLanczos(start_vector=s.to(dev), max_iters=60, tol=1e-7))
    dense = torch.from_numpy(np.asarray(__import__("scipy.sparse", fromlist=["csr_matrix"]).csr_matrix(
        (v.astype(np.float64), ix, ip)).todense())) + 1e-3 * torch.eye(m * m, dtype=torch.float64)
    ev64, V64, _ = olz.lanczos_eigs(lambda X: dense @ X, m * m, s.double(), 60, 1e-7, torch.float64)
    assert np.allclose(vals.cpu().double().numpy(), ev64[-8:].numpy(), rtol=1e-5)
    exact = olz.laplacian_2d_eigenvalues(m, 1e-3)
    print("top-8 ritz err vs closed form:", np.abs(vals.cpu().double().numpy() - exact[-8:]).max(), flush=True)
    Vd = vecs.to_dense().cpu().double()
    res = dense @ Vd - Vd * vals.cpu().double()[None, :]
    assert float(res.norm(dim=0).max()) < 0.2


@pytest.mark.parametrize("mgs", [True, False])
def test_arnoldi_matches_oracle(dev, mgs):
    rs = np.random.RandomState(21)
    n = 120
    A = torch.from_numpy(np.tril(rs.normal(size=(n, n))) / np.sqrt(n) + np.diag(np.linspace(1, 2, n))).float()
    s = torch.randn(n, 1, generator=torch.Generator().manual_seed(3))
    Qo, Ho, io = oar.arnoldi(lambda V: A @ V, n, s, 20, 1e-7)
    Q, H, idx, info = run_arnoldi(ca.Dense(A.to(dev)), s.to(dev), 20, 1e-7, mgs=mgs)
    assert idx == 20
    # late Hessenberg columns of any two FP32 evaluation orders drift apart ~3x per
    # step (the oracle in FP32 vs FP64 does too), so compare the first six columns
    assert rel(H[0][:7, :6], Ho[0][:7, :6]) < 1e-4
    assert rel(Q[0, :6, :n].T, Qo[0][:, :6]) < 1e-4
    # A Q_m = Q_{m+1} H (arnoldi.py invariant, tests/algorithms/test_arnoldi.py:235-243)
    Qd = Q[0, :, :n].T.cpu().double()
    assert rel(A.double() @ Qd[:, :-1], Qd @ H[0].cpu().double()) < 1e-5


def test_arnoldi_eigs_extremes_vs_oracle(dev):
    """Extreme Ritz values of Arnoldi (n-2 steps) on a normal matrix with spectrum [1, 3]
    match the oracle's and the true extremes (full-dimension FP32 MGS Arnoldi produces a
    spurious Ritz value in the reference algorithm itself, so it is not used here)."""
    n = 12
    d = np.linspace(1, 3, n)
    A = spd(d, 48)
    vals, vecs = eig(ca.Dense(A.to(dev)), n - 2, "LM", HipArnoldi(start_vector=torch.ones(n, device=dev),
                                                                   max_iters=n - 2, tol=1e-7))
    got = np.sort(vals.cpu().numpy().real)
    ev_o, _, _ = oar.arnoldi_eigs(lambda V: A @ V, n, torch.ones(n), n - 2, 1e-7)
    ref = np.sort(ev_o.real.numpy())
    # the converged extremes agree with the oracle and the true spectrum; interior Ritz
    # values after n-2 FP32 MGS steps are not converged and differ with evaluation order
    assert abs(got[0] - ref[0]) < 2e-3 and abs(got[-1] - ref[-1]) < 2e-3
    assert abs(got[-1] - 3.0) < 1e-3 and abs(got[0] - 1.0) < 1e-3


def test_lanczos_full_config_matches_golden(dev):
    """The Lanczos workload at full size (1024^2-grid Laplacian + 1e-3 I, 120 steps, seed-42 start): the
    engine's top-32 Ritz values against the oracle's FP64 Lanczos on the same start vector
    (tests/golden/full_config.json, made by tests/golden/make_full_config.py), 1e-5 relative."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "full_config.json")))
    ref = np.array(gold["lanczos_laplacian_1024sq"]["top32_ritz_fp64"])
    m = 1024
    ip, ix, v = oop.laplacian_2d_csr(m, 0.0)
    A = ca.Sparse.from_csr(torch.from_numpy(ip).to(dev), torch.from_numpy(ix).to(dev), torch.from_numpy(v).to(dev),
                           (m * m, m * m))
    op = ca.SelfAdjoint(A + 1e-3 * ca.I_like(A))
    s0 = torch.randn(m * m, generator=torch.Generator().manual_seed(42))
    vals, _ = eig(op, 32, "LM", HipLanczos(start_vector=s0.to(dev), max_iters=120, tol=1e-7))
    got = np.sort(vals.cpu().double().numpy())
    err = float(np.max(np.abs(got - ref) / np.abs(ref)))
    print(f"top-32 Ritz values vs FP64 oracle: max rel err {err:.2e}", flush=True)
    assert err < 1e-5
