"""Pins the learned-preconditioner oracle against mathematics and library routines
(finite differences, LAPACK dpstrf, torch.optim.Adam, dense inverses, cdist)."""
import numpy as np
import pytest
import scipy.linalg
import scipy.linalg.lapack
import scipy.spatial.distance
import torch

from oracle import learned as ol


def test_mlp_grads_match_finite_differences():
    din, dout = 3, 5
    p = torch.from_numpy(ol.init_params(din, dout, 3)).double()
    g = torch.Generator().manual_seed(0)
    X = torch.rand(7, din, generator=g, dtype=torch.float64)
    dG = torch.randn(7, dout, generator=g, dtype=torch.float64)
    grad = ol.mlp_grads(p, X, dG, din, dout)
    f = lambda q: float((ol.mlp_forward(q, X, din, dout)[0] * dG).sum())  # noqa: E731
    idx = torch.randperm(p.numel(), generator=g)[:40]
    for i in idx.tolist():
        e = torch.zeros_like(p)
        e[i] = 1e-6
        fd = (f(p + e) - f(p - e)) / 2e-6
        assert abs(fd - float(grad[i])) <= 1e-6 * max(1.0, abs(fd))


def test_mlp_forward_layers():
    """A hand-built net: W1 = I-like, zero biases; gelu(0) = 0 and gelu(z) -> z for large z."""
    din, dout = 2, 2
    lay, total = ol.param_layout(din, dout)
    p = torch.zeros(total, dtype=torch.float64)
    pr = ol.unpack(p, din, dout)
    pr["W1"][0, 0] = 1.0
    pr["W1"][1, 1] = 1.0
    pr["W2"][0, 0] = 1.0
    pr["W2"][1, 1] = 1.0
    pr["W3"][0, 1] = 2.0
    pr["W3"][1, 0] = 3.0
    pr["b3"][:] = torch.tensor([0.5, -0.5])
    X = torch.tensor([[40.0, 0.0], [0.0, 50.0]], dtype=torch.float64)
    G, _, _ = ol.mlp_forward(p, X, din, dout)
    assert torch.allclose(G, torch.tensor([[0.5, 79.5], [150.5, -0.5]], dtype=torch.float64))


def test_energy_loss_limits_and_invariance():
    g = torch.Generator().manual_seed(1)
    n, m, k = 50, 4, 6
    U, _ = torch.linalg.qr(torch.randn(n, m, generator=g, dtype=torch.float64))
    w = torch.tensor([4.0, 3.0, 2.0, 1.0], dtype=torch.float64)
    R = torch.randn(m, k, generator=g, dtype=torch.float64)
    G_in = U @ R + 1e-9 * torch.randn(n, k, generator=g, dtype=torch.float64)  # spans U
    assert float(ol.energy_loss(G_in, U, w, 0.0)) < 1e-8
    Q, _ = torch.linalg.qr(torch.cat([U, torch.randn(n, k, generator=g, dtype=torch.float64)], 1))
    G_out = Q[:, m:m + k]  # orthogonal to U
    assert abs(float(ol.energy_loss(G_out, U, w, 0.0)) - 1.0) < 1e-12
    G = torch.randn(n, k, generator=g, dtype=torch.float64)
    M = torch.randn(k, k, generator=g, dtype=torch.float64) + 3 * torch.eye(k, dtype=torch.float64)
    assert abs(float(ol.energy_loss(G, U, w, 0.0)) - float(ol.energy_loss(G @ M, U, w, 0.0))) < 1e-10
    # partial capture: G = first two target vectors -> (4 + 3) / 10 captured
    assert abs(float(ol.energy_loss(U[:, :2], U, w, 0.0)) - 0.3) < 1e-12


def test_energy_grad_finite_differences():
    g = torch.Generator().manual_seed(2)
    n, m, k = 30, 3, 4
    U, _ = torch.linalg.qr(torch.randn(n, m, generator=g, dtype=torch.float64))
    w = torch.tensor([3.0, 2.0, 1.0], dtype=torch.float64)
    G = torch.randn(n, k, generator=g, dtype=torch.float64)
    L, dG = ol.energy_loss_grad(G, U, w, 1e-6)
    for (i, a) in [(0, 0), (5, 2), (29, 3), (11, 1)]:
        E = torch.zeros_like(G)
        E[i, a] = 1e-6
        fd = (float(ol.energy_loss(G + E, U, w, 1e-6)) - float(ol.energy_loss(G - E, U, w, 1e-6))) / 2e-6
        assert abs(fd - float(dG[i, a])) < 1e-6


def test_adam_matches_torch():
    g = torch.Generator().manual_seed(3)
    p0 = torch.randn(20, generator=g, dtype=torch.float64)
    grads = [torch.randn(20, generator=g, dtype=torch.float64) for _ in range(4)]
    prm = torch.nn.Parameter(p0.clone())
    opt = torch.optim.Adam([prm], lr=3e-3)
    p, m, v = p0.numpy().copy(), np.zeros(20), np.zeros(20)
    for t, gr in enumerate(grads, 1):
        prm.grad = gr.clone()
        opt.step()
        p, m, v = ol.adam_step(p, gr.numpy(), m, v, t, 3e-3)
    assert np.allclose(p, prm.detach().numpy(), rtol=0, atol=1e-14)


def test_pivoted_cholesky_vs_lapack():
    rs = np.random.RandomState(4)
    B = rs.normal(size=(40, 40))
    A = B @ B.T + np.diag(np.linspace(1, 40, 40))
    L, piv, dpiv, d = ol.pivoted_cholesky(A, 40)
    c, lpiv, rank, info = scipy.linalg.lapack.dpstrf(A, lower=1)
    assert info == 0 and rank == 40
    assert np.array_equal(piv, lpiv - 1)
    Ll = np.tril(c)
    assert np.allclose(L[piv], Ll, atol=1e-10)
    assert np.allclose(L @ L.T, A, atol=1e-9)
    assert np.all(np.abs(d) < 1e-8)
    assert piv[0] == np.argmax(np.diag(A))


def test_pivoted_cholesky_low_rank_exact():
    rs = np.random.RandomState(5)
    V = rs.normal(size=(60, 5))
    A = V @ V.T
    L, piv, dpiv, d = ol.pivoted_cholesky(A, 5)
    assert np.allclose(L @ L.T, A, atol=1e-9)
    assert np.all(d < 1e-9)
    # residual diagonal is non-increasing at every step of a longer run
    _, _, dp, _ = ol.pivoted_cholesky(A + 1e-3 * np.eye(60), 8)
    assert np.all(np.diff(dp) <= 1e-12)


def test_nystrom_and_woodbury():
    rs = np.random.RandomState(6)
    V = rs.normal(size=(80, 6))
    K = V @ V.T
    G = rs.normal(size=(80, 6))  # generic sketch spans range(K)
    F = ol.nystrom_factor(K, G, jitter_rel=0.0)
    assert np.allclose(F @ F.T, K, atol=1e-8)
    F2 = rs.normal(size=(80, 7))
    s = 0.3
    X = rs.normal(size=(80, 3))
    ref = np.linalg.solve(s * np.eye(80) + F2 @ F2.T, X)
    assert np.allclose(ol.woodbury_apply(F2, s, X), ref, atol=1e-10)


def test_rbf_dense_vs_cdist():
    rs = np.random.RandomState(7)
    P = rs.uniform(size=(30, 4))
    K = ol.rbf_dense(P, 0.3)
    ref = np.exp(-scipy.spatial.distance.cdist(P, P, "sqeuclidean") / (2 * 0.09))
    assert np.allclose(K, ref, atol=1e-14)


@pytest.mark.parametrize("din,dout", [(4, 64), (3, 40)])
def test_param_count(din, dout):
    _, total = ol.param_layout(din, dout)
    assert total == din * 128 + 128 + 128 * 128 + 128 + 128 * dout + dout
