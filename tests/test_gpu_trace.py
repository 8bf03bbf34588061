"""Probe streams, Hutchinson and stochastic Lanczos quadrature against the oracle
(diagonal_estimation.py:158-210, slq.py:37-75, unary.py:46-60) and closed forms."""
import ctypes

import numpy as np
import pytest
import scipy.linalg
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd.linalg import HipHutch, HipLanczos, HipSLQ, diag, hutch_trace_csr, logdet
from cola_amd.linalg._engine import probes
from cola_amd.linalg.slq import slq_quadratures
from cola_amd.linalg.unary import LanczosUnary
from oracle import hutch as ohu
from oracle import operators as oop
from oracle import rng as orng
from oracle import unary as oun

pytestmark = pytest.mark.gpu


def spd(diag, seed):
    rs = np.random.RandomState(seed)
    Q, _ = np.linalg.qr(rs.normal(size=(len(diag), len(diag))))
    return torch.from_numpy(Q.T @ np.diag(diag) @ Q).float()


@pytest.mark.parametrize("n,B,seed", [(1000, 64, 0), (777, 100, 3300923923), (5, 130, 2**63 + 5)])
def test_probe_streams_bitexact(dev, n, B, seed):
    Z, _ = probes("rademacher", n, B, seed, dev)
    assert torch.equal(Z[:, :n].cpu(), orng.rademacher(seed, n, B))
    G, _ = probes("normal", n, B, seed, dev)
    assert torch.allclose(G[:, :n].cpu().double(), orng.gaussian(seed, n, B, torch.float64), atol=2e-5)
    W = torch.empty(((B + 63) // 64, n), dtype=torch.int64, device=dev)
    N.call("ck_rng_rademacher_bits", n, (B + 63) // 64, ctypes.c_uint64(seed), N.ptr(W), N.stream_ptr(dev))
    assert np.array_equal(W.cpu().numpy().view(np.uint64), orng.rademacher_bits(seed, n, B))


def test_hutch_hadamard_exact(dev):
    """Rows of a Hadamard matrix satisfy sum_b z_b z_b^T = n I: the estimate is exact."""
    n = 64
    H = torch.from_numpy(scipy.linalg.hadamard(n)).float()
    A = torch.randn(n, n, generator=torch.Generator().manual_seed(1))
    Z = H.to(dev).contiguous()  # (B=n, n) probe block
    AZ = (A.to(dev) @ Z.T).T.contiguous()
    s = torch.zeros(n, device=dev)
    s2 = torch.zeros(n, device=dev)
    N.call("ck_hutch_accumulate", n, n, N.ptr(Z), n, N.ptr(AZ), n, N.ptr(s), N.ptr(s2), N.stream_ptr(dev))
    assert torch.allclose((s / n).cpu(), torch.diag(A), atol=1e-4)


def test_hutch_csr_packed_vs_oracle(dev):
    m = 64
    ip, ix, v = oop.laplacian_2d_csr(m, 0.0)
    n = m * m
    A = ca.Sparse.from_csr(torch.from_numpy(ip).to(dev), torch.from_numpy(ix).to(dev), torch.from_numpy(v).to(dev),
                           (n, n))
    op = A + 1e-3 * ca.I_like(A)
    tr = float(hutch_trace_csr(op, 128, seed=7))
    Z = orng.rademacher(7, n, 128, torch.float64).T
    mm = oop.shifted(lambda X: oop.csr_matmat(ip, ix, torch.from_numpy(v).double(), X), 1e-3)
    ref = float(ohu.trace_from_probes(mm, Z))
    assert abs(tr - ref) <= 1e-9 * abs(ref)
    # exact trace 4.001 * n; Rademacher error is O(sqrt(2 * sum offdiag^2 / B))
    assert abs(tr - 4.001 * n) < 6 * np.sqrt(2 * 4 * n / 128)
    # per-row sums and sums of squares of the fused kernel
    W = torch.empty((2, n), dtype=torch.int64, device=dev)
    N.call("ck_rng_rademacher_bits", n, 2, ctypes.c_uint64(7), N.ptr(W), N.stream_ptr(dev))
    s = torch.zeros(n, device=dev)
    s2 = torch.zeros(n, device=dev)
    N.call("ck_hutch_csr_bits", op.hip_op()[0], 2, N.ptr(W), N.ptr(s), N.ptr(s2), N.stream_ptr(dev))
    ds, dq = ohu.hutch_with_probes(mm, Z)
    assert torch.allclose(s.cpu().double(), ds, rtol=1e-5, atol=1e-4)
    assert torch.allclose(s2.cpu().double(), dq, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("probes", [64, 192])
def test_trace_csr_bits_group_counts(dev, probes):
    """One and several 64-probe words per row (the single-word path has its own loop)."""
    m = 40
    ip, ix, v = oop.laplacian_2d_csr(m, 0.0)
    n = m * m
    A = ca.Sparse.from_csr(torch.from_numpy(ip).to(dev), torch.from_numpy(ix).to(dev), torch.from_numpy(v).to(dev),
                           (n, n))
    tr = float(hutch_trace_csr(A, probes, seed=11))
    Z = orng.rademacher(11, n, probes, torch.float64).T
    ref = float(ohu.trace_from_probes(lambda X: oop.csr_matmat(ip, ix, torch.from_numpy(v).double(), X), Z))
    assert abs(tr - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("groups", [1, 3])
def test_trace_seeded_equals_stored_words(dev, groups):
    """Recomputing the probe words inside the pass gives the stored-word estimate bit for bit,
    on short rows (the batched path) and rows longer than 8 (the loop)."""
    import ctypes
    g = torch.Generator().manual_seed(groups)
    n = 700
    D = torch.randn(n, n, generator=g) * (torch.rand(n, n, generator=g) < 0.006)
    D[5, :40] = 1.0  # one long row
    D += torch.eye(n)
    r, c = D.nonzero(as_tuple=True)
    A = ca.Sparse(D[r, c].to(dev), r.to(dev), c.to(dev), (n, n))
    op, keep = (A + 0.5 * ca.I_like(A)).hip_op()
    W = torch.empty((groups, n), dtype=torch.int64, device=dev)
    s = N.stream_ptr(dev)
    N.call("ck_rng_rademacher_bits", n, groups, ctypes.c_uint64(123), N.ptr(W), s)
    wb = N.q("ck_trace_workspace_bytes", n)
    work = torch.empty(wb // 8 + 1, dtype=torch.float64, device=dev)
    out1 = torch.empty(1, dtype=torch.float64, device=dev)
    out2 = torch.empty(1, dtype=torch.float64, device=dev)
    N.call("ck_trace_csr_rademacher_bits", op, groups, N.ptr(W), N.ptr(out1), N.ptr(work), wb, s)
    N.call("ck_trace_csr_rademacher_seeded", op, groups, ctypes.c_uint64(123), N.ptr(out2), N.ptr(work), wb, s)
    assert float(out1) == float(out2)
    Z = orng.rademacher(123, n, 64 * groups, torch.float64).T
    ref = float(ohu.trace_from_probes(lambda X: D.double() @ X + 0.5 * X, Z))
    assert abs(float(out2) - ref) <= 1e-9 * abs(ref)


def test_hipHutch_matches_oracle_with_same_probes(dev):
    n = 300
    A = spd(np.linspace(0.5, 4, n), 11)
    op = ca.Dense(A.to(dev))
    d, info = ca.linalg.hutchinson_diag_estimate(op, 0, tol=5e-2, rand="rademacher")
    probe_fn = lambda key, n_, bs: orng.rademacher(key, n_, bs).T  # noqa: E731
    do, io = ohu.hutchinson_diag_estimate(lambda V: A @ V, n, k=0, tol=5e-2, rand="rademacher", probe_fn=probe_fn)
    assert info["iterations"] == io["iterations"]
    assert torch.allclose(d.cpu(), do, rtol=1e-4, atol=1e-4)


def test_lanczos_unary_matches_oracle(dev):
    n = 60
    A = spd(np.linspace(0.2, 5, n), 5)
    V = torch.randn(n, 3, generator=torch.Generator().manual_seed(9))
    op = ca.SelfAdjoint(ca.Dense(A.to(dev)))
    got = LanczosUnary(op, torch.log, max_iters=n, tol=1e-7) @ V.to(dev)
    ref, _ = oun.lanczos_unary_matmat(lambda X: A.double() @ X, n, torch.log, V.double(), n, 1e-7, torch.float64,
                                      torch.float64)
    assert float((got.cpu().double() - ref).norm() / ref.norm()) < 2e-4


def test_slq_exact_with_unit_probes(dev):
    """With probes sqrt(n) e_i and full Lanczos, n * mean_i tau^2 log(theta) = sum log(lambda)
    (tests/algorithms/test_stochastic_lanczos_quad.py:48-64 spectrum)."""
    n = 10
    diag = 1 - np.linspace(0, 1, n + 1)[:-1]**0.5 + 0.05
    A = spd(diag, 4)
    op = ca.PSD(ca.Dense(A.to(dev)))
    Z = (np.sqrt(n) * torch.eye(n)).to(dev)
    sc = torch.full((n, ), float(n), dtype=torch.float64, device=dev)
    q, _ = slq_quadratures(op, Z, n, 1e-10, 0, None, 1, sc)
    assert abs(float(q.mean()) - float(np.sum(np.log(diag)))) < 1e-4


def test_slq_kron_vs_oracle(dev):
    """SLQ log-determinant of K1 x K2 x K3 + 0.01 I (16^3) with 32 Rademacher probes x 30
    steps: engine vs the oracle on identical probes; closed form from factor spectra."""
    g = torch.Generator().manual_seed(0)
    facs = []
    for _ in range(3):
        t = torch.sort(torch.rand(16, generator=g)).values
        facs.append(torch.exp(-(t[:, None] - t[None, :])**2 / (2 * 0.1**2)))
    K = ca.Kronecker(*[ca.Dense(f.to(dev)) for f in facs])
    op = ca.PSD(K + 0.01 * ca.I_like(K))
    alg = HipSLQ(num_probes=32, max_iters=30, key=5)
    est = float(logdet(op, alg))
    n = 16**3
    Z = orng.rademacher(5, n, 32, torch.float64).T
    mm = lambda X: oop.kron_matmat([f.double() for f in facs], X) + 0.01 * X  # noqa: E731
    from oracle.lanczos import lanczos as olz_l
    _, al, be, _ = olz_l(mm, n, Z, 30, 1e-6, torch.float64)
    q, _, _ = oun.slq_quadratic_forms(al, be, (Z**2).sum(0), torch.log)
    ref = float(q.mean())
    # cola's logdet returns |tr log A| (logdet.py:111-117, logdet.py:48): det < 1 here
    assert abs(est - abs(ref)) < 1e-4 * abs(ref)
    lam = [torch.linalg.eigvalsh(f.double()) for f in facs]
    exact = float(torch.log((lam[0][:, None, None] * lam[1][None, :, None] * lam[2][None, None, :]) + 0.01).sum())
    print("slq vs exact rel err", abs(est - abs(exact)) / abs(exact), flush=True)
    assert abs(est - abs(exact)) < 0.05 * abs(exact)


def test_logdet_lanczos_hutch(dev):
    n = 50
    diag = np.linspace(0.5, 3, n)
    A = spd(diag, 12)
    op = ca.PSD(ca.Dense(A.to(dev)))
    ld = float(logdet(op, HipLanczos(max_iters=n, tol=1e-7), HipHutch(tol=1e-2, rand="rademacher")))
    assert abs(ld - np.sum(np.log(diag))) < 0.1 * abs(np.sum(np.log(diag)))


def test_diag_structural(dev):
    d = torch.arange(1., 5.)
    K = ca.Kronecker(ca.Diagonal(d.to(dev)), ca.Dense(torch.diag(d).to(dev)))
    assert torch.allclose(diag(K, 0).cpu(), torch.kron(d, d))


def test_slq_kron64_full_config_matches_golden(dev):
    """The SLQ workload at full size (64^3 Kronecker + 0.01 I, 32 Rademacher probes of key 5, 30 steps):
    the engine's SLQ log-determinant against the oracle's FP64 value on the same probes
    (tests/golden/full_config.json, made by tests/golden/make_full_config.py), 1e-4 relative."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "full_config.json")))["slq_kron64"]
    g = torch.Generator().manual_seed(0)
    facs = []
    for _ in range(3):
        t = torch.sort(torch.rand(64, generator=g)).values
        facs.append(torch.exp(-(t[:, None] - t[None, :])**2 / (2 * 0.1**2)))
    K = ca.Kronecker(*[ca.Dense(f.to(dev)) for f in facs])
    op = ca.PSD(K + 0.01 * ca.I_like(K))
    est = float(logdet(op, HipSLQ(num_probes=32, max_iters=30, key=5)))
    ref = gold["logdet_fp64"]
    err = abs(est - abs(ref)) / abs(ref)  # cola's logdet returns |tr log A| (logdet.py:111-117)
    print(f"SLQ logdet {est:.6f} vs FP64 oracle {ref:.6f}: rel err {err:.2e}", flush=True)
    assert err < 1e-4


@pytest.mark.parametrize("k", [1, -1, 6, -9])
def test_hutch_kth_diagonal_matches_oracle_with_same_probes(dev, k):
    """diagonal_estimation.py:186-195 for k != 0: estimator (A z)_j z_{j+k} on the engine against
    the oracle restatement with the engine's probe stream, plus the reference's errors length."""
    n = 300
    A = spd(np.linspace(0.5, 4, n), 12)
    op = ca.Dense(A.to(dev))
    d, info = ca.linalg.hutchinson_diag_estimate(op, k, tol=5e-2, rand="rademacher")
    probe_fn = lambda key, n_, bs: orng.rademacher(key, n_, bs).T  # noqa: E731
    do, io = ohu.hutchinson_diag_estimate(lambda V: A @ V, n, k=k, tol=5e-2, rand="rademacher", probe_fn=probe_fn)
    assert d.shape == (n - abs(k),)
    assert info["iterations"] == io["iterations"]
    assert len(info["errors"]) == len(io["errors"])
    assert torch.allclose(d.cpu(), do, rtol=1e-4, atol=1e-4)


def test_hutch_kth_diagonal_exact_through_diag(dev):
    """diag(A, k, Exact()) returns A[j, j+k] for every k (exact_diag, diagonal_estimation.py:117-128)."""
    n = 150
    A = torch.randn(n, n, generator=torch.Generator().manual_seed(4))
    op = ca.LinearOperator(torch.float32, (n, n), matmat=lambda V: A.to(dev) @ V, device=dev)
    for k in (0, 2, -3, 149, -149):
        got = diag(op, k, ca.linalg.Exact())
        assert torch.allclose(got.cpu(), torch.diagonal(A, k), atol=1e-5), k


def test_hutch_stop_rule_block_count_from_known_variance(dev):
    """Rows of I + c (S + S^T) (S the cyclic shift) have Rademacher estimator variance 2c^2;
    2c^2 = 0.765 makes the mean relative stderr cross tol = 3e-2 between 8 and 9 blocks of 100
    probes, so the adaptive stop (diagonal_estimation.py:197-208) runs exactly 9 blocks: a
    stderr formula with a wrong factor stops elsewhere."""
    n = 1000
    c = float(np.sqrt(0.765 / 2))
    A = torch.eye(n)
    idx = torch.arange(n)
    A[idx, (idx + 1) % n] = c
    A[idx, (idx - 1) % n] = c
    d, info = ca.linalg.hutchinson_diag_estimate(ca.Dense(A.to(dev)), 0, tol=3e-2, rand="rademacher")
    assert info["iterations"] == 10
    assert len(info["errors"]) == 9 and info["errors"][-1] == info["errors"][-2]
    assert float((d.cpu() - 1).abs().max()) < 6 * np.sqrt(0.765 / 900)
