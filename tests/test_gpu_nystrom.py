"""Nystrom preconditioner on the engine (cola/linalg/preconditioning/preconditioners.py:87-186):
the reference's known answers (tests/algorithms/test_precond.py:60-162), the oracle's
restatement on the same sketch, the device Jacobi eigensolver against LAPACK, and the
preconditioner inside the engine's CG against the oracle's CG with the oracle's preconditioner."""
import numpy as np
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd.linalg import NystromPrecond, get_nys_approx
from cola_amd.linalg import preconditioning as pre
from cola_amd.linalg.cg import cg
from oracle import cg as ocg
from oracle import nystrom as onys

pytestmark = pytest.mark.gpu


def spectrum(coeff, scale, size):
    """utils_for_tests.py:188-196 (real case)."""
    x = np.linspace(0, 1, num=size + 1)[:-1].astype(np.float32)
    return (scale * (1 - x**coeff)).astype(np.float32)


def pd_from_diag(diag, seed):
    """utils_for_tests.py:164-173."""
    rs = np.random.RandomState(seed)
    Q, _ = np.linalg.qr(rs.normal(size=(diag.shape[0], diag.shape[0])).astype(np.float32))
    return (Q.T @ np.diag(diag) @ Q).astype(np.float32)


def relative_error(v, w):
    """utils_for_tests.py:117-123."""
    v = v.double().cpu() if torch.is_tensor(v) else torch.from_numpy(np.asarray(v, np.float64))
    w = w.double().cpu() if torch.is_tensor(w) else torch.from_numpy(np.asarray(w, np.float64))
    return float((v - w).norm() / max(float((v.norm() + w.norm()) / 2.0), 1e-16))


def _sym_cases():
    rs = np.random.RandomState(0)
    out = []
    for k in (1, 2, 7, 33, 64):
        M = rs.normal(size=(k, k))
        out.append((f"random{k}", M + M.T))
    Q, _ = np.linalg.qr(rs.normal(size=(40, 40)))
    out.append(("repeated", Q @ np.diag(np.repeat([3.0, 1.0, -2.0, 0.5], 10)) @ Q.T))
    out.append(("diagonal", np.diag(rs.normal(size=17))))
    out.append(("zero", np.zeros((9, 9))))
    out.append(("graded", Q[:24, :24] @ np.diag(10.0**-np.arange(24)) @ Q[:24, :24].T))
    return out


@pytest.mark.parametrize("name,S", _sym_cases(), ids=[c[0] for c in _sym_cases()])
def test_sym_eigh_against_lapack(dev, name, S):
    """ck_sym_eigh_f64 on three copies (batched): eigenvalues against LAPACK (descending), the
    residual A E^T - E^T diag(w) and the orthonormality of E."""
    k = S.shape[0]
    nb = 3
    A = torch.from_numpy(np.stack([S] * nb)).to(dev).contiguous()
    w = torch.empty((nb, k), dtype=torch.float64, device=dev)
    E = torch.empty((nb, k, k), dtype=torch.float64, device=dev)
    sweeps = torch.zeros(nb, dtype=torch.int32, device=dev)
    N.call("ck_sym_eigh_f64", nb, k, N.ptr(A), k, k * k, N.ptr(w), k, N.ptr(E), k, k * k, 40, N.ptr(sweeps),
           N.stream_ptr(dev))
    ref = np.sort(np.linalg.eigvalsh(S))[::-1]
    scale = max(np.abs(S).max(), 1e-300)
    for b in range(nb):
        wb, Eb = w[b].cpu().numpy(), E[b].cpu().numpy()
        assert int(sweeps[b]) >= 0, "did not converge"
        assert np.abs(wb - ref).max() <= 1e-13 * scale * k
        assert np.linalg.norm(S @ Eb.T - Eb.T * wb[None, :]) <= 1e-13 * scale * k
        assert np.abs(Eb @ Eb.T - np.eye(k)).max() <= 1e-13 * k
    assert torch.equal(w[0], w[2]) and torch.equal(E[0], E[1])


def test_nys_approx_diagonal_known_answer(dev):
    """test_precond.py:145-162."""
    A = ca.Dense(torch.diag(torch.tensor([3., 4., 5., 7., 1.])).to(dev))
    Omega = torch.eye(5)[:, [0, 4, 2]].to(dev)
    Lam, U = get_nys_approx(A, Omega, eps=1e-6)
    soln = torch.diag(torch.tensor([3., 0., 5., 0., 1.]))
    assert relative_error(soln, U @ (Lam[:, None] * U.T)) < 1e-7


def test_nys_approx_full_rank_known_answer(dev):
    """test_precond.py:125-142: rank n reproduces a random PD matrix (rel < 500 * 1e-7)."""
    A = torch.from_numpy(pd_from_diag(spectrum(0.5, 1.0, 10), 3))
    Omega = torch.randn(10, 10, generator=torch.Generator().manual_seed(4))
    Lam, U = get_nys_approx(ca.Dense(A.to(dev)), Omega.to(dev), eps=1e-6)
    assert relative_error(A, U @ (Lam[:, None] * U.T)) < 500 * 1e-7


def test_nys_precond_closed_form(dev):
    """test_precond.py:91-122 through the engine's apply (P = Nys @ I on the device)."""
    diag = spectrum(0.5, 1.0, 10)
    A = torch.diag(torch.from_numpy(diag))
    rank, mu = 5, 1e-8
    eye = torch.eye(10)
    Om, Om2 = eye[:, :rank], eye[:, rank:]
    B = ca.Dense(A.to(dev))
    Nys = NystromPrecond(B, rank=rank)
    Nys._create_approx(B, Om.clone().to(dev), mu=mu, eps=1e-8, adjust_mu=False)
    P = (Nys @ eye.to(dev)).cpu()
    Psqrt = P**0.5
    approx = Psqrt @ A @ Psqrt
    d = torch.from_numpy(diag)
    soln_P = float(d[rank - 1]) * Om @ torch.diag(1 / (d[:rank] + mu)) @ Om.T + eye - Om @ Om.T
    inv = (A + mu * eye)[rank:, rank:]
    soln = (float(d[rank - 1]) + mu) * Om @ Om.T + Om2 @ (inv @ Om2.T)
    assert relative_error(soln, approx) < 1e-6
    assert relative_error(soln_P, P) < 1e-6


def test_nys_sqrt_inverse(dev):
    """test_precond.py:60-83: sqrt(P)^2 = P and inverse(P) P = I (rank 5 of a seeded PD matrix of
    order 10, the default keyed sketch)."""
    A = torch.from_numpy(pd_from_diag(spectrum(0.5, 1.0, 10), 21))
    Nys = NystromPrecond(ca.Dense(A.to(dev)), rank=5)
    eye = torch.eye(10, device=dev)
    Pd = Nys @ eye
    S = pre.sqrt(Nys) @ eye
    assert relative_error(Pd, S @ S) < 1e-6
    Inv = pre.inverse(Nys) @ eye
    assert relative_error(eye, Inv @ Pd) < 1e-6
    # the keyed sketch is the reference backend's: torch.manual_seed(PRNGKey(42)); torch.randn
    g = torch.Generator(device=dev).manual_seed(ca.linalg.PRNGKey(42))
    Om = torch.randn(10, 5, generator=g, device=dev)
    ref = onys.NystromPrecond(lambda V: A.numpy() @ V, Om.cpu().numpy())
    assert relative_error(Pd, ref.matmat(np.eye(10, dtype=np.float32))) < 1e-5


def rbf(n, ell, seed):
    X = torch.rand(n, 3, generator=torch.Generator().manual_seed(seed))
    return torch.exp(-torch.cdist(X, X)**2 / (2 * ell * ell))


@pytest.mark.parametrize("n,rank", [(1000, 32), (3000, 64), (257, 7)])
def test_nystrom_matches_oracle(dev, n, rank):
    """An RBF kernel matrix sketched by the same Gaussian Omega: the captured eigenvalues and the
    preconditioner's action against the oracle (NumPy, FP64 SVD of the FP32 factor)."""
    K = rbf(n, 0.3, n)
    Omega = torch.randn(n, rank, generator=torch.Generator().manual_seed(rank))
    P = NystromPrecond(ca.Dense(K.to(dev)), rank=rank)
    P._create_approx(ca.Dense(K.to(dev)), Omega.to(dev), mu=1e-3, eps=1e-8, adjust_mu=True)
    Kd = K.double().numpy()
    ref = onys.NystromPrecond(lambda V: Kd @ V, Omega.double().numpy(), mu=1e-3, eps=1e-8, adjust_mu=True)
    lam = P.Lambda.double().cpu().numpy()
    top = lam[:rank // 2]
    assert np.abs(top - ref.Lambda[:rank // 2]).max() / ref.Lambda[0] < 1e-5
    V = torch.randn(n, 4, generator=torch.Generator().manual_seed(1))
    got = (P @ V.to(dev)).double().cpu().numpy()
    want = ref.matmat(V.double().numpy())
    print(n, rank, "lambda", lam[:3], "rel apply diff", np.linalg.norm(got - want) / np.linalg.norm(want), flush=True)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-4


def test_nystrom_preconditioned_cg(dev):
    """P from NystromPrecond(K, 64, mu=noise) inside the engine's CG on (K + noise I): fewer than
    half the steps of plain CG, the oracle's step count (its CG with its preconditioner on the
    same sketch) within 2, and an FP64-checked residual no worse than FP32 CG reaches there."""
    n, noise = 2000, 1e-2
    K = rbf(n, 0.4, 7)
    A = ca.PSD(ca.Dense(K.to(dev)) + noise * ca.I_like(ca.Dense(K.to(dev))))
    P = NystromPrecond(ca.Dense(K.to(dev)), rank=64, mu=noise, adjust_mu=False)
    b = torch.randn(n, 2, generator=torch.Generator().manual_seed(3))
    _, plain = cg(A, b.to(dev), tol=1e-5, max_iters=3000)
    x, info = cg(A, b.to(dev), P=P, tol=1e-5, max_iters=3000)
    A64 = K.double() + float(np.float32(noise)) * torch.eye(n, dtype=torch.float64)

    def res(z):
        return float(((A64 @ z.double().cpu() - b.double()).norm(dim=0) / b.double().norm(dim=0)).max())

    g = torch.Generator(device=dev).manual_seed(ca.linalg.PRNGKey(42))
    Om = torch.randn(n, 64, generator=g, device=dev).cpu()
    Kn = K.numpy()
    ref = onys.NystromPrecond(lambda V: Kn @ V, Om.numpy(), mu=noise, eps=1e-8, adjust_mu=False)
    s32 = float(np.float32(noise))
    xo, _, ko, _ = ocg.run_batched_cg(lambda V: K @ V + s32 * V, b, None, 3000, 1e-5,
                                      lambda V: torch.from_numpy(ref.matmat(V.numpy())).float())
    print("steps plain", plain["iterations"], "nystrom", info["iterations"], "oracle", ko, "fp64 residual",
          res(x), "oracle", res(xo), flush=True)
    assert info["iterations"] < plain["iterations"] // 2
    assert abs(int(info["iterations"]) - 1 - ko) <= 2
    assert res(x) <= 2 * max(res(xo), 1e-5)
