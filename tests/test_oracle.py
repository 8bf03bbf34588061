"""The oracle pinned against the reference tests' known answers, published
constants and closed forms -- never against a retyped copy of its own formula."""
import json
import os

import numpy as np
import pytest
import scipy.io
import scipy.linalg
import scipy.sparse
import torch

from oracle import arnoldi as oar
from oracle import cg as ocg
from oracle import hutch as ohu
from oracle import lanczos as olz
from oracle import operators as oop
from oracle import rng as orng
from oracle import tridiag as otd
from oracle import unary as oun

GOLD = os.path.join(os.path.dirname(__file__), "golden")
KAT = json.load(open(os.path.join(GOLD, "kat.json")))


def spd(diag, seed):
    rs = np.random.RandomState(seed)
    Q, _ = np.linalg.qr(rs.normal(size=(len(diag), len(diag))))
    return torch.from_numpy(Q.T @ np.diag(diag) @ Q)


# ----------------------------------------------------------------------------- rng
def test_prng_key_chain():
    k = KAT["prng_keys"]
    assert orng.PRNGKey(42) == k["PRNGKey42"]
    assert orng.next_key(orng.PRNGKey(42)) == k["next_key_of_42"]
    assert orng.PRNGKey(0) == k["PRNGKey0"]


def test_splitmix64_published_outputs():
    """mix64(state + gamma*i) for state 0 gives SplitMix64's published sequence."""
    outs = [int(orng.mix64(np.uint64((orng.GOLDEN * i) % 2**64))) for i in (1, 2, 3)]
    assert [f"{o:016x}" for o in outs] == KAT["splitmix64_seed0"]["outputs_hex"]


def test_rademacher_and_gaussian_streams():
    Z = orng.rademacher(3, 5000, 130)
    assert set(np.unique(Z.numpy())) == {-1.0, 1.0}
    assert abs(float(Z.mean())) < 0.01
    # probe p reads bit p%64 of word (p//64): check one by hand
    w = orng.stream_word(3, 2, 17)
    assert Z[128 + 1, 17] == (1.0 - 2.0 * float((int(w) >> 1) & 1))
    G = orng.gaussian(9, 20000, 4, torch.float64)
    assert abs(float(G.mean())) < 0.02 and abs(float(G.std()) - 1) < 0.02
    assert abs(float((G**4).mean()) - 3) < 0.15


# ----------------------------------------------------------------------------- operators
def test_csr_known_answer():
    k = KAT["csr_3x4"]
    ip, ix, v = oop.coo_to_csr(np.array(k["data"], np.float32), k["rows"], k["cols"], tuple(k["shape"]))
    out = oop.csr_matmat(ip, ix, torch.from_numpy(v), torch.tensor(k["x1"])[:, None])
    assert torch.allclose(out[:, 0], torch.tensor(k["Ax1"]))
    D = scipy.sparse.csr_matrix((v, ix, ip), shape=tuple(k["shape"])).toarray()
    assert np.allclose(np.array(k["x2"]) @ D, k["x2A"])


def test_kron_matmat_vs_torch_kron():
    k = KAT["kron_2x3_4x2"]
    A, B = torch.tensor(k["A"]).double(), torch.tensor(k["B"]).double()
    full = torch.kron(A, B)
    X = torch.randn(full.shape[1], 3, dtype=torch.float64)
    assert torch.allclose(oop.kron_matmat([A, B], X), full @ X)
    g = torch.Generator().manual_seed(0)
    facs = [torch.randn(d, d, generator=g, dtype=torch.float64) for d in (3, 4, 5)]
    X = torch.randn(60, 2, generator=g, dtype=torch.float64)
    assert torch.allclose(oop.kron_matmat(facs, X), torch.kron(torch.kron(facs[0], facs[1]), facs[2]) @ X)


def test_1138_bus_csr_vs_scipy():
    M = scipy.io.mmread(os.path.join(GOLD, "1138_bus.mtx")).tocsr()
    coo = M.tocoo()
    ip, ix, v = oop.coo_to_csr(coo.data, coo.row, coo.col, M.shape)
    assert np.array_equal(ip, M.indptr) and np.array_equal(ix, M.indices)
    X = torch.randn(M.shape[0], 2, dtype=torch.float64)
    assert torch.allclose(oop.csr_matmat(ip, ix, torch.from_numpy(v), X), torch.from_numpy(M @ X.numpy()))


def test_laplacian_csr_and_closed_form():
    ip, ix, v = oop.laplacian_2d_csr(12, 1e-3, dtype=np.float64)
    D = scipy.sparse.csr_matrix((v, ix, ip)).toarray()
    assert np.allclose(D, D.T)
    assert np.allclose(np.linalg.eigvalsh(D), olz.laplacian_2d_eigenvalues(12, 1e-3))
    k = KAT["laplacian_1024"]
    assert np.allclose(olz.laplacian_2d_eigenvalues(1024, 1e-3)[-32:], k["top32"])
    assert len(oop.laplacian_2d_csr(64)[2]) == 5 * 64 * 64 - 4 * 64


# ----------------------------------------------------------------------------- CG
def test_cg_diag345():
    k = KAT["cg_diag345"]
    A = torch.diag(torch.tensor(k["A_diag"]))
    x, info = ocg.cg(lambda V: A @ V, torch.tensor(k["b"]), tol=1e-7, max_iters=5)
    assert torch.allclose(x, torch.tensor(k["x"]), atol=1e-6)


def test_cg_cond_fun_table():
    for c in KAT["cg_cond_fun"]["cases"]:
        res = torch.tensor(c["res"])
        assert ocg.cond_fun((None, c["k"], res, None), c["tol"], c["max_iters"]) == c["flag"]


def test_cg_repeated_eigenvalues():
    k = KAT["cg_repeated_eig"]
    A = spd(np.array(k["diag"]), 0).float()
    x, _, steps, _ = ocg.run_batched_cg(lambda V: A @ V, torch.ones(30, 1), None, 100, 1e-11, None)
    assert steps < k["max_k_exclusive"]


def test_cg_random_spd_fp64_and_info():
    A = spd(np.linspace(0.1, 2, 80), 3)
    b = torch.randn(80, 4, dtype=torch.float64)
    x, r, k, info = ocg.run_batched_cg(lambda V: A @ V, b, None, 500, 1e-10, None)
    assert torch.allclose(x, torch.linalg.solve(A, b), atol=1e-8)
    assert info["iterations"] == k + 1 and len(info["errors"]) == k  # (k+1)+1 entries, first two dropped
    assert torch.allclose(r, b - A @ x, atol=1e-8)


def test_cg_zero_rhs_safe_division():
    A = spd(np.linspace(1, 2, 20), 1)
    b = torch.zeros(20, 2, dtype=torch.float64)
    b[:, 0] = 1.0
    x, _, _, _ = ocg.run_batched_cg(lambda V: A @ V, b, None, 100, 1e-8, None)
    assert torch.all(x[:, 1] == 0)


def test_cg_preconditioner_identity_equivalence():
    A = spd(np.linspace(0.5, 5, 40), 2)
    b = torch.randn(40, 1, dtype=torch.float64)
    x1, _, k1, _ = ocg.run_batched_cg(lambda V: A @ V, b, None, 100, 1e-9, None)
    x2, _, k2, _ = ocg.run_batched_cg(lambda V: A @ V, b, None, 100, 1e-9, lambda V: V)
    assert k1 == k2 and torch.allclose(x1, x2)
    # exact inverse as preconditioner converges in one step
    Ainv = torch.linalg.inv(A)
    x3, _, k3, _ = ocg.run_batched_cg(lambda V: A @ V, b, None, 100, 1e-9, lambda V: Ainv @ V)
    assert k3 <= 2 and torch.allclose(x3, torch.linalg.solve(A, b), atol=1e-9)


# ----------------------------------------------------------------------------- Lanczos
@pytest.mark.parametrize("idx", [0, 1, 2])
def test_lanczos_known_cases(idx):
    c = KAT["lanczos_cases"]["cases"][idx]
    A = torch.tensor(c["A"], dtype=torch.float32)
    Q, alpha, beta, info = olz.lanczos(lambda V: A @ V, 3, torch.tensor(c["rhs"], dtype=torch.float32)[:, None], 3,
                                       1e-7)
    assert info["iterations"] - 1 == c["iters"]
    assert torch.allclose(beta[0], torch.tensor(c["beta"], dtype=torch.float32), atol=1e-6)
    assert torch.allclose(alpha[0], torch.tensor(c["alpha"], dtype=torch.float32), atol=1e-6)


def test_lanczos_invariants_and_max_eig():
    A = spd(1 - np.linspace(0, 1, 41)[:-1]**0.5 + 1e-2, 21)
    s = torch.randn(40, 1, dtype=torch.float64)
    Q, alpha, beta, _ = olz.lanczos(lambda V: A @ V, 40, s, 40, 1e-12, torch.float64)
    T = olz.tridiag_dense(alpha[0], beta[0])
    assert torch.allclose(Q[0].T @ Q[0], torch.eye(Q.shape[2], dtype=torch.float64), atol=1e-10)
    assert torch.allclose(Q[0].T @ A @ Q[0], T, atol=1e-10)
    k = KAT["lanczos_max_eig"]
    D = torch.diag(torch.tensor(k["diag"], dtype=torch.float64))
    ev, _, _ = olz.lanczos_eigs(lambda V: D @ V, 4, torch.ones(4, dtype=torch.float64), 4, 1e-7, torch.float64)
    assert abs(float(ev[-1]) - k["max"]) < 1e-10


def test_lanczos_laplacian_full_dimension():
    """Full Lanczos on a 6x6-grid Laplacian recovers the closed-form spectrum."""
    ip, ix, v = oop.laplacian_2d_csr(6, 1e-3, dtype=np.float64)
    D = torch.from_numpy(scipy.sparse.csr_matrix((v, ix, ip)).toarray())
    s = torch.randn(36, dtype=torch.float64)
    ev, _, _ = olz.lanczos_eigs(lambda V: D @ V, 36, s, 36, 1e-14, torch.float64)
    exact = olz.laplacian_2d_eigenvalues(6, 1e-3)
    # Lanczos finds each distinct eigenvalue once; compare against the distinct set
    distinct = np.unique(np.round(exact, 9))
    assert all(np.min(np.abs(distinct - e)) < 1e-8 for e in ev.numpy())


# ----------------------------------------------------------------------------- Arnoldi
def test_arnoldi_factorisation_invariants():
    rs = np.random.RandomState(48)
    n = 30
    A = torch.from_numpy(np.tril(rs.normal(size=(n, n))) + np.diag(np.linspace(1, 2, n)))
    s = torch.randn(n, 1, dtype=torch.float64)
    Q, H, info = oar.arnoldi(lambda V: A @ V, n, s, 12, 1e-12, torch.float64)
    Q, H = Q[0], H[0]
    assert torch.allclose(A @ Q[:, :-1], Q @ H, atol=1e-10)
    assert torch.allclose(Q.T @ Q, torch.eye(13, dtype=torch.float64), atol=1e-10)
    assert torch.all(torch.tril(H, -2) == 0)


def test_arnoldi_triangular_eigs():
    rs = np.random.RandomState(4)
    n = 10
    L = np.tril(rs.normal(size=(n, n)) * 0.2)
    d = np.linspace(1, 2, n)
    np.fill_diagonal(L, d)
    A = torch.from_numpy(L)
    ev, _, _ = oar.arnoldi_eigs(lambda V: A @ V, n, torch.ones(n, dtype=torch.float64), n, 1e-12, torch.float64)
    assert np.allclose(np.sort(ev.real.numpy()), d, atol=1e-8)


# ----------------------------------------------------------------------------- SLQ / unary / Hutchinson
def test_slq_unit_probes_exact_logdet():
    n = 10
    diag = 1 - np.linspace(0, 1, n + 1)[:-1]**0.5 + 0.05
    A = spd(diag, 4)
    Z = np.sqrt(n) * torch.eye(n, dtype=torch.float64)
    est = oun.slq_fwd(lambda V: A @ V, n, torch.log, Z, n, 1e-12, torch.float64, torch.float64)
    assert abs(float(est) - np.sum(np.log(diag))) < 1e-9


def test_slq_stochastic_10x10():
    """tests/algorithms/test_stochastic_lanczos_quad.py:48-64 (rel error < 0.1)."""
    n = 10
    diag = 1 - np.linspace(0, 1, n + 1)[:-1]**0.5
    diag = diag.astype(np.float32) + 1e-3
    A = spd(diag, 0).float()
    num = oun.num_samples_from_vtol(1 / np.sqrt(70))
    assert num == 70 - 1 or num == 70  # int() truncation of 1/vtol^2 (slq.py:74)
    Z = orng.gaussian(1, n, num).T
    est = oun.slq_fwd(lambda V: A @ V, n, torch.log, Z, n, 1e-7)
    ref = float(np.sum(np.log(diag)))
    assert abs(float(est) - ref) / abs(ref) < 0.1


def test_slq_vtol_truncation():
    assert oun.num_samples_from_vtol(32**-0.5) == 31
    assert oun.num_samples_from_vtol(0.17677) == 32


def test_lanczos_unary_log_matches_dense():
    A = spd(np.linspace(0.3, 3, 25), 8)
    V = torch.randn(25, 3, dtype=torch.float64)
    out, _ = oun.lanczos_unary_matmat(lambda X: A @ X, 25, torch.log, V, 25, 1e-12, torch.float64)
    w, U = torch.linalg.eigh(A)
    assert torch.allclose(out, U @ torch.diag(torch.log(w)) @ U.T @ V, atol=1e-8)


def test_hutch_hadamard_exact_and_closed_form_trace():
    n = 32
    H = torch.from_numpy(scipy.linalg.hadamard(n)).double()
    A = torch.randn(n, n, dtype=torch.float64)
    assert abs(float(ohu.trace_from_probes(lambda V: A @ V, H.T)) - float(torch.trace(A))) < 1e-10
    ip, ix, v = oop.laplacian_2d_csr(16, 1e-3, dtype=np.float64)
    mm = lambda X: oop.csr_matmat(ip, ix, torch.from_numpy(v), X)  # noqa: E731
    Z = orng.rademacher(1, 256, 256, torch.float64).T
    tr = float(ohu.trace_from_probes(mm, Z))
    assert abs(tr - 4.001 * 256) < 6 * np.sqrt(2 * 4 * 256 / 256)


def test_hutchinson_diag_estimate_converges():
    A = spd(np.linspace(1, 3, 120), 6).float()
    d, info = ohu.hutchinson_diag_estimate(lambda V: A @ V, 120, tol=3e-2, rand="rademacher")
    assert float((d - torch.diag(A)).abs().mean()) < 0.3
    assert info["iterations"] >= 2


# ----------------------------------------------------------------------------- tridiagonal
def test_tridiag_vs_dense():
    rs = np.random.RandomState(0)
    d, e = rs.normal(size=30), rs.normal(size=29)
    w, v = otd.eigh_tridiag(d, e)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.allclose(w, np.linalg.eigvalsh(T))
    assert np.allclose(T @ v, v * w[None, :], atol=1e-12)


# ----------------------------------------------------------------------------- pins added for the stop rule,
# the k-th diagonal and the per-probe quadrature
def circulant_two_neighbours(n, c):
    """I + c (S + S^T) with S the cyclic shift: every row has exactly two off-diagonal entries c."""
    A = torch.eye(n, dtype=torch.float64)
    idx = torch.arange(n)
    A[idx, (idx + 1) % n] = c
    A[idx, (idx - 1) % n] = c
    return A


def test_hutch_stop_rule_block_count_from_known_variance():
    """Rademacher estimator of row i: (Az)_i z_i = A_ii + sum_{j != i} A_ij z_i z_j, variance
    sum_{j != i} A_ij^2 = 2c^2.  With A_ii = 1 the mean relative stderr after b blocks of 100
    probes is sqrt(2c^2 / (100 b)) (diagonal_estimation.py:197-203); 2c^2 = 0.765 puts
    tol = 3e-2 between b = 8 (0.0309) and b = 9 (0.0292), so the loop runs exactly 9 blocks.
    A stderr without the 1/(b*bs) factor, or with bs in place of b*bs, stops elsewhere."""
    n = 1000
    A = circulant_two_neighbours(n, np.sqrt(0.765 / 2))
    d, info = ohu.hutchinson_diag_estimate(lambda V: A @ V, n, dtype=torch.float64, tol=3e-2, rand="rademacher")
    assert info["iterations"] == 10  # 9 blocks + the first condition evaluation
    # torch_tqdm.py:56-62: errors [e2..e9] plus the final error again
    assert len(info["errors"]) == 9 and info["errors"][-1] == info["errors"][-2]
    assert 0.0292 * 0.97 < info["errors"][-1] <= 3e-2
    assert float((d - 1).abs().max()) < 6 * np.sqrt(0.765 / 900)


@pytest.mark.parametrize("k", [0, 1, -1, 5, -7])
def test_hutch_kth_diagonal_exact_with_hadamard_probes(k):
    """Hadamard probes satisfy sum_b z_b z_b^T = n I, so one block of n probes gives the k-th
    diagonal A[j, j+k] exactly: pins the roll direction and the slicing of :192-195."""
    n = 64
    H = torch.from_numpy(scipy.linalg.hadamard(n)).double()
    A = torch.randn(n, n, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    d, info = ohu.hutchinson_diag_estimate(lambda V: A @ V, n, dtype=torch.float64, k=k, tol=3e-2,
                                           max_iters=1, rand="rademacher", probe_fn=lambda key, n_, bs: H)
    assert d.shape == (n - abs(k),)
    assert torch.allclose(d, torch.diagonal(A, k), atol=1e-12)


def test_slq_quadratic_forms_exact_with_full_lanczos():
    """With Lanczos run to full dimension the Gauss quadrature is exact: the per-probe form
    ||z||^2 sum_j tau_j^2 log(theta_j) equals z^T log(A) z from a dense FP64 eigendecomposition,
    for unit probes (whose sum is log det A) and for Rademacher probes (||z||^2 = n)."""
    n = 24
    A = spd(np.linspace(0.4, 6.0, n), 21)
    w, U = torch.linalg.eigh(A)
    logA = U @ torch.diag(torch.log(w)) @ U.T
    for Z in (torch.eye(n, dtype=torch.float64),
              orng.rademacher(5, n, 6, torch.float64).T):
        _, alpha, beta, _ = olz.lanczos(lambda V: A @ V, n, Z, n, 1e-13, torch.float64)
        q, _, _ = oun.slq_quadratic_forms(alpha, beta, (Z**2).sum(0), torch.log)
        exact = torch.einsum("ib,ij,jb->b", Z, logA, Z)
        assert torch.allclose(q, exact, rtol=1e-9, atol=1e-9)
    _, alpha, beta, _ = olz.lanczos(lambda V: A @ V, n, torch.eye(n, dtype=torch.float64), n, 1e-13, torch.float64)
    q, _, _ = oun.slq_quadratic_forms(alpha, beta, torch.ones(n, dtype=torch.float64), torch.log)
    assert abs(float(q.sum()) - float(torch.log(w).sum())) < 1e-9


def test_slq_quadratic_forms_agrees_with_slq_fwd_for_rademacher():
    """slq.py:51 scales by n, the per-probe form by ||z||^2: equal for Rademacher probes, also
    with a truncated (m < n) Lanczos run."""
    n = 200
    A = spd(np.linspace(0.5, 9.0, n), 8)
    Z = orng.rademacher(9, n, 16, torch.float64).T
    _, alpha, beta, _ = olz.lanczos(lambda V: A @ V, n, Z, 25, 1e-10, torch.float64)
    q, _, _ = oun.slq_quadratic_forms(alpha, beta, (Z**2).sum(0), torch.log)
    est = oun.slq_fwd(lambda V: A @ V, n, torch.log, Z, 25, 1e-10, torch.float64, torch.float64)
    assert abs(float(q.mean()) - float(est)) < 1e-9 * abs(float(est))


# ------------------------------------------------------------------------------- GMRES
def _pd_from_diag(diag, seed):
    """cola/utils/utils_for_tests.py:164-173 with a seeded generator: Q^T diag Q, Q from QR of a normal matrix."""
    rs = np.random.RandomState(seed)
    Q, _ = np.linalg.qr(rs.normal(size=(len(diag), len(diag))))
    return Q.T @ np.diag(diag) @ Q


def test_oracle_gmres_easy_known_answer():
    """tests/algorithms/test_gmres.py:72-88 (easy): diag(3, 4, 5) x = 1 with 3 steps -> (1/3, 1/4, 1/5)."""
    from oracle import gmres as og
    A = torch.diag(torch.tensor([3., 4., 5.]))
    x, info = og.gmres(lambda V: A @ V, 3, torch.ones(3, 1), None, 3, 1e-8)
    assert torch.allclose(x[:, 0], torch.tensor([1 / 3, 1 / 4, 1 / 5]), rtol=1e-6)


def test_oracle_gmres_random_spectrum():
    """tests/algorithms/test_gmres.py (random): spectrum 1 - x^0.5 on 25 points, max_iters = n - 5,
    tol 1e-8, three rhs of ones: relative error below 5e-4."""
    from oracle import gmres as og
    x_ = np.linspace(0, 1, num=26)[:-1]
    diag = (1 - x_**0.5).astype(np.float32)
    A = torch.from_numpy(_pd_from_diag(diag, 7)).float()
    rhs = torch.ones(25, 3)
    soln = torch.linalg.solve(A.double(), rhs.double())
    x, _ = og.gmres(lambda V: A @ V, 25, rhs, None, 20, 1e-8)
    assert float((x.double() - soln).norm() / soln.norm()) < 5e-4


def test_oracle_gmres_equals_fom_from_independent_arnoldi():
    """With m < n steps and no padded rows, the reference's normal equations of the square m x m
    Hessenberg block give x0 + Q_m H_m^-1 beta e1: against that formula built from an independent
    FP64 Arnoldi with full re-orthogonalisation (numpy)."""
    from oracle import gmres as og
    rs = np.random.RandomState(11)
    n, m = 60, 15
    A = rs.normal(size=(n, n)) / np.sqrt(n) + 3 * np.eye(n)
    b = rs.normal(size=n)
    x, _ = og.gmres(lambda V: torch.from_numpy(A) @ V, n, torch.from_numpy(b), None, m, 1e-12, torch.float64)
    Q = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    Q[:, 0] = b / np.linalg.norm(b)
    for j in range(m):
        w = A @ Q[:, j]
        for _ in range(2):
            h = Q[:, :j + 1].T @ w
            w -= Q[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        Q[:, j + 1] = w / H[j + 1, j]
    e1 = np.zeros(m)
    e1[0] = np.linalg.norm(b)
    fom = Q[:, :m] @ np.linalg.solve(H[:m, :m], e1)
    assert np.linalg.norm(x.numpy() - fom) / np.linalg.norm(fom) < 1e-10


def test_oracle_gmres_early_stop_least_squares():
    """A diagonalisable non-symmetric matrix with three distinct eigenvalues: Arnoldi stops after three
    steps, the padded tail drops out, and x is the exact solution."""
    from oracle import gmres as og
    rs = np.random.RandomState(3)
    n = 12
    V = rs.normal(size=(n, n)) + 4 * np.eye(n)
    lam = np.repeat([1.0, 2.0, 5.0], 4)
    A = V @ np.diag(lam) @ np.linalg.inv(V)
    b = rs.normal(size=(n, 1))
    x, info = og.gmres(lambda W: torch.from_numpy(A) @ W, n, torch.from_numpy(b), None, 10, 1e-10, torch.float64)
    assert info["iterations"] <= 5
    assert np.linalg.norm(A @ x.numpy() - b) / np.linalg.norm(b) < 1e-8
