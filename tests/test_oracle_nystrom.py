"""The oracle's Nystrom restatement against the reference's known answers
(tests/algorithms/test_precond.py:60-162) and closed forms, on the CPU."""
import numpy as np

from oracle import nystrom as onys


def spectrum(coeff, scale, size):
    """utils_for_tests.py:188-196 (real case): 1 - x^coeff on size points of [0, 1)."""
    x = np.linspace(0, 1, num=size + 1)[:-1].astype(np.float32)
    return (scale * (1 - x**coeff)).astype(np.float32)


def pd_from_diag(diag, seed):
    """utils_for_tests.py:164-173: Q^T diag Q with Q from the QR of a seeded normal matrix."""
    rs = np.random.RandomState(seed)
    Q, _ = np.linalg.qr(rs.normal(size=(diag.shape[0], diag.shape[0])).astype(np.float32))
    return (Q.T @ np.diag(diag) @ Q).astype(np.float32)


def relative_error(v, w):
    """utils_for_tests.py:117-123."""
    return float(np.linalg.norm(v - w) / max((np.linalg.norm(v) + np.linalg.norm(w)) / 2.0, 1e-16))


def test_nys_approx_diagonal_known_answer():
    """test_precond.py:145-162: diag(3,4,5,7,1) sketched by e1, e5, e3 -> diag(3,0,5,0,1)."""
    A = np.diag(np.array([3., 4., 5., 7., 1.], np.float32))
    Omega = np.eye(5, dtype=np.float32)[:, [0, 4, 2]]
    Lam, U = onys.get_nys_approx(lambda V: A @ V, Omega, 1e-6)
    assert relative_error(np.diag(np.array([3., 0., 5., 0., 1.], np.float32)), U @ (Lam[:, None] * U.T)) < 1e-7
    assert np.allclose(np.sort(Lam), [1., 3., 5.], rtol=1e-6)


def test_nys_approx_full_rank_known_answer():
    """test_precond.py:125-142: rank n on a random PD matrix reproduces it (rel < 500 * 1e-7)."""
    A = pd_from_diag(spectrum(0.5, 1.0, 10), 3)
    Omega = np.random.RandomState(4).normal(size=(10, 10)).astype(np.float32)
    Lam, U = onys.get_nys_approx(lambda V: A @ V, Omega, 1e-6)
    assert relative_error(A, U @ (Lam[:, None] * U.T)) < 500 * 1e-7
    # a mistake in the shift nu (added and not removed) would show in the eigenvalues
    assert np.allclose(np.sort(Lam), np.sort(np.linalg.eigvalsh(A.astype(np.float64))), atol=2e-5)


def test_nys_precond_closed_form():
    """test_precond.py:91-122: diagonal A (spectrum 1 - x^0.5, n = 10), sketch = first 5 basis
    vectors, mu 1e-8, no mu adjustment: P = d_{r-1} Om diag(1/(d+mu)) Om^T + I - Om Om^T, and
    P^1/2 A P^1/2 = (d_{r-1} + mu) Om Om^T + Om2 (A + mu)[r:, r:] Om2^T."""
    diag = spectrum(0.5, 1.0, 10)
    A = np.diag(diag)
    rank, mu = 5, 1e-8
    eye = np.eye(10, dtype=np.float32)
    Om, Om2 = eye[:, :rank], eye[:, rank:]
    P_op = onys.NystromPrecond(lambda V: A @ V, Om.copy(), mu=mu, eps=1e-8, adjust_mu=False)
    P = P_op.matmat(eye)
    Psqrt = P**0.5
    approx = Psqrt @ A @ Psqrt
    soln_P = diag[rank - 1] * Om @ np.diag(1 / (diag[:rank] + mu)) @ Om.T + eye - Om @ Om.T
    inv = (A + mu * eye)[rank:, rank:]
    soln = (diag[rank - 1] + mu) * Om @ Om.T + Om2 @ (inv @ Om2.T)
    assert relative_error(soln, approx) < 1e-6
    assert relative_error(soln_P, P) < 1e-6


def test_nys_sqrt_inverse():
    """test_precond.py:60-83: sqrt(P)^2 = P and inverse(P) P = I for a rank-5 preconditioner of a
    random PD matrix of order 10 (seed 21)."""
    A = pd_from_diag(spectrum(0.5, 1.0, 10), 21)
    Omega = np.random.RandomState(42).normal(size=(10, 5)).astype(np.float32)
    P = onys.NystromPrecond(lambda V: A @ V, Omega)
    eye = np.eye(10, dtype=np.float32)
    Pd = P.matmat(eye)
    S = onys.sqrt(P).matmat(eye)
    assert relative_error(Pd, S @ S) < 1e-6
    Inv = onys.inverse(P).matmat(eye)
    assert relative_error(eye, Inv @ Pd) < 1e-6
