"""Nystrom preconditioner restated from cola/linalg/preconditioning/preconditioners.py:87-186
(NumPy, CPU).  Test infrastructure only: the engine never imports it.

Pinned by the reference's own known answers (tests/algorithms/test_precond.py:60-162): the
diagonal case with basis-vector sketches (exact Lambda and U), full rank on a random PD matrix
(U Lambda U^T = A), the preconditioner's closed form on a diagonal matrix, and sqrt/inverse
of the preconditioner squaring / multiplying back to itself / the identity."""
import numpy as np
import scipy.linalg


def get_nys_approx(A, Omega, eps):
    """preconditioners.py:143-153; A is a callable V -> A @ V."""
    Omega, _ = np.linalg.qr(Omega, mode="reduced")
    Y = A(Omega)
    nu = eps * np.linalg.norm(Y, ord="fro")
    Y = Y + nu * Omega
    C = np.linalg.cholesky(Omega.T @ Y)
    B = scipy.linalg.solve_triangular(C, Y.T, lower=True).T
    U, Sigma, _ = np.linalg.svd(B, full_matrices=False)
    Lambda = np.clip(Sigma**2.0 - nu, 0.0, None)
    return Lambda.astype(Omega.dtype), U.astype(Omega.dtype)


class NystromPrecond:
    """preconditioners.py:87-124 with an explicit sketch (the _create_approx path)."""

    def __init__(self, A, Omega, mu=1e-7, eps=1e-8, adjust_mu=True):
        self.Lambda, self.U = get_nys_approx(A, Omega, eps)
        self.adjusted_mu = amu = mu * np.max(self.Lambda) if adjust_mu else mu
        self.subspace_num = np.min(self.Lambda) + amu
        self.subspace_denom = self.Lambda + amu
        self.preconditioned_eigmax = np.min(self.Lambda) + amu
        self.preconditioned_eigmin = amu

    def matmat(self, V):
        """preconditioners.py:122-124 / 134-136."""
        scaling = (self.subspace_num / self.subspace_denom - 1)[:, None]
        return self.U @ (scaling * (self.U.T @ V)) + V


def sqrt(P):
    """preconditioners.py:165-171: the same U with sqrt(num) / sqrt(denom)."""
    Q = object.__new__(NystromPrecond)
    Q.U, Q.subspace_num, Q.subspace_denom = P.U, np.sqrt(P.subspace_num), np.sqrt(P.subspace_denom)
    return Q


def inverse(P):
    """preconditioners.py:180-186: num and denom swapped."""
    Q = object.__new__(NystromPrecond)
    Q.U, Q.subspace_num, Q.subspace_denom = P.U, P.subspace_denom, P.subspace_num
    return Q
