"""f(A) V through Lanczos, and the stochastic Lanczos quadrature estimators.

* ``lanczos_unary_matmat`` -- cola/linalg/unary/unary.py:46-60 (LanczosUnary._matmat)
* ``slq_fwd``               -- cola/linalg/tbd/slq.py:37-52
* ``stochastic_lanczos_quad`` -- cola/linalg/tbd/slq.py:55-75 (probe count int(1/vtol^2))
"""
import torch

from .lanczos import lanczos


def _batched_tridiag(alpha, beta):
    b, k = beta.shape
    T = torch.zeros((b, k, k), dtype=beta.dtype)
    ar = torch.arange(k)
    T[:, ar, ar] = beta
    if k > 1:
        T[:, ar[:-1], ar[1:]] = alpha
        T[:, ar[1:], ar[:-1]] = alpha
    return T


def lanczos_unary_matmat(matmat, n, f, V, max_iters=100, tol=1e-6, dtype=torch.float32, eig_dtype=None):
    """unary.py:46-60. V (n, bs) -> f(A) V (n, bs)."""
    Q, alpha, beta, info = lanczos(matmat, n, V, max_iters, tol, dtype)
    T = _batched_tridiag(alpha, beta)
    if eig_dtype is not None:
        T = T.to(eig_dtype)
    eigvals, P = torch.linalg.eigh(T)
    norms = torch.linalg.norm(V, dim=0)
    zero_thresh = 10 * torch.finfo(eigvals.dtype).eps * torch.max(eigvals.abs(), dim=1, keepdim=True).values
    f_eigvals = torch.where(eigvals.abs() > zero_thresh, f(eigvals), torch.zeros_like(eigvals))
    out = P.conj()[:, 0, :] * norms[:, None].to(P.dtype)
    out = (Q.to(P.dtype) @ P @ (f_eigvals * out)[..., None])[..., 0]
    return out.T.to(V.dtype), info


def slq_quadratic_forms(alpha, beta, norms_sq, f, eig_dtype=torch.float64, mp_eps=None):
    """The per-probe Gauss quadrature z^T f(A) z = ||z||^2 sum_j tau_j^2 f(theta_j).

    With z = V[:, b]: ``lanczos_unary_matmat`` followed by z^T(.) collapses to this
    because Q^T z = ||z|| e_1 (unary.py:58-59). The zero threshold is the same
    10*eps*max|theta| rule (unary.py:55-57)."""
    T = _batched_tridiag(alpha, beta).to(eig_dtype)
    eigvals, P = torch.linalg.eigh(T)
    eps = torch.finfo(eigvals.dtype).eps if mp_eps is None else mp_eps
    zero_thresh = 10 * eps * torch.max(eigvals.abs(), dim=1, keepdim=True).values
    f_eigvals = torch.where(eigvals.abs() > zero_thresh, f(eigvals), torch.zeros_like(eigvals))
    tau = P[:, 0, :]
    return norms_sq.to(eig_dtype) * torch.sum(tau**2 * f_eigvals, dim=-1), eigvals, tau


def slq_fwd(matmat, n, f, probes, max_iters, tol, dtype=torch.float32, eig_dtype=None):
    """slq.py:37-52 with explicit probes (n, num_samples).

    Note slq.py:48 thresholds with max(eigvals) (signed), slq.py:51 scales by n."""
    _mp = torch.finfo(dtype).eps
    _, alpha, beta, _ = lanczos(matmat, n, probes, max_iters, tol, dtype)
    T = _batched_tridiag(alpha, beta)
    if eig_dtype is not None:
        T = T.to(eig_dtype)
    eigvals, Q = torch.linalg.eigh(T)
    tau = Q[..., 0, :]
    const = 10 * _mp * torch.max(eigvals, dim=1, keepdim=True).values
    fn_vals = torch.where(eigvals.abs() > const, f(eigvals), torch.zeros_like(eigvals))
    approx = torch.sum(tau**2 * fn_vals, dim=-1)
    estimate = n * approx
    return torch.mean(estimate, dim=0)


def num_samples_from_vtol(vtol):
    """slq.py:74."""
    return max(int(1 / vtol**2), 1)
