"""Batched Lanczos with full double Gram-Schmidt, restated from
cola/linalg/decompositions/lanczos.py (torch, CPU).

State layout is the reference's: V (b, n, m+2) with column 0 zero and the start
vector in column 1; diag (b, m); subdiag (b, m+1) (lanczos.py:275-284).
"""
import numpy as np
import torch

from .cg import while_loop_winfo


def init_lanczos(rhs, max_iters, dtype):
    """lanczos.py:275-284. rhs is (n, b)."""
    diag = torch.zeros((rhs.shape[-1], max_iters), dtype=dtype)
    subdiag = torch.zeros((rhs.shape[-1], max_iters + 1), dtype=dtype)
    V = torch.zeros((rhs.shape[-1], rhs.shape[0], max_iters + 2), dtype=dtype)
    norm = torch.linalg.norm(rhs, dim=-2, keepdim=True)
    rhs = rhs / norm
    V[..., 1] = rhs.T.clone()
    return V, diag, subdiag, 1


def do_gram(V, w):
    """lanczos.py:293-296: w -= V (V^H w) over all m+2 columns."""
    aux = torch.sum(V.conj() * w[..., None], dim=-2, keepdim=True)
    return w - torch.sum(V * aux, dim=-1)


def do_double_gram(V, w):
    """lanczos.py:287-290."""
    return do_gram(V, do_gram(V, w))


def lanczos_fact(matmat, init_val, max_iters, tol):
    """lanczos.py:235-272. matmat acts on (n, b)."""

    def body_fun(state):
        V, diag, subdiag, i = state
        V = V.clone()
        diag = diag.clone()
        subdiag = subdiag.clone()
        update = torch.linalg.norm(V[..., i], dim=-1, keepdim=True)
        V[..., i] = V[..., i] / update
        new_vec = matmat(V[..., i].T).T
        upd = torch.sum(new_vec.conj() * V[..., i], dim=-1)
        diag[..., i - 1] = upd
        aux = diag[..., [i - 1]] * V[..., i] + subdiag[..., [i - 1]] * V[..., i - 1]
        new_vec = new_vec - aux
        new_vec = do_double_gram(V, new_vec)
        V[..., i + 1] = new_vec
        subdiag[..., i] = torch.linalg.norm(V[..., i + 1], dim=-1)
        return V, diag, subdiag, i + 1

    def error(state):
        *_, subdiag, i = state
        err = torch.tensor(1e-30, dtype=subdiag.real.dtype)
        rel_err = subdiag[..., i - 1].real / torch.maximum(subdiag[..., 1].real, err)
        return torch.max(rel_err, dim=0).values + (i <= 1) * 1.

    def cond_fun(state):
        *_, subdiag, i = state
        is_not_max = i <= max_iters
        is_large = (subdiag[..., i - 1].real > tol * subdiag[..., 1].real) | (i <= 1)
        return bool(is_not_max and torch.any(is_large))

    (V, diag, subdiag, i), info = while_loop_winfo(error, cond_fun, body_fun, init_val)
    return V, diag, subdiag, i, info


def lanczos(matmat, n, start_vector, max_iters=100, tol=1e-7, dtype=torch.float32):
    """lanczos.py:185-232 (torch branch, which trims to the executed iterations).

    Returns (Q (b, n, k), alpha (b, k-1) off-diagonal, beta (b, k) diagonal, info)."""
    max_iters = min(max_iters, n)
    rhs = start_vector[:, None] if start_vector.dim() == 1 else start_vector
    init_val = init_lanczos(rhs, max_iters, dtype)
    vec, beta, alpha, i, info = lanczos_fact(matmat, init_val, max_iters, tol)
    alpha, beta, Q, iters = alpha[..., 1:-1], beta, vec[..., 1:-1], i - 1
    alpha, beta, Q = alpha[..., :iters - 1], beta[..., :iters], Q[..., :iters]
    return Q, alpha, beta, info


def tridiag_dense(alpha, beta):
    """Tridiagonal(alpha, beta, alpha).to_dense() for one system."""
    T = torch.diag(beta)
    if alpha.numel():
        T = T + torch.diag(alpha, 1) + torch.diag(alpha, -1)
    return T


def lanczos_eigs(matmat, n, start_vector, max_iters=100, tol=1e-7, dtype=torch.float32, eig_dtype=None):
    """lanczos.py:34-61 for a single start vector: eigh(T) -> ascending -> Q @ eigvecs.

    eig_dtype=None keeps the reference's choice (A's dtype); torch.float64 gives the
    FP64 tridiagonal eigensolve the engine uses."""
    Q, alpha, beta, info = lanczos(matmat, n, start_vector, max_iters, tol, dtype)
    T = tridiag_dense(alpha[0], beta[0])
    if eig_dtype is not None:
        T = T.to(eig_dtype)
    eigvals, eigvecs = torch.linalg.eigh(T)
    idx = torch.argsort(eigvals)
    V = Q[0] @ eigvecs[:, idx].to(Q.dtype)
    return eigvals[idx], V, info


def get_slice(num, which):
    """cola/linalg/decompositions/decompositions.py:214-224."""
    if num == -1:
        raise ValueError(f"Number of eigenvalues {num} must be explicitly specified")
    if which == "SM":
        return slice(0, num, None)
    if which == "LM":
        return slice(-1 if num is None else -num, None, None)
    raise NotImplementedError(f"which={which} is not implemented")


def laplacian_2d_eigenvalues(m, shift=0.0):
    """Closed form of the 5-point Dirichlet Laplacian on an m x m grid (FP64, ascending)."""
    j = np.arange(1, m + 1)
    lam1 = 2.0 - 2.0 * np.cos(np.pi * j / (m + 1))
    return np.sort((lam1[:, None] + lam1[None, :]).ravel() + shift)
