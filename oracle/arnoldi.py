"""Arnoldi with modified Gram-Schmidt, restated from
cola/linalg/decompositions/arnoldi.py:166-205,289-335 (torch, CPU).

State layout is the reference's: Q (b, n, m+1), H (b, m+1, m).
"""
import torch

from .cg import while_loop_winfo


def init_arnoldi(rhs, max_iters, dtype):
    """arnoldi.py:327-335. rhs is (n, b)."""
    H = torch.zeros((rhs.shape[-1], max_iters + 1, max_iters), dtype=dtype)
    Q = torch.zeros((rhs.shape[-1], rhs.shape[-2], max_iters + 1), dtype=dtype)
    norm = torch.linalg.norm(rhs, dim=-2)
    rhs = rhs / norm
    Q[..., 0] = rhs.T.clone()
    return Q, H, 0, norm


def arnoldi_fact(matmat, n, init_val, max_iters, tol):
    """arnoldi.py:289-324."""
    max_iters = min(max_iters, n)

    def cond_fun(state):
        _, H, idx, norm = state
        is_not_max = idx < max_iters
        is_large = (norm > tol * H[:, 1, 0].real) | (idx <= 0)
        return bool(is_not_max and torch.any(is_large))

    def body_fun(state):
        Q, H, idx, _ = state
        Q, H = Q.clone(), H.clone()
        new_vec = matmat(Q[..., idx].T).T
        h_vec = torch.zeros((H.shape[0], H.shape[1]), dtype=new_vec.dtype)
        for jdx in range(idx + 1):  # arnoldi.py:304-311
            angle = torch.sum(Q[..., jdx].conj() * new_vec, dim=-1)
            h_vec[..., jdx] = angle
            new_vec = new_vec - h_vec[..., [jdx]] * Q[..., jdx]
        norm = torch.linalg.norm(new_vec, dim=-1, keepdim=True)
        new_vec = new_vec / torch.clip(norm, min=tol / 2.)
        h_vec[..., idx + 1] = norm[:, 0]
        H[..., idx] = h_vec
        Q[..., idx + 1] = new_vec
        return Q, H, idx + 1, norm[:, 0]

    (Q, H, idx, _), info = while_loop_winfo(lambda s: s[-1][0], cond_fun, body_fun, init_val)
    return Q, H, idx, info


def arnoldi(matmat, n, start_vector, max_iters=100, tol=1e-7, dtype=torch.float32):
    """arnoldi.py:166-205 (MGS branch). Returns Q (b, n, m+1), H (b, m+1, m), info."""
    rhs = start_vector[:, None] if start_vector.dim() == 1 else start_vector
    init_val = init_arnoldi(rhs, max_iters, dtype)
    Q, H, _, info = arnoldi_fact(matmat, n, init_val, max_iters, tol)
    return Q, H, info


def arnoldi_eigs(matmat, n, start_vector, max_iters=100, tol=1e-7, dtype=torch.float32):
    """arnoldi.py:35-62: eig of H[:-1] (complex), eigenvectors Q[:, :-1] @ vs."""
    Q, H, info = arnoldi(matmat, n, start_vector, max_iters, tol, dtype)
    Q, H = Q[0][:, :-1], H[0][:-1]
    eigvals, vs = torch.linalg.eig(H)
    return eigvals, Q.to(vs.dtype) @ vs, info
