"""GMRES restated from cola/linalg/inverse/gmres.py:39-124 (torch, CPU): Arnoldi (MGS) of the
residual, the normal equations of the square Hessenberg block with the reference's padding of
near-zero rows, soln = x0 + Q y.  Test infrastructure only: the engine never imports it."""
import torch

from .arnoldi import arnoldi


def gmres_fwd(matmat, n, rhs, x0, max_iters, tol, dtype=torch.float32):
    """gmres.py:85-124 (use_householder=False, use_triangular=False). rhs, x0 (n, k)."""
    res = rhs - matmat(x0)
    Q, H, info = arnoldi(matmat, n, res, max_iters, tol, dtype)
    Q, H = Q[:, :, :-1], H[:, :-1, :]
    beta = torch.linalg.norm(res, dim=-2)
    HT = H.conj().permute(0, 2, 1)
    largest_vals = torch.max(torch.abs(H), -1).values
    overall_max = torch.max(largest_vals.reshape(largest_vals.shape[0], -1), -1).values
    zero_thresh = 10 * tol * overall_max[:, None]
    padding = torch.where(largest_vals < zero_thresh, torch.ones_like(largest_vals), torch.zeros_like(largest_vals))
    added_diag = torch.vmap(torch.diag)(padding)
    y = torch.linalg.solve(HT @ H + added_diag, HT[..., 0, None]).squeeze(-1) * beta[:, None]
    y = torch.where(largest_vals < zero_thresh, torch.zeros_like(y), y)
    pred = (Q @ y[..., None]).permute(1, 0, 2)[:, :, 0]
    return x0 + pred, info


def gmres(matmat, n, rhs, x0=None, max_iters=100, tol=1e-7, dtype=torch.float32):
    """gmres.py:39-69: vector rhs promoted to (n, 1); x0 defaults to zeros."""
    is_vector = rhs.dim() == 1
    if x0 is None:
        x0 = torch.zeros_like(rhs)
    if is_vector:
        rhs, x0 = rhs[:, None], x0[:, None]
    soln, info = gmres_fwd(matmat, n, rhs, x0, max_iters, tol, dtype)
    return (soln[:, 0] if is_vector else soln), info
