"""HipGMRES: GMRES (cola/linalg/inverse/gmres.py:11-124) on the engine's Arnoldi kernels.

Drop-in for cola's ``GMRES`` dataclass (same fields and defaults) and ``gmres()``.
The residual r0 = b - A x0 starts a batched Arnoldi run (``run_arnoldi``: MGS by default,
the reference's order, or CGS2 with ``mgs=False``); the small least-squares problem is
solved on the device by Givens rotations in FP64 (``ck_hessenberg_lstsq_f64``), and
x = x0 + Q y is formed by ``ck_basis_combine_batched``.

The reference solves the normal equations of the square Hessenberg block, padding the
rows whose largest entry is below 10*tol*max|H| and zeroing their y entries
(gmres.py:108-121).  Those rows are the tail left by an Arnoldi run that stopped early,
so the engine solves the (k+1) x k least-squares problem of the k steps taken, or the
m x m square block when all m steps ran: the same solution without squaring the
condition number.  As in the reference, P is accepted and not applied
(gmres.py:93-97 never uses it), and the backward rule (gmres_bwd, gmres.py:72-82)
solves A db = dx with GMRES again, which is the exact gradient for symmetric A.
"""
from dataclasses import dataclass

import torch

from .. import _native as N
from .._blocks import ld_of
from ..autodiff import GradOperator, iterative_autograd, vjp_derivs
from .algorithm_base import Algorithm
from .arnoldi import run_arnoldi


@dataclass
class HipGMRES(Algorithm):
    """GMRES on the HIP engine.

    Args:
        tol (float): Arnoldi stopping tolerance (relative to the first subdiagonal entry).
        max_iters (int): Arnoldi steps.
        pbar (bool): accepted for compatibility; no progress bar is drawn.
        x0 (Tensor): (n,) or (n, b) initial guess.
        P (LinearOperator): accepted and not applied, as in the reference.
        mgs (bool): modified Gram-Schmidt (the reference's order) or two CGS passes.
        chunk (int): Arnoldi steps between reads of the device loop flag.
    """
    tol: float = 1e-6
    max_iters: int = 1000
    pbar: bool = False
    x0: object = None
    P: object = None
    mgs: bool = True
    chunk: int = 32

    def __call__(self, A, b):
        return gmres(A, b, x0=self.x0, max_iters=self.max_iters, tol=self.tol, P=self.P, pbar=self.pbar,
                     mgs=self.mgs, chunk=self.chunk)


def gmres(A, rhs, x0=None, max_iters=100, tol=1e-7, P=None, use_householder=False, use_triangular=False, pbar=False,
          mgs=True, chunk=32):
    """gmres.py:39-69: rhs (n,) or (n, b) on the HIP device; returns (soln, info)."""
    N.require_device(rhs, "right-hand side")
    if use_householder:
        raise NotImplementedError("the engine's Arnoldi is Gram-Schmidt (MGS or CGS2); Householder is not built")
    is_vector = rhs.dim() == 1
    if x0 is None:
        x0 = torch.zeros_like(rhs)
    if is_vector:
        rhs, x0 = rhs[:, None], x0[:, None]
    soln, info = gmres_fwd(A, rhs, x0, max_iters, tol, P, mgs, chunk)
    return (soln[:, 0] if is_vector else soln), info


def gmres_bwd(res, grads, unflatten, *args, **kwargs):
    """gmres.py:72-82: db from a second GMRES solve on the output gradient, dA = vjp of
    theta -> A(theta) @ soln with -db."""
    rhs, x0, max_iters, tol, P, *rest = args
    mgs = rest[0] if len(rest) > 0 else True
    chunk = rest[1] if len(rest) > 1 else 32
    y_grads = grads[0]
    op_args, output = res
    soln = output[0]
    A = unflatten(op_args)
    if y_grads is None:
        return GradOperator([None] * len(op_args)), None
    db, _ = _gmres_fwd(A, y_grads.contiguous(), x0, max_iters, tol, P, mgs, chunk)

    def fun(*theta):
        return unflatten(theta) @ soln

    return GradOperator(vjp_derivs(fun, op_args, -db)), db


def _gmres_fwd(A, rhs, x0, max_iters, tol, P, mgs=True, chunk=32):
    """gmres.py:85-124."""
    n, nb = rhs.shape
    dev = rhs.device
    x0 = x0.to(rhs.dtype)
    res = rhs - A @ x0 if bool(torch.any(x0 != 0)) else rhs.clone()
    Qb, H, steps, info = run_arnoldi(A, res.float().contiguous(), max_iters, tol, mgs, chunk)
    m = H.shape[2]
    k = min(int(steps), m)
    rows, cols = (m, m) if k >= m else (k + 1, k)
    beta = torch.linalg.norm(res.float(), dim=0).contiguous()
    y = torch.zeros((nb, max(cols, 1)), dtype=torch.float64, device=dev)
    wbytes = N.q("ck_hessenberg_lstsq_workspace_bytes", nb, rows, cols)
    work = torch.empty(max(wbytes, 8) // 8, dtype=torch.float64, device=dev)
    s = N.stream_ptr(dev)
    N.call("ck_hessenberg_lstsq_f64", nb, m, N.ptr(H), H.stride(0), rows, cols, N.ptr(beta), 10.0 * float(tol),
           N.ptr(work), work.numel() * 8, N.ptr(y), y.stride(0), s)
    ld = ld_of(n)
    out = torch.zeros((nb, ld), dtype=torch.float32, device=dev)
    if cols > 0:
        N.call("ck_basis_combine_batched", n, nb, cols, N.ptr(Qb), ld, Qb.stride(0), N.ptr(y), y.stride(0), N.ptr(out),
               ld, s)
    pred = out[:, :n].T
    info = dict(info)
    info["arnoldi_steps"] = k
    return x0 + pred.to(x0.dtype), info


@iterative_autograd(gmres_bwd)
def gmres_fwd(A, rhs, x0, max_iters, tol, P, mgs=True, chunk=32):
    return _gmres_fwd(A, rhs, x0, max_iters, tol, P, mgs, chunk)
