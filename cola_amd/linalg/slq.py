"""Stochastic Lanczos quadrature (cola/linalg/tbd/slq.py:37-75) on the engine.

One batched Lanczos run over all probes, FP64 tridiagonal eigensolves (first
eigenvector components only) and the quadrature reduction on the device:
estimate_b = scale_b * sum_j tau_bj^2 f(theta_bj), zero below 10*eps*max(theta)."""
import ctypes
from dataclasses import dataclass
from typing import Optional

import torch

from .. import _native as N
from ._engine import probes
from .algorithm_base import Algorithm
from .keys import PRNGKey
from .lanczos import run_lanczos, tridiag_eigh

F_KIND = {"log": 0, "identity": 1, "exp": 2, "sqrt": 3, "reciprocal": 4}


@dataclass
class HipSLQ(Algorithm):
    """logdet by stochastic Lanczos quadrature with an explicit probe count.

    Args:
        num_probes: probes per estimate (the reference derives int(1/vtol^2), slq.py:74).
        max_iters: Lanczos steps per probe.
        tol: Lanczos stop tolerance (lanczos.py:263-268).
        rand: 'rademacher' or 'normal'.
        key: probe-stream seed (default PRNGKey(0), the reference's SLQ default).
        scale: 'norm' weights each probe by ||z||^2 (Hutchinson form, unary.py:58);
            'n' uses the dimension like slq.py:51 (equal for Rademacher probes).
    """
    num_probes: int = 32
    max_iters: int = 30
    tol: float = 1e-6
    rand: str = "rademacher"
    key: Optional[int] = None
    scale: str = "norm"
    pbar: bool = False


def slq_quadratures(A, Z, max_iters, tol, f_kind=0, eps=None, max_mode=0, scale=None, chunk=32):
    """Per-probe quadratures for probes Z (n, B) on the device. Returns (q (B,) fp64, run)."""
    run = run_lanczos(A, Z, max_iters, tol, chunk)
    ev, _, fr, st = tridiag_eigh(run, vectors=False)
    if int(st.max()) != 0:
        raise RuntimeError("tridiagonal eigensolve did not converge")
    nb, k = ev.shape
    eps = torch.finfo(A.dtype).eps if eps is None else eps
    q = torch.empty(nb, dtype=torch.float64, device=Z.device)
    sc = scale.to(torch.float64).contiguous() if scale is not None else None
    N.call("ck_slq_quadrature", nb, k, N.ptr(ev), N.ptr(fr), N.ptr(sc), f_kind, float(eps), max_mode, N.ptr(q),
           N.stream_ptr(Z.device))
    return q, run


def slq_logdet(A, alg):
    """Mean over probes of ||z||^2 sum_j tau_j^2 log(theta_j) (or n * ..., slq.py:51).
    Differentiable with respect to the operator's tensors: the backward is the rule of
    slq.py:10-31 on this estimator's own probes, (1/m) sum_b (A^-1 z_b)^T (dA/dtheta) z_b."""
    from ..autodiff import iterative_autograd
    return iterative_autograd(_slq_logdet_bwd)(_slq_logdet)(A, alg)


def _slq_logdet_bwd(res, grads, unflatten, alg):
    from ..autodiff import GradOperator, vjp_derivs
    from .cg import cg
    op_args, _ = res
    A = unflatten(op_args)
    n = A.shape[0]
    seed = PRNGKey(0) if alg.key is None else int(alg.key)
    Zb, _ = probes(alg.rand, n, alg.num_probes, seed, A.device)
    Z = Zb[:, :n].T.contiguous()
    solves, _ = cg(A, Z, tol=1e-6, max_iters=100)
    g = grads[0]
    d_solves = (1.0 / alg.num_probes) * g.to(solves.dtype) * solves

    def fun(*theta):
        return unflatten(theta) @ Z

    return (GradOperator(vjp_derivs(fun, op_args, d_solves)), )


def _slq_logdet(A, alg):
    n = A.shape[0]
    seed = PRNGKey(0) if alg.key is None else int(alg.key)
    Zb, _ = probes(alg.rand, n, alg.num_probes, seed, A.device)
    Z = Zb[:, :n].T
    if alg.scale == "n":
        sc = torch.full((alg.num_probes, ), float(n), dtype=torch.float64, device=A.device)
    else:
        sc = (Zb[:, :n].double()**2).sum(1)
    q, run = slq_quadratures(A, Z, alg.max_iters, alg.tol, 0, None, 0, sc)
    return q.mean(), {"iterations": run.info["iterations"], "per_probe": q}


def stochastic_lanczos_quad(A, fun, max_iters=100, tol=1e-5, vtol=0.1, pbar=False, key=None):
    """slq.py:55-75: Gaussian probes, num_samples = max(int(1/vtol^2), 1), estimate
    n * mean_b sum_j tau_bj^2 f(theta_bj) with the signed-max threshold of slq.py:48.
    Differentiable with respect to the operator's tensors by slq_bwd (slq.py:10-31), which,
    like the reference, assumes f = log.

    ``fun`` may be one of the names in F_KIND or a torch function applied on the device."""
    num = max(int(1 / vtol**2), 1)
    return slq_fwd(A, fun, num_samples=num, max_iters=max_iters, tol=tol, pbar=pbar, key=key)


def _gaussian_probes(A, num, key):
    n = A.shape[0]
    seed = PRNGKey(0) if key is None else int(key)
    Zb, _ = probes("normal", n, num, seed, A.device)
    return Zb[:, :n].T


def slq_bwd(res, grads, unflatten, *args, **kwargs):
    """slq.py:10-31: d logdet/d theta ~ (1/m) sum_j (A^-1 z_j)^T (dA/dtheta) z_j with m Gaussian
    probes from key (default PRNGKey(0)), solved by CG at tol 1e-6 with at most 100 steps,
    scaled by the output gradient."""
    from ..autodiff import GradOperator, vjp_derivs
    from .cg import cg
    op_args, _ = res
    num_samples = kwargs["num_samples"]
    A = unflatten(op_args)
    Z = _gaussian_probes(A, num_samples, kwargs.get("key"))
    solves, _ = cg(A, Z.contiguous(), tol=1e-6, max_iters=100)
    g = grads[0] if isinstance(grads, (tuple, list)) else grads
    d_solves = (1.0 / num_samples) * g.to(solves.dtype) * solves

    def fun(*theta):
        return unflatten(theta) @ Z

    return (GradOperator(vjp_derivs(fun, op_args, d_solves)), )


def _slq_fwd(A, fun, num_samples, max_iters, tol, pbar, key):
    n = A.shape[0]
    Z = _gaussian_probes(A, num_samples, key)
    sc = torch.full((num_samples, ), float(n), dtype=torch.float64, device=A.device)
    if isinstance(fun, str):
        q, _ = slq_quadratures(A, Z, max_iters, tol, F_KIND[fun], None, 1, sc)
        return q.mean()
    run = run_lanczos(A, Z, max_iters, tol)
    ev, _, fr, _ = tridiag_eigh(run, vectors=False)
    const = 10 * torch.finfo(A.dtype).eps * ev.max(dim=1, keepdim=True).values
    fv = torch.where(ev.abs() > const, fun(ev), torch.zeros_like(ev))
    return (n * (fr**2 * fv).sum(-1)).mean()


def slq_fwd(A, fun, num_samples, max_iters, tol, pbar, key):
    """slq.py:36-52 with the backward rule slq_bwd (keyword arguments as in the reference)."""
    from ..autodiff import iterative_autograd
    return iterative_autograd(slq_bwd)(_slq_fwd)(A, fun, num_samples=num_samples, max_iters=max_iters, tol=tol,
                                                   pbar=pbar, key=key)


_ = ctypes
