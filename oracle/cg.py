"""Batched preconditioned CG restated from cola/linalg/inverse/cg.py (torch, CPU).

Layout follows the reference: right-hand sides are the columns of an (n, B) array.
Reductions are over axis -2 (``xnp.norm(.., axis=-2, keepdims=True)``).
"""
import time

import numpy as np
import torch

SMALL = 1e-40  # cg.py:8


def do_safe_div(num, denom):
    """cg.py:173-178: denominators with |d| < 1e-40 become 1e-40."""
    is_zero = denom.abs() < torch.tensor(SMALL, dtype=num.real.dtype)
    denom = torch.where(is_zero, torch.tensor(SMALL, dtype=denom.dtype), denom)
    return num / denom


def initialize(matmat, b, P, x0):
    """cg.py:122-130."""
    r0 = b - matmat(x0)
    z0 = P(r0)
    p0 = z0
    gamma0 = torch.sum(r0.conj() * z0, dim=-2, keepdim=True)
    alpha0 = torch.zeros_like(gamma0)
    beta0 = torch.zeros_like(gamma0)
    return (x0, 0, r0, p0, alpha0, beta0, gamma0)


def cond_fun(state, tol, max_iters):
    """cg.py:133-138: continue while any column's ||r|| > tol and k < max_iters."""
    _, k, r, *_ = state
    rs = torch.linalg.norm(r, dim=-2, keepdim=True)
    return bool(torch.any(rs > tol)) and (k < max_iters)


def update_alpha(gamma, p, Ap, has_converged):
    """cg.py:157-162."""
    denom = torch.sum(p.conj() * Ap, dim=-2, keepdim=True)
    alpha = do_safe_div(gamma, denom)
    return torch.where(has_converged, torch.zeros((), dtype=p.dtype), alpha)


def update_gamma_beta(r, z, gamma0, has_converged):
    """cg.py:165-170."""
    gamma1 = torch.sum(r.conj() * z, dim=-2, keepdim=True)
    beta = do_safe_div(gamma1, gamma0)
    beta = torch.where(has_converged, torch.zeros((), dtype=r.dtype), beta)
    return gamma1, beta


def take_cg_step(state, matmat, P):
    """cg.py:141-154."""
    x0, k, r0, p0, _, _, gamma0 = state
    eps = torch.tensor(SMALL, dtype=p0.real.dtype)
    has_converged = torch.linalg.norm(r0, dim=-2, keepdim=True) < eps
    Ap0 = matmat(p0)
    alpha = update_alpha(gamma0, p0, Ap0, has_converged)
    x1 = x0 + alpha * p0
    r1 = r0 - alpha * Ap0
    z1 = P(r1)
    gamma1, beta = update_gamma_beta(r1, z1, gamma0, has_converged)
    p1 = z1 + beta * p0
    return (x1, k + 1, r1, p1, alpha, beta, gamma1)


def while_loop_winfo(errorfn, cond, body, init_val):
    """cola/utils/torch_tqdm.py:7-71 without the progress bar.

    info['iterations'] counts cond evaluations; info['errors'] records errorfn at
    each evaluation plus once after the loop, with the first two entries dropped."""
    info = {'iterations': 0, 'errors': []}
    t0 = time.time()
    val = init_val
    while True:
        info['errors'].append(float(errorfn(val)))
        info['iterations'] += 1
        if not cond(val):
            break
        val = body(val)
    info['iteration_time'] = (time.time() - t0) / info['iterations']
    info['errors'].append(float(errorfn(val)))
    info['errors'] = np.array(info['errors'][2:])
    return val, info


def run_batched_cg(matmat, b, x0=None, max_iters=1000, tol=1e-6, P=None):
    """cg.py:94-119. Returns (x, r, k, info)."""
    if x0 is None:
        x0 = torch.zeros_like(b)
    if P is None:
        P = lambda v: v  # noqa: E731  (I_like(A), cg.py:65-66)
    mult = torch.linalg.norm(b, dim=-2, keepdim=True)
    b_norm = do_safe_div(b, mult)
    init_val = initialize(matmat, b_norm, P, x0)
    r0 = init_val[2]
    tol_b = tol * torch.linalg.norm(r0, dim=-2, keepdim=True) + tol

    state, info = while_loop_winfo(lambda s: torch.linalg.norm(s[2], dim=-2).mean(),
                                   lambda s: cond_fun(s, tol_b, max_iters),
                                   lambda s: take_cg_step(s, matmat, P), init_val)
    return state[0] * mult, state[2] * mult, state[1], info


def cg(matmat, rhs, x0=None, P=None, tol=1e-6, max_iters=5000):
    """cg.py:39-69: vector rhs is promoted to (n, 1); returns (soln, info)."""
    is_vector = rhs.dim() == 1
    if x0 is None:
        x0 = torch.zeros_like(rhs)
    if is_vector:
        rhs, x0 = rhs[:, None], x0[:, None]
    soln, _, _, info = run_batched_cg(matmat, rhs, x0, max_iters, tol, P)
    return (soln.reshape(-1) if is_vector else soln), info
