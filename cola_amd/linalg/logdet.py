"""logdet / slogdet on the engine (cola/linalg/logdet/logdet.py:29-184).

Iterative rule (logdet.py:111-117): log(A) through Lanczos, then its trace by
Hutchinson; plus the structural rules for Kronecker, Diagonal, Identity,
ScalarMul, square Products, BlockDiag, Triangular and Permutation.  ``HipSLQ`` is the engine's direct stochastic
Lanczos quadrature log-determinant (explicit probe count, FP64 quadrature)."""
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from ..annotations import PSD
from ..dispatch import dispatch
from ..ops import BlockDiag, Diagonal, Identity, Kronecker, LinearOperator, Permutation, Product, ScalarMul, Triangular
from .algorithm_base import Algorithm, Auto
from .decompositions import LU, Cholesky, cholesky, is_small, plu
from .lanczos import HipLanczos
from .slq import HipSLQ, slq_logdet
from .trace import HipHutch, trace
from .unary import log


def logdet(A, log_alg=Auto(), trace_alg=Auto()):
    """logdet.py:29-49."""
    _, ld = slogdet(A, log_alg, trace_alg)
    return ld


@dispatch.abstract
def slogdet(A: LinearOperator, log_alg: Algorithm = Auto(), trace_alg: Algorithm = Auto()):
    """(sign, log|det|) with det(A) = sign * exp(logdet) (logdet.py:52-78)."""


@dispatch(precedence=-1)
def slogdet(A: LinearOperator, log_alg: Auto, trace_alg: Algorithm):
    """logdet.py:82-94: N^2 <= 1e6 -> a dense Cholesky (PSD) or LU factorization; larger PSD
    operators -> log(A) by Lanczos on the engine (the reference's Arnoldi branch for large
    non-PSD operators is not built: NotImplementedError)."""
    small = is_small(A)
    if small:
        return slogdet(A, Cholesky() if A.isa(PSD) else LU(), trace_alg)
    if A.isa(PSD):
        return slogdet(A, HipLanczos(**log_alg.__dict__), trace_alg)
    raise NotImplementedError("log-determinants of large non-PSD operators (Arnoldi) are not built")


@dispatch(precedence=-1)
def slogdet(A: LinearOperator, log_alg: Cholesky, trace_alg: Algorithm):
    """logdet.py:97-102: twice the log-determinant of the Cholesky factor."""
    assert A.isa(PSD), "Cholesky only valid for PSD matrices, wrap in cola_amd.PSD if necessary"
    sign, ld = slogdet(cholesky(A), log_alg, trace_alg)
    return sign * torch.conj(torch.as_tensor(sign)), 2 * ld


@dispatch(precedence=-1)
def slogdet(A: LinearOperator, log_alg: LU, trace_alg: Algorithm):
    """logdet.py:105-108: the product P L U's log-determinant (the permutation contributes
    (1, 0), as in the reference: the sign ignores its parity)."""
    P, L, U = plu(A)
    return slogdet(P @ L @ U, log_alg, trace_alg)


@dispatch(precedence=-1)
def slogdet(A: LinearOperator, log_alg: HipLanczos, trace_alg: Algorithm):
    """logdet.py:111-117: trace(log(A))."""
    logA = log(A, log_alg)
    tr = trace(logA, trace_alg if not isinstance(trace_alg, Auto) else HipHutch(**trace_alg.__dict__))
    mag = torch.abs(tr)
    return tr / mag, mag


@dispatch(precedence=-1)
def slogdet(A: LinearOperator, log_alg: HipSLQ, trace_alg: Algorithm):
    """Stochastic Lanczos quadrature in one batched Lanczos run (engine algorithm)."""
    est, _ = slq_logdet(A, log_alg)
    mag = torch.abs(est)
    return est / mag, mag


@dispatch(cond=lambda A, *_: all([(Ai.shape[-2] == Ai.shape[-1]) for Ai in A.Ms]))
def slogdet(A: Product, log_alg: Algorithm, trace_alg: Algorithm):
    """logdet.py:121-124."""
    signs, lds = zip(*[slogdet(Ai, log_alg, trace_alg) for Ai in A.Ms])
    out_s = 1.0
    for s_ in signs:
        out_s = out_s * s_
    return out_s, sum(lds)


@dispatch
def slogdet(A: Identity, log_alg: Algorithm, trace_alg: Algorithm):
    zero = torch.zeros((), dtype=A.dtype, device=A.device)
    return 1. + zero, zero


@dispatch
def slogdet(A: ScalarMul, log_alg: Algorithm, trace_alg: Algorithm):
    c = A.c
    return c / torch.abs(c), torch.log(torch.abs(c))


@dispatch
def slogdet(A: Diagonal, log_alg: Algorithm, trace_alg: Algorithm):
    mag = torch.abs(A.diag)
    return torch.prod(A.diag / mag), torch.sum(torch.log(mag))


@dispatch
def slogdet(A: Kronecker, log_alg: Algorithm, trace_alg: Algorithm):
    """logdet.py:150-158: logdet(A1 x ... x Ak) = sum_i (N / n_i) logdet(A_i)."""
    signs, lds = zip(*[slogdet(Ai, log_alg, trace_alg) for Ai in A.Ms])
    sizes = [Ai.shape[-1] for Ai in A.Ms]
    prod = int(np.prod(sizes))
    out_s = 1.0
    for s_, n_ in zip(signs, sizes):
        out_s = out_s * s_**(prod / n_)
    return out_s, sum(ld_ * prod / n_ for ld_, n_ in zip(lds, sizes))


@dispatch
def slogdet(A: BlockDiag, log_alg: Algorithm, trace_alg: Algorithm):
    """logdet.py:161-167: sum of the blocks' log-determinants, weighted by multiplicity."""
    signs, lds = zip(*[slogdet(Ai, log_alg, trace_alg) for Ai in A.Ms])
    out_s = 1.0
    for s_, m in zip(signs, A.multiplicities):
        out_s = out_s * s_**m
    return out_s, sum(ld_ * m for ld_, m in zip(lds, A.multiplicities))


@dispatch
def slogdet(A: Triangular, log_alg: Algorithm, trace_alg: Algorithm):
    """logdet.py:170-176: the diagonal's product."""
    d = torch.diagonal(A.A)
    mag = torch.abs(d)
    return torch.prod(d / mag), torch.sum(torch.log(mag))


@dispatch
def slogdet(A: Permutation, log_alg: Algorithm, trace_alg: Algorithm):
    """logdet.py:179-184: (1, 0), as the reference returns (the permutation's parity is not
    counted there either)."""
    zero = torch.zeros((), dtype=A.dtype, device=A.device)
    return 1. + zero, zero
