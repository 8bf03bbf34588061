"""Dense Cholesky and LU decompositions (cola/linalg/decompositions/decompositions.py:29-56,
148-199) and the small-operator branch of ``Auto`` they serve.

The reference's ``Auto`` sends an operator with N^2 <= 1e6 to a dense factorization instead
of an iterative method (inv.py:73-92, logdet.py:82-94): Cholesky for PSD operators, LU for the
others.  Here the factorizations run on the operator's device through ``torch.linalg`` (the
matrix is at most 1000 x 1000) and return the structured factors the reference returns:
``Triangular`` operators and a ``Permutation``, whose inverses and log-determinants have their
own rules.  Structured operators keep their structure (Identity, Diagonal, ScalarMul,
Kronecker and BlockDiag, as decompositions.py:157-199).
"""
from dataclasses import dataclass

import numpy as np
import torch

from ..dispatch import dispatch
from ..ops import BlockDiag, Diagonal, Identity, Kronecker, LinearOperator, Permutation, ScalarMul, Triangular
from .algorithm_base import Algorithm

# decompositions above this many entries go to the iterative methods (inv.py:79, logdet.py:85)
SMALL_ENTRIES = 1e6


def is_small(A):
    return bool(np.prod(A.shape) <= SMALL_ENTRIES)


@dataclass
class Cholesky(Algorithm):
    """A = L L^* with L lower triangular (decompositions.py:29-41)."""

    def __call__(self, A):
        return cholesky(A)


@dataclass
class LU(Algorithm):
    """A = P L U with P a permutation, L lower and U upper triangular (decompositions.py:44-56)."""

    def __call__(self, A):
        return plu(A)


@dispatch
def cholesky(A: LinearOperator):
    """decompositions.py:148-154: the lower factor of the dense matrix, as a Triangular."""
    return Triangular(torch.linalg.cholesky(A.to_dense()), lower=True)


@dispatch
def cholesky(A: Identity):
    """decompositions.py:157-159."""
    return A


@dispatch
def cholesky(A: Diagonal):
    """decompositions.py:162-164: the entrywise square root."""
    return Diagonal(torch.sqrt(A.diag))


@dispatch
def cholesky(A: ScalarMul):
    """decompositions.py:162-164: sqrt(c) I."""
    return ScalarMul(torch.sqrt(A.c), shape=A.shape, dtype=A.dtype, device=A.device)


@dispatch
def cholesky(A: Kronecker):
    """decompositions.py:167-170: the Kronecker product of the factors' Cholesky factors."""
    return Kronecker(*[cholesky(M) for M in A.Ms])


@dispatch
def cholesky(A: BlockDiag):
    """decompositions.py:173-175: block by block."""
    return BlockDiag(*[cholesky(M) for M in A.Ms], multiplicities=A.multiplicities)


def _perm_of(Pm):
    """Index form of a permutation matrix: (Pm v)_i = v_{perm[i]} (Permutation's convention)."""
    return torch.argmax(Pm.abs(), dim=1)


@dispatch
def plu(A: LinearOperator):
    """decompositions.py:178-186: A = P L U of the dense matrix with partial pivoting."""
    Pm, L, U = torch.linalg.lu(A.to_dense())
    return Permutation(_perm_of(Pm), A.dtype), Triangular(L, lower=True), Triangular(U, lower=False)


@dispatch
def plu(A: Identity):
    """decompositions.py:189-191."""
    return A, A, A


@dispatch
def plu(A: Diagonal):
    """decompositions.py:194-197: I, sqrt(D), sqrt(D)."""
    S = Diagonal(torch.sqrt(A.diag))
    return Identity(A.shape, A.dtype, A.device), S, S


@dispatch
def plu(A: ScalarMul):
    """decompositions.py:194-197: I, sqrt(c) I, sqrt(c) I."""
    S = ScalarMul(torch.sqrt(A.c), shape=A.shape, dtype=A.dtype, device=A.device)
    return Identity(A.shape, A.dtype, A.device), S, S


def adjoint(L):
    """L^* with the factor's structure kept (a Triangular's adjoint is triangular the other way;
    Kronecker and BlockDiag factor-wise; Diagonal, ScalarMul and Identity elementwise)."""
    if isinstance(L, Triangular):
        return Triangular(L.A.mH, lower=not L.lower)
    if isinstance(L, Kronecker):
        return Kronecker(*[adjoint(M) for M in L.Ms])
    if isinstance(L, BlockDiag):
        return BlockDiag(*[adjoint(M) for M in L.Ms], multiplicities=L.multiplicities)
    if isinstance(L, Diagonal):
        return Diagonal(L.diag.conj())
    if isinstance(L, ScalarMul):
        return ScalarMul(L.c.conj(), shape=L.shape, dtype=L.dtype, device=L.device)
    if isinstance(L, Identity):
        return L
    return L.H
