"""inv / solve dispatch (cola/linalg/inverse/inv.py:24-165) onto the engine."""
import numpy as np
import torch

from ..annotations import PSD, Unitary
from ..dispatch import dispatch
from ..ops import (BlockDiag, Diagonal, Identity, Kronecker, LinearOperator, Permutation, Product, ScalarMul, Triangular,
                   TriangularInv)
from .algorithm_base import Algorithm, Auto, IterativeOperatorWInfo
from .batched import BatchedInverse, HipBatchedCG
from .cg import HipCG
from .decompositions import LU, Cholesky, adjoint, cholesky, is_small, plu
from .gmres import HipGMRES
from .mixed import HipMixedCG


def solve(A, b, alg=Auto()):
    """inv.py:24-39: inv(A, alg) @ b."""
    return inv(A, alg) @ b


@dispatch.abstract
def inv(A: LinearOperator, alg: Algorithm = Auto()):
    """(Lazily) the inverse of A; ``inv(A, alg) @ b`` solves A x = b."""


@dispatch(precedence=-1)
def inv(A: LinearOperator, alg: HipCG):
    """inv.py:66-69."""
    assert A.isa(PSD), "CG only valid for PSD matrices, wrap in cola_amd.PSD if desired"
    return IterativeOperatorWInfo(A, alg)


@dispatch(precedence=-1)
def inv(A: LinearOperator, alg: HipMixedCG):
    """Engine algorithm: FP64-residual refinement around FP32 (P)CG."""
    assert A.isa(PSD), "CG only valid for PSD matrices, wrap in cola_amd.PSD if desired"
    return IterativeOperatorWInfo(A, alg)


@dispatch
def inv(A: LinearOperator, alg: HipGMRES):
    """inv.py:61-63."""
    return IterativeOperatorWInfo(A, alg)


@dispatch(precedence=-1)
def inv(A: LinearOperator, alg: Auto):
    """inv.py:73-92: N^2 <= 1e6 -> a dense Cholesky (PSD) or LU factorization; larger
    operators -> CG on the engine (PSD) or GMRES on the engine's Arnoldi kernels."""
    small = is_small(A)
    if A.isa(PSD):
        return inv(A, Cholesky() if small else HipCG(**alg.__dict__))
    return inv(A, LU() if small else HipGMRES(**alg.__dict__))


@dispatch(precedence=-1)
def inv(A: LinearOperator, alg: Cholesky):
    """inv.py:95-99: inv(L^*) inv(L) for A = L L^*."""
    assert A.isa(PSD), "Cholesky only valid for PSD matrices, wrap in cola_amd.PSD if desired"
    L = cholesky(A)
    return inv(adjoint(L), alg) @ inv(L, alg)


@dispatch(precedence=-1)
def inv(A: LinearOperator, alg: LU):
    """inv.py:101-104: inv(U) inv(L) inv(P) for A = P L U."""
    P, L, U = plu(A)
    return inv(U, alg) @ inv(L, alg) @ inv(P, alg)


@dispatch(cond=lambda A, *_: A.isa(Unitary))
def inv(A: LinearOperator, alg: Algorithm):
    """inv.py:108-110."""
    from ..fns import lazify
    return lazify(A.to_dense().T.conj())


@dispatch
def inv(A: Identity, alg: Algorithm):
    """inv.py:113-115."""
    return A


@dispatch
def inv(A: ScalarMul, alg: Algorithm):
    """inv.py:118-120."""
    return ScalarMul(1 / A.c, shape=A.shape, dtype=A.dtype, device=A.device)


@dispatch(cond=lambda A, *_: all([M.shape[-2] == M.shape[-1] for M in A.Ms]))
def inv(A: Product, alg: Algorithm):
    """inv.py:128-131."""
    return Product(*reversed([inv(M, alg) for M in A.Ms]))


@dispatch
def inv(A: Kronecker, alg: Algorithm):
    """inv.py:139-141: factor-wise."""
    return Kronecker(*[inv(M, alg) for M in A.Ms])


@dispatch
def inv(A: BlockDiag, alg: Algorithm):
    """inv.py:135-136: block by block."""
    return BlockDiag(*[inv(M, alg) for M in A.Ms], multiplicities=A.multiplicities)


@dispatch
def inv(A: BlockDiag, alg: HipBatchedCG):
    """Engine algorithm: every block in one batched CG (ck_bcg_*) instead of one loop per block."""
    return BatchedInverse(A, alg)


@dispatch
def inv(A: Diagonal, alg: Algorithm):
    """inv.py:144-146."""
    return Diagonal(1. / A.diag)


@dispatch
def inv(A: Permutation, alg: Algorithm):
    """inv.py:123-125."""
    return Permutation(torch.argsort(A.perm), A.dtype)


@dispatch
def inv(A: Triangular, alg: Algorithm):
    """inv.py:149-151: triangular solves (TriangularInv, inv.py:154-165)."""
    return TriangularInv(A)


_ = np
