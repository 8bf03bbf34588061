"""Operator algebra (cola/fns.py:32-252): lazify, densify, dot, add, mul, kron."""
import torch

from .ops import Dense, Diagonal, Identity, Kronecker, LinearOperator, Product, ScalarMul, Sum


def lazify(A):
    """fns.py:32-38."""
    return A if isinstance(A, LinearOperator) else Dense(A)


def densify(A):
    """fns.py:41-47."""
    return A.to_dense() if isinstance(A, LinearOperator) else A


def dot(A, B):
    """fns.py:63-95."""
    if isinstance(B, Identity):
        return A
    if isinstance(A, Identity):
        return B
    a = A.Ms if isinstance(A, Product) else (A, )
    b = B.Ms if isinstance(B, Product) else (B, )
    return Product(*(a + b))


def add(A, B):
    """fns.py:98-122."""
    A, B = lazify(A), lazify(B)
    a = A.Ms if isinstance(A, Sum) else (A, )
    b = B.Ms if isinstance(B, Sum) else (B, )
    return Sum(*(a + b))


def mul(A, c):
    """fns.py:125-137: c*A is Product(ScalarMul(c), A); scalars merge."""
    if isinstance(c, torch.Tensor):
        c = c.item() if c.numel() == 1 else c
    if isinstance(A, ScalarMul):
        return ScalarMul(A.c * c, A.shape, A.dtype, A.device)
    S = ScalarMul(c, (A.shape[-2], A.shape[-2]), A.dtype, A.device)
    return Product(S, A)


def kron(A, B):
    """fns.py:196-221."""
    A, B = lazify(A), lazify(B)
    if isinstance(A, Diagonal) and isinstance(B, Diagonal):
        return Diagonal((A.diag[:, None] * B.diag[None, :]).reshape(-1))
    a = A.Ms if isinstance(A, Kronecker) else (A, )
    b = B.Ms if isinstance(B, Kronecker) else (B, )
    return Kronecker(*(a + b))
