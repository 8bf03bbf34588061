"""cola_amd: CoLA's Krylov engine (CG, Lanczos, Arnoldi, SLQ, Hutchinson) on MI355X.

Keeps cola's LinearOperator / annotation / cola.linalg API; the solvers run as
hand-written gfx950 HIP kernels behind a C ABI (include/cola_krylov.h).
"""
from . import annotations, io, linalg, ops
from .annotations import PSD, Hermitian, SelfAdjoint, Stiefel, Unitary
from .fns import add, densify, dot, kron, lazify, mul
from .ops import (BlockDiag, Dense, DenseBF16, Diagonal, I_like, Identity, Kernel, Kronecker, KronSum, LinearOperator,
                  Permutation, Product, Triangular, TriangularInv,
                  RBFKernel, ScalarMul, Sparse, StationaryKernelFn, Sum, Tridiagonal, WoodburyInverse, matern12,
                  matern32, matern52, rbf)

__version__ = "0.1.0"
