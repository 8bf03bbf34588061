"""Linear algebra on the engine: inv/solve, eig, logdet, trace/diag, unary functions.

``CG``, ``Lanczos``, ``Arnoldi`` and ``Hutch`` name the engine's algorithms, so
code written against cola's names runs on the HIP kernels unchanged."""
from .algorithm_base import Algorithm, Auto, IterativeOperatorWInfo
from .arnoldi import HipArnoldi, arnoldi, arnoldi_eigs
from .batched import BatchedMixedSolver, BatchedPrecond, BatchedSystems, HipBatchedCG, autotune
from .cg import HipCG, cg, run_batched_cg
from .decompositions import LU, Cholesky, cholesky, plu
from .eig import eig, eigmax, eigmin, get_slice
from .gmres import HipGMRES, gmres
from .inv import inv, solve
from .keys import PRNGKey, next_key
from .lanczos import HipLanczos, lanczos, lanczos_eigs
from .learned import (LearnedPreconditioner, PivotedCholeskyPreconditioner, PointMLP, make_training_instance,
                      rbf_matrix, select_preconditioner, train)
from .logdet import logdet, slogdet
from .mixed import HipMixedCG, KroneckerEigInverse, MixedSolver, mixed_cg, solve_concurrent
from .preconditioning import NystromPrecond, NystromPrecondLazy, get_nys_approx
from .slq import HipSLQ, stochastic_lanczos_quad
from .trace import Exact, HipHutch, diag, hutch_trace_csr, hutchinson_diag_estimate, trace
from .unary import LanczosUnary, apply_unary, exp, isqrt, log, pow, sqrt

CG = HipCG
Lanczos = HipLanczos
Arnoldi = HipArnoldi
Hutch = HipHutch
GMRES = HipGMRES

__all__ = ["Algorithm", "Auto", "IterativeOperatorWInfo", "Cholesky", "LU", "cholesky", "plu", "HipCG", "CG", "cg", "HipGMRES", "GMRES", "gmres", "run_batched_cg", "inv", "solve",
           "HipLanczos", "Lanczos", "lanczos", "lanczos_eigs", "HipArnoldi", "Arnoldi", "arnoldi", "arnoldi_eigs",
           "eig", "eigmax", "eigmin", "get_slice", "logdet", "slogdet", "HipSLQ", "stochastic_lanczos_quad",
           "HipHutch", "Hutch", "Exact", "diag", "trace", "hutch_trace_csr", "hutchinson_diag_estimate",
           "HipMixedCG", "KroneckerEigInverse", "MixedSolver", "mixed_cg", "solve_concurrent", "LanczosUnary", "apply_unary", "log", "exp", "sqrt", "isqrt", "pow", "PRNGKey", "next_key",
           "PointMLP", "LearnedPreconditioner", "PivotedCholeskyPreconditioner", "select_preconditioner",
           "make_training_instance", "train", "rbf_matrix",
           "HipBatchedCG", "BatchedSystems", "BatchedMixedSolver", "BatchedPrecond", "autotune",
           "NystromPrecond", "NystromPrecondLazy", "get_nys_approx"]
