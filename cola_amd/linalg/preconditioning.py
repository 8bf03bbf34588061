"""Nystrom preconditioners (cola/linalg/preconditioning/preconditioners.py:87-186) built and
applied on the engine.

Construction, ``get_nys_approx`` (preconditioners.py:143-153), with the sketch Omega kept
as a feature-major (rank, ld) block:
  * Omega <- orth(Omega): two Cholesky-QR passes (``ck_gram_f64`` in FP64, a Cholesky of
    the rank x rank Gram matrix, ``ck_rows_gemm_f32``).  Only the sketch's range enters
    the result, so this basis serves where the reference takes Householder QR's Q;
  * Y = A Omega by the engine's operator apply; nu = eps ||Y||_F; Y += nu Omega;
  * C = chol(Omega^T Y) and B = Y C^-T;
  * the SVD of B from the eigenpairs (sigma^2, V) of B^T B (FP64 Gram, Jacobi on the
    device, ``ck_sym_eigh_f64``): U = B V diag(1/sigma), Lambda = max(sigma^2 - nu, 0),
    in descending order as the SVD returns them.
The preconditioner y = V + U diag(num/denom - 1) U^T V is the engine's low-rank op
(CK_OP_WOODBURY_INV with F = U, scale 1 and core -diag(num/denom - 1) in FP64), the same
kernels that apply the Woodbury preconditioners.  rank <= 64.
"""
import numpy as np
import torch

from .. import _native as N
from .._blocks import ld_of, to_blocks
from ..ops import LinearOperator, _op
from ._engine import apply_blocks, lower, workspace
from .keys import PRNGKey
from .learned import dense_sketch_bf16, gram_blocks, rows_gemm, spd_inverse

MAX_RANK = 64


def _cholqr_blocks(Xb, n, passes=2):
    """An orthonormal basis of the columns of the feature-major block Xb (k, ld)."""
    ld = Xb.stride(0)
    for _ in range(passes):
        G = gram_blocks(Xb, Xb, n)
        try:
            fac = spd_inverse(G, linvT32=True)
        except np.linalg.LinAlgError:  # rank-deficient sketch: a relative jitter keeps the factor finite
            fac = spd_inverse(G, jitter_rel=1e-10, linvT32=True)
        Q = torch.zeros_like(Xb)
        rows_gemm(Xb, 1, ld, fac["linvT32"], Q, 1, ld, n)
        Xb = Q
    return Xb


def sym_eigh(S, max_sweeps=40):
    """Eigenpairs of a small symmetric FP64 device matrix: (w descending, E with row j the
    eigenvector of w[j]), by ``ck_sym_eigh_f64``."""
    k = S.shape[0]
    if k > MAX_RANK:
        raise ValueError(f"the device eigensolver takes order <= {MAX_RANK}, got {k}")
    S = S.to(torch.float64).contiguous()
    dev = S.device
    w = torch.empty(k, dtype=torch.float64, device=dev)
    E = torch.empty((k, k), dtype=torch.float64, device=dev)
    sweeps = torch.zeros(1, dtype=torch.int32, device=dev)
    N.call("ck_sym_eigh_f64", 1, k, N.ptr(S), k, k * k, N.ptr(w), k, N.ptr(E), k, k * k, max_sweeps, N.ptr(sweeps),
           N.stream_ptr(dev))
    return w, E, sweeps


def _nys_factor(A, Omega, eps, bf16_sketch=False):
    """get_nys_approx on the engine; returns (Lambda FP64, U (n, rank) fp32, nu FP64).
    bf16_sketch: form A Omega on BF16 MFMA for a dense fp32 A (rank 32 or 64, n a multiple
    of 256) from the bf16-rounded basis, which then serves as Omega throughout."""
    N.require_device(Omega, "sketch")
    n, r = Omega.shape
    if r > MAX_RANK:
        raise ValueError(f"Nystrom rank must be <= {MAX_RANK}, got {r}")
    dev = Omega.device
    ld = ld_of(n)
    Qb = _cholqr_blocks(to_blocks(Omega.to(torch.float32)), n)
    lowered = lower(A)
    got = dense_sketch_bf16(lowered, Qb, n) if bf16_sketch else None
    if got is not None:
        Yb, Qb = got
    else:
        work, wb = workspace([lowered[0]] if lowered is not None else [], r, dev)
        Yb = torch.zeros_like(Qb)
        apply_blocks(A, lowered, Qb, Yb, n, r, work, wb)
    nu = eps * torch.sqrt(torch.diagonal(gram_blocks(Yb, Yb, n)).sum())
    Yb[:, :n] += nu.to(torch.float32) * Qb[:, :n]
    fac = spd_inverse(gram_blocks(Qb, Yb, n), linvT32=True)
    Bb = torch.zeros_like(Yb)
    rows_gemm(Yb, 1, ld, fac["linvT32"], Bb, 1, ld, n)
    w, E, _ = sym_eigh(gram_blocks(Bb, Bb, n))
    sigma = torch.sqrt(torch.clamp(w, min=0.0))
    inv_sigma = torch.where(sigma > 0, 1.0 / torch.where(sigma > 0, sigma, torch.ones_like(sigma)),
                            torch.zeros_like(sigma))
    W = (E.T * inv_sigma[None, :]).to(torch.float32)
    U = torch.empty((n, r), dtype=torch.float32, device=dev)
    rows_gemm(Bb, 1, ld, W, U, r, 1, n)
    Lam = torch.clamp(w - nu, min=0.0)
    return Lam, U, nu


def get_nys_approx(A, Omega, eps):
    """preconditioners.py:143-153: (Lambda (rank,), U (n, rank)) with A ~ U diag(Lambda) U^T."""
    Lam, U, _ = _nys_factor(A, Omega, eps)
    return Lam.to(Omega.dtype), U.to(Omega.dtype)


class _NystromApply(LinearOperator):
    """V + U diag(num/denom - 1) U^T V on the engine's low-rank op."""

    def _set_apply(self, U, num, denom):
        self.U = U.to(torch.float32).contiguous()
        self.subspace_num = num
        self.subspace_denom = denom
        scaling = (num / denom - 1.0).to(torch.float64)
        self._core = torch.diag(-scaling).contiguous()

    @property
    def subspace_scaling(self):
        return (self.subspace_num / self.subspace_denom - 1.0)[:, None]

    def hip_op(self):
        U = self.U
        if not (U.is_cuda and U.shape[1] <= 128):
            return None
        op = _op(N.CK_OP_WOODBURY_INV, U.shape[0], scale=1.0)
        op.a, op.lda, op.aux, op.rank = U.data_ptr(), U.stride(0), self._core.data_ptr(), U.shape[1]
        return op, [U, self._core]

    def _matmat(self, V):
        if self._use_hip(V):
            return self._hip_apply(V)
        U = self.U.to(torch.float64)
        sc = self.subspace_scaling.to(torch.float64)
        return (U @ (sc * (U.T @ V.to(torch.float64))) + V.to(torch.float64)).to(V.dtype)


class NystromPrecond(_NystromApply):
    """preconditioners.py:87-124.

    Args:
        A (LinearOperator): positive definite (n, n) operator on the device.
        rank (int): rank of the Nystrom approximation (<= 64).
        mu (float): regularisation of the systems (A + mu) x = b the preconditioner serves.
        eps (float): shift used in the construction.
        adjust_mu (bool): scale mu by the largest captured eigenvalue.
        key (int): seed of the Gaussian sketch (default PRNGKey(42)), drawn as the
            reference's torch backend draws it (torch_fns.py:233-241).
        bf16_sketch (bool): form A Omega on BF16 MFMA when A is dense fp32 (engine option;
            the reference multiplies in the operator's dtype).
    """

    def __init__(self, A, rank, mu=1e-7, eps=1e-8, adjust_mu=True, key=None, bf16_sketch=False):
        self.device = A.device
        super().__init__(A.dtype, A.shape)
        self.bf16_sketch = bf16_sketch
        key = PRNGKey(42) if key is None else key
        g = torch.Generator(device=A.device).manual_seed(int(key))
        Omega = torch.randn(A.shape[0], rank, generator=g, dtype=torch.float32, device=A.device)
        self._create_approx(A=A, Omega=Omega, mu=mu, eps=eps, adjust_mu=adjust_mu)

    def _create_approx(self, A, Omega, mu, eps, adjust_mu):
        """preconditioners.py:111-120."""
        Lam, U, _ = _nys_factor(A, Omega, eps, getattr(self, "bf16_sketch", False))
        self.Lambda = Lam.to(torch.float32)
        self.adjusted_mu = amu = mu * Lam.max() if adjust_mu else torch.tensor(float(mu), dtype=torch.float64,
                                                                                 device=Lam.device)
        num = Lam.min() + amu
        self.preconditioned_eigmax = num
        self.preconditioned_eigmin = amu
        self._set_apply(U, num, Lam + amu)


class NystromPrecondLazy(_NystromApply):
    """preconditioners.py:127-136: a Nystrom operator given by U, num and denom."""

    def __init__(self, dtype, shape, U, subspace_num, subspace_denom):
        self.device = U.device
        super().__init__(dtype, shape)
        self._set_apply(U, subspace_num, subspace_denom)


def sqrt(A):
    """preconditioners.py:165-171."""
    if not isinstance(A, _NystromApply):
        raise TypeError("sqrt here takes a Nystrom preconditioner; use cola_amd.linalg.sqrt for operators")
    return NystromPrecondLazy(A.dtype, A.shape, A.U.clone(), torch.sqrt(A.subspace_num), torch.sqrt(A.subspace_denom))


def inverse(A):
    """preconditioners.py:180-186: num and denom swapped."""
    if not isinstance(A, _NystromApply):
        raise TypeError("inverse here takes a Nystrom preconditioner; use cola_amd.linalg.inv for operators")
    return NystromPrecondLazy(A.dtype, A.shape, A.U.clone(), A.subspace_denom.clone(), A.subspace_num.clone())

