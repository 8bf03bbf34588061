"""f(A) through Lanczos on the engine (cola/linalg/unary/unary.py:37-60,94-334).

``LanczosUnary(A, f) @ V`` runs one batched Lanczos over V's columns, an FP64
tridiagonal eigensolve per column on the device, the zero-eigenvalue threshold
10*eps*max|theta| of unary.py:55-57 (eps of A's dtype), and assembles
f(A) v = ||v|| Q P (f(theta) * P[0, :]) (unary.py:58-59) with a batched basis
combination kernel.  ``f`` is applied to the FP64 Ritz values on the device.
"""
import torch

from .. import _native as N
from ..annotations import PSD, SelfAdjoint
from ..dispatch import dispatch
from ..ops import LinearOperator
from .algorithm_base import Algorithm, Auto
from .lanczos import HipLanczos, run_lanczos, tridiag_eigh


class LanczosUnary(LinearOperator):
    def __init__(self, A, f, max_iters=100, tol=1e-6, chunk=32, **_):
        self.device = A.device
        super().__init__(A.dtype, A.shape, annotations={SelfAdjoint})
        self.A, self.f = A, f
        self.max_iters, self.tol, self.chunk = max_iters, tol, chunk
        self.info = {}

    def _matmat(self, V):
        run = run_lanczos(self.A, V, self.max_iters, self.tol, self.chunk)
        self.info.update(run.info)
        ev, vecs, _, st = tridiag_eigh(run, vectors=True)
        if int(st.max()) != 0:
            raise RuntimeError("tridiagonal eigensolve did not converge")
        norms = torch.linalg.vector_norm(V.double(), dim=0)
        eps = torch.finfo(self.dtype).eps
        thresh = 10 * eps * ev.abs().max(dim=1, keepdim=True).values
        fv = torch.where(ev.abs() > thresh, self.f(ev), torch.zeros_like(ev))
        w = vecs[:, 0, :] * norms[:, None]  # (nb, k): P[0, :] * ||v||
        C = torch.bmm(vecs, (fv * w)[..., None])[..., 0].contiguous()  # (nb, k) = P (f * w)
        nb, k = C.shape
        out = torch.empty((nb, run.ld), dtype=torch.float32, device=V.device)
        N.call("ck_basis_combine_batched", run.n, nb, k, N.ptr(run.V[:, 1]), run.ld, run.ldb, N.ptr(C), C.stride(0),
               N.ptr(out), run.ld, N.stream_ptr(V.device))
        return out[:, :run.n].T.to(V.dtype)


@dispatch.abstract
def apply_unary(f, A: LinearOperator, alg: Algorithm = Auto()):
    """f(A) as a lazy operator (unary.py:94-108)."""


@dispatch(precedence=-1)
def apply_unary(f, A: LinearOperator, alg: Auto):
    """unary.py:114-131 restricted to the engine: self-adjoint operators use Lanczos."""
    if A.isa(SelfAdjoint) or A.isa(PSD):
        return apply_unary(f, A, HipLanczos(**alg.__dict__))
    raise NotImplementedError("the engine evaluates f(A) for self-adjoint operators (Lanczos)")


@dispatch(precedence=-1)
def apply_unary(f, A: LinearOperator, alg: HipLanczos):
    """unary.py:134-137."""
    assert A.isa(SelfAdjoint), "Lanczos only valid for SelfAdjoint, wrap in cola_amd.SelfAdjoint if desired"
    kw = dict(alg.__dict__)
    kw.pop("start_vector", None)
    return LanczosUnary(A, f, max_iters=kw.get("max_iters", 100), tol=kw.get("tol", 1e-6), chunk=kw.get("chunk", 32))


def log(A, alg=Auto()):
    """unary.py:249-262."""
    return apply_unary(torch.log, A, alg)


def exp(A, alg=Auto()):
    return apply_unary(torch.exp, A, alg)


def sqrt(A, alg=Auto()):
    return apply_unary(torch.sqrt, A, alg)


def isqrt(A, alg=Auto()):
    return apply_unary(lambda x: 1.0 / torch.sqrt(x), A, alg)


def pow(A, alpha, alg=Auto()):
    return apply_unary(lambda x: x**alpha, A, alg)
