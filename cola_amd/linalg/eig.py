"""eig dispatch (cola/linalg/eig/eigs.py:19-182) onto the engine."""
import numpy as np
import torch

from ..annotations import SelfAdjoint, Unitary
from ..dispatch import dispatch
from ..ops import Dense, Diagonal, Identity, LinearOperator
from .algorithm_base import Algorithm, Auto
from .arnoldi import HipArnoldi, arnoldi_eigs
from .lanczos import HipLanczos, lanczos_eigs


def get_slice(num, which):
    """cola/linalg/decompositions/decompositions.py:214-224."""
    if num == -1:
        raise ValueError(f"Number of eigenvalues {num} must be explicitly specified")
    if which == "SM":
        return slice(0, num, None)
    if which == "LM":
        return slice(-1 if num is None else -num, None, None)
    raise NotImplementedError(f"which={which} is not implemented")


@dispatch.abstract
def eig(A: LinearOperator, k: int, which: str = "LM", alg: Algorithm = Auto()):
    """Eigenvalues and eigenvectors (eigs.py:19-41)."""


@dispatch(precedence=-1)
def eig(A: LinearOperator, k: int, which: str, alg: Auto):
    """eigs.py:76-96 restricted to the engine: Lanczos when self-adjoint, else Arnoldi."""
    if A.isa(SelfAdjoint):
        return eig(A, k, which, HipLanczos(**alg.__dict__))
    return eig(A, k, which, HipArnoldi(**alg.__dict__))


@dispatch
def eig(A: LinearOperator, k: int, which: str, alg: HipLanczos):
    """eigs.py:106-111: only the requested Ritz vectors are formed."""
    assert A.isa(SelfAdjoint), "Lanczos only valid for SelfAdjoint, wrap in cola_amd.SelfAdjoint if desired"
    eig_slice = get_slice(k, which)
    kw = dict(alg.__dict__)
    kw.pop("pbar", None)
    vals, vecs, _ = lanczos_eigs(A, kw.pop("start_vector"), kw["max_iters"], kw["tol"], key=kw["key"],
                                 eig_slice=eig_slice, chunk=kw["chunk"], reorth=kw.get("reorth", "full"),
                                 block=kw.get("block", 16))
    return vals, vecs


@dispatch
def eig(A: LinearOperator, k: int, which: str, alg: HipArnoldi):
    """eigs.py:99-103."""
    eig_slice = get_slice(k, which)
    vals, vecs, _ = arnoldi_eigs(A, alg.start_vector, alg.max_iters, alg.tol, key=alg.key, mgs=alg.mgs,
                                 chunk=alg.chunk)
    return vals[eig_slice], Dense(vecs.to_dense()[:, eig_slice])


@dispatch
def eig(A: Identity, k: int, which: str, alg: Algorithm):
    eig_slice = get_slice(k, which)
    vals = torch.ones(A.shape[0], dtype=A.dtype, device=A.device)
    return vals[eig_slice], Unitary(Dense(torch.eye(A.shape[0], dtype=A.dtype, device=A.device)[:, eig_slice]))


@dispatch
def eig(A: Diagonal, k: int, which: str, alg: Algorithm):
    eig_slice = get_slice(k, which)
    idx = torch.argsort(A.diag)
    vals = A.diag[idx]
    vecs = torch.eye(A.shape[0], dtype=A.dtype, device=A.device)[:, idx]
    return vals[eig_slice], Unitary(Dense(vecs[:, eig_slice]))


def eigmax(A, alg=Auto()):
    """eigs.py:44-57."""
    vals, _ = eig(A, 1, "LM", alg)
    return vals[0]


def eigmin(A, alg=Auto()):
    """eigs.py:60-73."""
    vals, _ = eig(A, 1, "SM", alg)
    return vals[0]


_ = np
