"""HipArnoldi: Arnoldi factorisation A Q = Q H (+ h e^T) on the engine.

Drop-in for cola's ``Arnoldi`` (cola/linalg/decompositions/decompositions.py:60-88)
and ``arnoldi``/``arnoldi_eigs`` (cola/linalg/decompositions/arnoldi.py:35-62,166-205,
289-335).  ``mgs=True`` keeps the reference's modified Gram-Schmidt order: one
dependent dot/axpy kernel per basis column, each also forming the partial sums
of the next column's dot.  ``mgs=False`` uses two classical Gram-Schmidt passes
(two batched GEMV sweeps each), which is faster for long bases and agrees with
MGS to rounding.  Returns Q (n, m+1) and H (m+1, m) like the reference (the
reference's eig path drops the last row/column, arnoldi.py:58).
"""
import time
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np
import torch

from .. import _native as N
from .._blocks import ld_of, to_blocks
from ..annotations import Stiefel
from ..ops import Dense
from ._engine import lower, read_ints, workspace
from .algorithm_base import Algorithm
from .lanczos import default_start


@dataclass
class HipArnoldi(Algorithm):
    """Arnoldi decomposition on the HIP engine (fields as cola's Arnoldi, plus mgs/chunk)."""
    start_vector: Any = None
    max_iters: int = 1_000
    tol: float = 1e-6
    pbar: bool = False
    key: Optional[int] = None
    mgs: bool = True
    chunk: int = 32

    def __call__(self, A):
        return arnoldi(A, self.start_vector, self.max_iters, self.tol, self.pbar, self.key, self.mgs, self.chunk)


def run_arnoldi(A, start, max_iters, tol, mgs=True, chunk=32):
    """Returns (Q (nb, m+1, ld) basis, H (nb, m+1, m) fp32, steps, info)."""
    N.require_device(start, "start vector")
    n = A.shape[0]
    m = min(int(max_iters), n)
    if start.dim() == 1:
        start = start[:, None]
    nb = start.shape[1]
    dev = start.device
    ld = ld_of(n)
    ldb = (m + 1) * ld
    Q = torch.zeros((nb, m + 1, ld), dtype=torch.float32, device=dev)
    H = torch.zeros((nb, m + 1, m), dtype=torch.float32, device=dev)
    state = torch.zeros(N.q("ck_arnoldi_state_bytes", nb, m), dtype=torch.uint8, device=dev)
    sb = N.q("ck_arnoldi_scratch_bytes", n, nb, m)
    scratch = torch.empty((sb + 3) // 4, dtype=torch.float32, device=dev)
    aop = lower(A)
    work, wb = workspace([aop[0] if aop else None], nb, dev)
    s = N.stream_ptr(dev)
    t0 = time.time()
    S = to_blocks(start, ld)
    N.call("ck_arnoldi_start", n, nb, N.ptr(S), ld, N.ptr(Q), ld, ldb, N.ptr(H), N.ptr(state), m, N.ptr(scratch), s)
    step = 0
    idx, done = read_ints(state, 2, dev)
    while not done and step < m:
        nst = min(int(chunk), m - step)
        if aop is not None:
            N.call("ck_arnoldi_iterate", aop[0], n, nb, m, N.ptr(Q), ld, ldb, N.ptr(H), N.ptr(state), float(tol), step,
                   nst, int(mgs), N.ptr(scratch), N.ptr(work), wb, s)
        else:
            for st in range(step, step + nst):
                qv = torch.as_strided(Q, (n, nb), (1, ldb), st * ld)
                W = A._matmat(qv)
                torch.as_strided(Q, (nb, n), (ldb, 1), (st + 1) * ld).copy_(W.T)
                N.call("ck_arnoldi_post", n, nb, m, N.ptr(Q), ld, ldb, N.ptr(H), N.ptr(state), float(tol), st,
                       int(mgs), N.ptr(scratch), s)
        step += nst
        idx, done = read_ints(state, 2, dev)
    errs = state[N.CTRL_BYTES:N.CTRL_BYTES + 4 * (idx + 1)].view(torch.float32).cpu().numpy().astype(np.float64)
    info = {"iterations": idx + 1, "errors": np.concatenate([errs, errs[-1:]])[2:],
            "iteration_time": (time.time() - t0) / (idx + 1)}
    return Q, H, idx, info


def arnoldi(A, start_vector=None, max_iters=100, tol=1e-7, pbar=False, key=None, mgs=True, chunk=32):
    """arnoldi.py:166-205: (Stiefel(Dense(Q)), Dense(H), info), Q (n, m+1), H (m+1, m)."""
    if start_vector is None:
        start_vector = default_start(A, key)[:, 0]
    n = A.shape[0]
    Qb, H, _, info = run_arnoldi(A, start_vector, max_iters, tol, mgs, chunk)
    if start_vector.dim() == 1:
        return Stiefel(Dense(Qb[0, :, :n].T)), Dense(H[0]), info
    return Qb[:, :, :n].transpose(1, 2), H, info


def arnoldi_eigs(A, start_vector=None, max_iters=100, tol=1e-7, pbar=False, key=None, mgs=True, chunk=32):
    """arnoldi.py:35-62: eig(H[:-1]) (complex) and eigenvectors Q[:, :-1] @ vs.

    The small m x m non-symmetric eigenproblem is solved by LAPACK on the host
    (torch.linalg.eig); the engine produces Q and H."""
    if start_vector is None:
        start_vector = default_start(A, key)[:, 0]
    n = A.shape[0]
    Qb, H, _, info = run_arnoldi(A, start_vector, max_iters, tol, mgs, chunk)
    Hm = H[0, :-1].cpu().double()
    eigvals, vs = torch.linalg.eig(Hm)
    Q = Qb[0, :-1, :n].T
    vs_d = vs.to(torch.complex64).to(Q.device)
    V = Q.to(torch.complex64) @ vs_d
    return eigvals.to(torch.complex64).to(Q.device), Dense(V), info
