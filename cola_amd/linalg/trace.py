"""Hutchinson diagonal/trace estimation on the engine
(cola/linalg/trace/diagonal_estimation.py:34-58,158-210; diag_trace.py:22-147).

``HipHutch`` keeps cola's ``Hutch`` fields and estimator: blocks of
min(100, n) probes (the reference overrides ``bs`` with this, diagonal_estimation.py:176),
keys chained with ``next_key`` from ``PRNGKey(42)``, Gaussian or Rademacher
probes, the running sum / sum of squares of z * (A z), and the adaptive stop on
the mean relative standard error (:197-208).  Probes come from the engine's
counter-based streams (seeded by the key chain) instead of torch.randn.
Set ``bs_override=True`` to use ``bs`` probes per block as the field suggests.

``hutch_trace_csr`` is the engine's fused path for a CSR operator: packed
Rademacher probes (64 per 64-bit word) and a popcount contraction, so the
probe block is never stored and each nonzero costs one word gather.
"""
import ctypes
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from .. import _native as N
from .._blocks import ld_of
from ..dispatch import dispatch
from ..ops import BlockDiag, Dense, Diagonal, Identity, Kronecker, KronSum, LinearOperator, ScalarMul, Sum
from ._engine import apply_blocks, lower, probes, workspace
from .algorithm_base import Algorithm, Auto
from .keys import PRNGKey, next_key


@dataclass
class HipHutch(Algorithm):
    """Hutchinson estimator on the HIP engine (fields as cola's Hutch)."""
    tol: float = 3e-2
    max_iters: int = 10_000
    bs: int = 100
    rand: str = "normal"
    pbar: bool = False
    key: Optional[int] = None
    bs_override: bool = False

    def __call__(self, A, k=0):
        return hutchinson_diag_estimate(A, k, **self.__dict__)[0]


@dataclass
class Exact(Algorithm):
    """Exact diagonal from blocks of unit vectors (diagonal_estimation.py:13-31,117-128)."""
    bs: int = 100
    pbar: bool = False

    def __call__(self, A, k=0):
        # out[t] = A[t + max(0, -k), t + max(0, k)] from blocks of unit columns (exact_diag,
        # diagonal_estimation.py:117-128, with its chunk shifting written as an index map)
        n = A.shape[0]
        k = int(k)
        if abs(k) >= n:
            raise ValueError(f"diagonal {k} is outside a {n} x {n} operator")
        m = n - abs(k)
        r0, c0 = max(0, -k), max(0, k)
        bs = min(100, n)
        out = torch.empty(m, dtype=A.dtype, device=A.device)
        for t in range(0, m, bs):
            ts = torch.arange(t, min(t + bs, m), device=A.device)
            E = torch.zeros((n, ts.numel()), dtype=A.dtype, device=A.device)
            E[ts + c0, torch.arange(ts.numel(), device=A.device)] = 1.0
            out[ts] = (A @ E)[ts + r0, torch.arange(ts.numel(), device=A.device)]
        return out


def hutchinson_diag_estimate(A, k=0, bs=100, tol=3e-2, max_iters=10000, pbar=False, rand="normal", key=None,
                             bs_override=False):
    """diagonal_estimation.py:158-210. Returns (mean k-th diagonal (n - |k|,), info).

    For k != 0 the estimator of diagonal entry j is (A z)_j z_{j+k} (the reference's
    ``roll(z, -k)`` with the wrapped rows zeroed, :192-195): the accumulation kernel reads
    the probe block k rows further on (k > 0) or A z |k| rows further on (k < 0)."""
    assert tol > 1e-3, "tolerance chosen too high for stochastic diagonal estimation"
    assert rand in ["normal", "rademacher"], "rand must be 'normal' or 'rademacher'"
    n = A.shape[0]
    k = int(k)
    if abs(k) >= n:
        raise ValueError(f"diagonal {k} is outside a {n} x {n} operator")
    m = n - abs(k)
    dev = A.device
    if dev.type != "cuda":
        raise ValueError("the Krylov engine runs on the HIP device; A is on the CPU")
    bs = min(int(bs), n) if bs_override else min(100, n)
    key = PRNGKey(42) if key is None else int(key)
    aop = lower(A)
    work, wb = workspace([aop[0] if aop else None], bs, dev)
    ld = ld_of(n)
    dsum = torch.zeros(m, dtype=torch.float32, device=dev)
    dsq = torch.zeros(m, dtype=torch.float32, device=dev)
    AZ = torch.empty((bs, ld), dtype=torch.float32, device=dev)
    # row offsets of the two factors: (A z)_{j + a} * z_{j + c}, j < m
    a_off, z_off = (0, k) if k >= 0 else (-k, 0)
    s = N.stream_ptr(dev)
    errors = []
    t0 = time.time()
    i = 0
    while True:
        key = next_key(key)
        Z, _ = probes(rand, n, bs, key, dev)
        apply_blocks(A, aop, Z, AZ, n, bs, work, wb)
        N.call("ck_hutch_accumulate", m, bs, N.ptr(Z[:, z_off:]), ld, N.ptr(AZ[:, a_off:]), ld, N.ptr(dsum),
               N.ptr(dsq), s)
        i += 1
        cnt = i * bs
        mean = dsum / cnt
        stderr = torch.sqrt((dsq / cnt - mean**2) / cnt)
        err = float(torch.mean(stderr / torch.maximum(mean.abs(), torch.full_like(mean, 0.1))))
        errors.append(err)
        if not (i < max_iters and err > tol):
            break
    # torch_tqdm.py:38-62: one error per condition evaluation (the first, at zero blocks, is
    # undefined), the final error appended once more, and the first two entries dropped
    info = {"iterations": i + 1, "errors": np.array(errors[1:] + errors[-1:]),
            "iteration_time": (time.time() - t0) / (i + 1)}
    return dsum / (i * bs), info


def hutch_trace_csr(A, num_probes=64, seed=0):
    """(1/B) sum_b z_b^T A z_b for a CSR operator (with scale/shift) and B = 64*g
    Rademacher probes of the engine stream `seed`; FP64 result on the device."""
    got = lower(A)
    if got is None or got[0].kind != N.CK_OP_CSR_F32:
        raise ValueError("hutch_trace_csr needs a Sparse operator (optionally scaled/shifted) on the device")
    op = got[0]
    n = A.shape[0]
    groups = (num_probes + 63) // 64
    if groups * 64 != num_probes:
        raise ValueError("num_probes must be a multiple of 64 for packed probes")
    dev = A.device
    s = N.stream_ptr(dev)
    wb = N.q("ck_trace_workspace_bytes", n)
    key = (str(dev), wb, s.value)
    work = _TRACE_WORK.get(key)
    if work is None:
        # per-stream scratch of the block partials (reused: launches on one stream are ordered)
        work = _TRACE_WORK[key] = torch.empty((wb + 7) // 8, dtype=torch.float64, device=dev)
    out = torch.empty(1, dtype=torch.float64, device=dev)
    # the probe words (those of ck_rng_rademacher_bits(seed)) are recomputed inside the pass
    N.call("ck_trace_csr_rademacher_seeded", op, groups, ctypes.c_uint64(seed & (2**64 - 1)), N.ptr(out), N.ptr(work),
           wb, s)
    return out[0]


_TRACE_WORK = {}


@dispatch.abstract
def diag(A: LinearOperator, k: int = 0, alg: Algorithm = Auto()):
    """Diagonal of A (diag_trace.py:22-39)."""


@dispatch(precedence=-1)
def diag(A: LinearOperator, k: int, alg: Auto):
    """diag_trace.py:43-50."""
    tol = alg.__dict__.get("tol", 1e-6)
    if tol < 1 / np.sqrt(10 * np.prod(A.shape)):
        return diag(A, k, Exact())
    return diag(A, k, HipHutch(**alg.__dict__))


@dispatch(precedence=-1)
def diag(A: LinearOperator, k: int, alg: HipHutch | Exact):
    """diag_trace.py:53-55."""
    return alg(A, k)


@dispatch
def diag(A: Dense, k: int, alg: Algorithm):
    return torch.diagonal(A.A, k)


@dispatch
def diag(A: Identity, k: int, alg: Algorithm):
    return torch.ones(A.shape[0], dtype=A.dtype, device=A.device) if k == 0 else \
        torch.zeros(A.shape[0] - abs(k), dtype=A.dtype, device=A.device)


@dispatch
def diag(A: Diagonal, k: int, alg: Algorithm):
    return A.diag if k == 0 else torch.zeros(A.shape[0] - abs(k), dtype=A.dtype, device=A.device)


@dispatch
def diag(A: Sum, k: int, alg: Algorithm):
    return sum(diag(M, k, alg) for M in A.Ms)


@dispatch
def diag(A: ScalarMul, k: int, alg: Algorithm):
    return A.c * diag(Identity(A.shape, A.dtype, A.device), k, alg)


@dispatch
def diag(A: BlockDiag, k: int, alg: Algorithm):
    """diag_trace.py:87-91: the blocks' diagonals in order, multiplicities repeated (k = 0)."""
    assert k == 0, "only the main diagonal of a BlockDiag is formed"
    return torch.cat([diag(M, k, alg) for M, m in zip(A.Ms, A.multiplicities) for _ in range(m)])


@dispatch
def diag(A: KronSum, k: int, alg: Algorithm):
    """diag_trace.py:112-118: outer sum of the factors' diagonals (k = 0)."""
    assert k == 0
    ds = [diag(M, k, alg) for M in A.Ms]
    out = ds[0]
    for d in ds[1:]:
        out = (out[:, None] + d[None, :]).reshape(-1)
    return out


@dispatch
def diag(A: Kronecker, k: int, alg: Algorithm):
    """diag_trace.py:103-109."""
    assert k == 0
    ds = [diag(M, k, alg) for M in A.Ms]
    out = ds[0]
    for d in ds[1:]:
        out = (out[:, None] * d[None, :]).reshape(-1)
    return out


@dispatch.abstract
def trace(A: LinearOperator, alg: Algorithm = Auto()):
    """tr(A) (diag_trace.py:121-136)."""


@dispatch
def trace(A: LinearOperator, alg: Algorithm):
    """diag_trace.py:139-142."""
    assert A.shape[0] == A.shape[1], "Can't trace non square matrix"
    return diag(A, 0, alg).sum()


@dispatch
def trace(A: Kronecker, alg: Algorithm):
    """diag_trace.py:145-147."""
    out = 1.0
    for M in A.Ms:
        out = out * trace(M, alg)
    return out
