"""Shared host-side plumbing of the Krylov drivers: operator application on
column blocks, control-block reads, workspace sizing and probe generation."""
import ctypes

import torch

from .. import _native as N
from .._blocks import ld_of


def lower(A):
    """The ck_op descriptor of A (with keepalive tensors), or None for a generic operator."""
    got = A.hip_op()
    return got if got is not None else None


def workspace(ops, nrhs, device):
    """One scratch tensor big enough for every op in `ops` (they run one after another)."""
    wb = 0
    for op in ops:
        if op is not None:
            wb = max(wb, N.q("ck_op_workspace_bytes", op, nrhs))
    if wb == 0:
        return None, 0
    return torch.empty((wb + 3) // 4, dtype=torch.float32, device=device), wb


def apply_blocks(A, lowered, Xb, Yb, n, nrhs, work, wb, ldx=None, ldy=None):
    """Yb[:, :n] = A @ Xb[:, :n] for column blocks (engine kernel when lowered, else A._matmat)."""
    ldx = ldx or Xb.stride(0)
    ldy = ldy or Yb.stride(0)
    if lowered is not None:
        N.call("ck_op_apply", lowered[0], nrhs, N.ptr(Xb), ldx, N.ptr(Yb), ldy, N.ptr(work), wb,
               N.stream_ptr(Xb.device))
        return
    Xv = torch.as_strided(Xb, (n, nrhs), (1, ldx))
    Y = A._matmat(Xv)
    torch.as_strided(Yb, (nrhs, n), (ldy, 1)).copy_(Y.T)


def read_ints(state, count, device):
    """Synchronous read of the first `count` int32 of a device control block."""
    host = torch.empty(count, dtype=torch.int32, pin_memory=True)
    host.copy_(state[:4 * count].view(torch.int32), non_blocking=True)
    torch.cuda.current_stream(device).synchronize()
    return host.tolist()


def probes(kind, n, num, seed, device):
    """(num, ld) block of probes from the engine's counter-based streams."""
    ld = ld_of(n)
    Z = torch.empty((num, ld), dtype=torch.float32, device=device)
    if ld > n:
        Z[:, n:].zero_()
    fn = {"rademacher": "ck_rng_rademacher_f32", "normal": "ck_rng_gaussian_f32",
          "gaussian": "ck_rng_gaussian_f32"}[kind]
    N.call(fn, n, num, ctypes.c_uint64(seed & (2**64 - 1)), N.ptr(Z), ld, N.stream_ptr(device))
    return Z, ld
