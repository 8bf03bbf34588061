"""Column-block buffers: the engine's vector layout.

A block of B vectors of length n is a contiguous (B, ld) fp32 tensor, ld = n
rounded up to a multiple of 64 floats so every vector starts 256-byte aligned
(float4 loads never straddle vectors).  The reference's layout for the same
data is an (n, B) row-major array; ``to_blocks``/``from_blocks`` convert with the
engine's transpose kernel.
"""
import torch

from . import _native as N


def ld_of(n):
    return (n + 63) // 64 * 64


def empty_blocks(B, n, device, dtype=torch.float32, zero=False):
    ld = ld_of(n)
    t = (torch.zeros if zero else torch.empty)((B, ld), dtype=dtype, device=device)
    return t, ld


def to_blocks(X, ld=None):
    """(n, B) or (n,) tensor on the device -> contiguous (B, ld) fp32 block."""
    N.require_device(X, "right-hand side")
    if X.dim() == 1:
        X = X[:, None]
    n, B = X.shape
    ld = ld or ld_of(n)
    X = X.to(torch.float32)
    out = torch.empty((B, ld), dtype=torch.float32, device=X.device)
    if ld > n:
        out[:, n:].zero_()
    if B == 1 and X.is_contiguous():
        out[0, :n].copy_(X[:, 0])
        return out
    Xc = X.contiguous()
    N.call("ck_transpose_f32", n, B, N.ptr(Xc), B, N.ptr(out), ld, 1, N.stream_ptr(X.device))
    return out


def from_blocks(Xb, n):
    """(B, ld) block -> (n, B) row-major tensor."""
    B, ld = Xb.shape
    out = torch.empty((n, B), dtype=torch.float32, device=Xb.device)
    if B == 1:
        out[:, 0].copy_(Xb[0, :n])
        return out
    N.call("ck_transpose_f32", n, B, N.ptr(Xb), ld, N.ptr(out), B, 0, N.stream_ptr(Xb.device))
    return out
