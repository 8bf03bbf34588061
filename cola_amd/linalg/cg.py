"""HipCG: batched (preconditioned) conjugate gradients on the engine.

Drop-in for cola's ``CG`` (cola/linalg/inverse/cg.py:11-36): same fields, same
``(soln, info)`` return, same arithmetic and stopping rule (run_batched_cg,
cg.py:94-178).  The loop condition lives on the device: the host enqueues
``chunk`` steps at a time (through ``ck_cg_iterate`` when the operator lowers
to an engine kernel, otherwise step by step around ``A._matmat``) and reads
the 8-byte control block once per chunk instead of syncing twice per step
(cola/utils/torch_tqdm.py:42,88).
"""
import time
from dataclasses import dataclass

import numpy as np
import torch

from .. import _native as N
from .._blocks import from_blocks, ld_of, to_blocks
from ..autodiff import GradOperator, iterative_autograd, vjp_derivs
from ._engine import apply_blocks, lower, read_ints, workspace
from .algorithm_base import Algorithm


@dataclass
class HipCG(Algorithm):
    """Conjugate gradients on the HIP engine.

    Args:
        tol (float): relative tolerance (threshold tol*||r0|| + tol on the
            normalised residual, cg.py:101).
        max_iters (int): step cap.
        pbar (bool): accepted for compatibility; no progress bar is drawn.
        x0 (Tensor): (n,) or (n, b) initial guess.
        P (LinearOperator): preconditioner, applied as z = P r (an approximation of A^-1).
        chunk (int): steps enqueued between two reads of the device loop flag.
        fused (bool): run small dense systems (with no or a diagonal preconditioner) as one
            persistent kernel launch (ck_cg_fused); other systems use chained launches.
    """
    tol: float = 1e-6
    max_iters: int = 1000
    pbar: bool = False
    x0: object = None
    P: object = None
    chunk: int = 16
    fused: bool = True

    def __call__(self, A, b):
        return cg(A, b, x0=self.x0, P=self.P, tol=self.tol, max_iters=self.max_iters, chunk=self.chunk,
                  fused=self.fused)


def cg(A, rhs, x0=None, P=None, tol=1e-6, max_iters=5000, pbar=False, chunk=16, fused=True):
    """cg.py:39-69 on the engine. rhs (n,) or (n, b) on the HIP device; returns (soln, info).
    Differentiable with respect to the operator's tensors (cg_bwd, cg.py:72-86)."""
    is_vector = rhs.dim() == 1
    if is_vector:
        rhs = rhs[:, None]
        if x0 is not None:
            x0 = x0[:, None]
    X, _, _, info = run_cg(A, rhs, x0, max_iters, tol, P, chunk, fused)
    return (X.reshape(-1) if is_vector else X), info


def cg_bwd(res, grads, unflatten, *args, **kwargs):
    """cg.py:72-86: the output gradient is solved with the same CG settings (including the
    forward's x0, as the reference does), db = A^-1 dx, and the operator's gradient is the
    vector-Jacobian product of theta -> A(theta) @ x with -db."""
    x0, max_iters, tol, P, chunk, fused = args[1:7]
    y_grads = grads[0]
    op_args, output = res
    soln = output[0]
    A = unflatten(op_args)
    if y_grads is None:
        return GradOperator([None] * len(op_args)), None
    db, *_ = run_batched_cg(A, y_grads.contiguous(), x0, max_iters, tol, P, chunk=chunk, fused=fused)

    def fun(*theta):
        return unflatten(theta) @ soln

    d_params = vjp_derivs(fun, op_args, -db)
    return GradOperator(d_params), db


@iterative_autograd(cg_bwd)
def run_cg(A, b, x0, max_iters, tol, P, chunk=16, fused=True):
    """cg.py:89-91: run_batched_cg with the backward rule cg_bwd."""
    return run_batched_cg(A, b, x0, max_iters, tol, P, chunk=chunk, fused=fused)


STATUS_NAMES = {N.CK_OK: "ok", N.CK_ERR_MAXITER: "maxiter", N.CK_ERR_BREAKDOWN: "breakdown", N.CK_ERR_NAN: "nan",
                N.CK_ERR_TIMEOUT: "timeout"}


def _status_name(code):
    return STATUS_NAMES.get(int(code), f"status {int(code)}")


_FUSED_BUFFERS = {}


def _fused_buffers(dev, n, B, max_iters):
    """State, workspace and pinned host copy of the state for ck_cg_fused, kept per (device,
    stream, shape): solves enqueued on one stream run in order, so they can share them."""
    stream = torch.cuda.current_stream(dev).cuda_stream
    key = (dev.index, stream, n, B, int(max_iters))
    bufs = _FUSED_BUFFERS.get(key)
    if bufs is None:
        if len(_FUSED_BUFFERS) >= 32:
            _FUSED_BUFFERS.clear()
        nbytes = N.q("ck_cg_state_bytes", B, max_iters)
        wb = N.q("ck_cg_fused_workspace_bytes", n, B)
        bufs = (torch.empty(nbytes, dtype=torch.uint8, device=dev),
                torch.empty((wb + 7) // 8, dtype=torch.float64, device=dev),
                torch.empty(nbytes, dtype=torch.uint8, pin_memory=True), N.q("ck_cg_stride", B))
        _FUSED_BUFFERS[key] = bufs
    return bufs


def fused_solve(A, b, x0=None, max_iters=1000, tol=1e-6, P=None, nblocks=0, spin_limit=0, wait_offset=0,
                cooperative=False):
    """cg.py:94-119 as one persistent launch (ck_cg_fused) when A lowers to a dense fp32
    operator and P is None or diagonal; returns (x, r, k, info) or None when unsupported.
    b (n, B) is read in place (any strides), x and r are written in the same layout."""
    aop = lower(A)
    pop = lower(P) if P is not None else None
    if aop is None or (P is not None and pop is None):
        return None
    n, B = b.shape
    if N.q("ck_cg_fused_supported", aop[0], pop[0] if pop else None, B) != N.CK_OK:
        return None
    dev = b.device
    bf = b if b.dtype == torch.float32 else b.float()
    xr = torch.empty((2, n, B), dtype=torch.float32, device=dev)
    x, r = xr[0], xr[1]
    state, work, host, stride = _fused_buffers(dev, n, B, max_iters)
    x0f = None
    if x0 is not None:
        x0f = x0.reshape(n, B)
        x0f = x0f if x0f.dtype == torch.float32 else x0f.float()
    t0 = time.time()
    N.call("ck_cg_fused", aop[0], pop[0] if pop else None, B, N.ptr(bf), bf.stride(0), bf.stride(1), N.ptr(x0f),
           x0f.stride(0) if x0f is not None else 0, x0f.stride(1) if x0f is not None else 0, N.ptr(x), x.stride(0),
           x.stride(1), N.ptr(r), r.stride(0), r.stride(1), N.ptr(state), float(tol), int(max_iters), N.ptr(work),
           work.numel() * 8, int(nblocks), int(spin_limit), int(wait_offset), int(bool(cooperative)),
           N.stream_ptr(dev))
    host.copy_(state, non_blocking=True)
    torch.cuda.current_stream(dev).synchronize()
    ctrl = host[:N.CTRL_BYTES].view(torch.int32)
    k, status = int(ctrl[0]), int(ctrl[5])
    if status == N.CK_ERR_TIMEOUT:
        raise N.KrylovError("ck_cg_fused", status, "a grid-wide wait of the persistent CG timed out")
    f = host[N.CTRL_BYTES:].view(torch.float32).numpy()
    errs = f[N.CK_CG_ERRORS * stride:N.CK_CG_ERRORS * stride + k + 1].astype(np.float64)
    rn = f[N.CK_CG_RNORM * stride:N.CK_CG_RNORM * stride + B]
    tl = f[N.CK_CG_TOL * stride:N.CK_CG_TOL * stride + B]
    elapsed = time.time() - t0
    info = {"iterations": k + 1, "errors": np.concatenate([errs, errs[-1:]])[2:], "iteration_time": elapsed / (k + 1),
            "converged": bool((rn <= tl).all()), "status": _status_name(status), "fused": True}
    return x, r, k, info


class CGState:
    """Device buffers of one batched solve (column blocks, control block, partials)."""

    def __init__(self, n, B, max_iters, device, precond):
        self.n, self.B, self.device = n, B, device
        self.ld = ld_of(n)
        mk = lambda: torch.zeros((B, self.ld), dtype=torch.float32, device=device)  # noqa: E731
        self.bn, self.x, self.r, self.p, self.Ap = mk(), mk(), mk(), mk(), mk()
        self.z = mk() if precond else self.r
        self.state = torch.zeros(N.q("ck_cg_state_bytes", B, max_iters), dtype=torch.uint8, device=device)
        self.partials = torch.empty(N.q("ck_cg_partials_bytes", n, B) // 4, dtype=torch.float32, device=device)
        self.stride = N.q("ck_cg_stride", B)

    def floats(self, index, count):
        off = N.CTRL_BYTES + 4 * self.stride * index
        return self.state[off:off + 4 * count].view(torch.float32)


def run_batched_cg(A, b, x0=None, max_iters=1000, tol=1e-6, P=None, chunk=16, return_state=False, fused=True):
    """cg.py:94-119. b is (n, B) on the HIP device. Returns (x, r, k, info)."""
    N.require_device(b, "right-hand side")
    if b.dim() == 1:
        b = b[:, None]
    n, B = b.shape
    if B > 1024:
        raise ValueError("at most 1024 right-hand sides per solve")
    if fused and not return_state:
        out = fused_solve(A, b, x0, max_iters, tol, P)
        if out is not None:
            return out
    st = CGState(n, B, max_iters, b.device, P is not None)
    Bb = to_blocks(b, st.ld)
    x0b = to_blocks(x0.reshape(n, B), st.ld) if x0 is not None else None
    k, done, info = cg_blocks(A, Bb, st, x0b, max_iters, tol, P, chunk)
    x = from_blocks(st.x, n)
    r = from_blocks(st.r, n)
    if return_state:
        info["state"] = st
    return x, r, k, info


def cg_blocks(A, Bb, st, x0b=None, max_iters=1000, tol=1e-6, P=None, chunk=16, lowered=None, want_info=True,
              read=True, prepared=False, finish=True, final=False, x64=None):
    """Batched CG on column blocks: solves A X = Bb[:, :n] into st.x (st.r the residual).

    ``st`` (a CGState) is reused across calls; ``lowered`` may carry precomputed
    (aop, pop, work, wb). Returns (k, done, info).  With ``read=False`` exactly
    min(chunk, max_iters) steps are enqueued (each stops on the device flag) and
    nothing is read back: returns (None, None, {}).  ``prepared``: st.r already holds
    the normalised rhs, st.x zeros and the state its column norms (``Bb`` unused).
    ``finish=False`` leaves x and r normalised (the caller un-normalises).  ``final`` (with
    ``read=False``): the last enqueued step stops after the x/r update, for a caller that
    takes only x after exactly that many steps (ck_cg_iterate_final); with ``x64`` as well
    (an FP64 (B, ld) block) that step adds mult * x to x64 instead of updating x and r
    (ck_cg_iterate_final_f64), and the returned info holds ``x64_updated``."""
    n, B, dev, ld = st.n, st.B, st.device, st.ld
    if lowered is None:
        aop = lower(A)
        pop = lower(P) if P is not None else None
        work, wb = workspace([aop[0] if aop else None, pop[0] if pop else None], B, dev)
    else:
        aop, pop, work, wb = lowered
    s = N.stream_ptr(dev)
    t0 = time.time()
    if prepared:
        if x0b is not None:
            raise ValueError("a prepared start has x = 0")
    else:
        N.call("ck_cg_normalize", n, B, N.ptr(Bb), Bb.stride(0), N.ptr(st.bn), ld, N.ptr(st.state),
               N.ptr(st.partials), s)
    if prepared:
        pass
    elif x0b is None:
        st.x.zero_()
        st.r.copy_(st.bn)
    else:
        # cg.py:123: r0 = b_norm - A x0 (x0 taken as given, not normalised)
        st.x.copy_(x0b)
        apply_blocks(A, aop, st.x, st.Ap, n, B, work, wb)
        torch.sub(st.bn, st.Ap, out=st.r)
    if P is not None:
        apply_blocks(P, pop, st.r, st.z, n, B, work, wb)
    N.call("ck_cg_start", n, B, N.ptr(st.r), N.ptr(st.z), N.ptr(st.p), ld, N.ptr(st.state), float(tol),
           int(max_iters), N.ptr(st.partials), s)
    fused = aop is not None and (P is None or pop is not None)
    steps = 0
    # the loop flag is first read after one chunk: a column block that starts converged costs
    # one chunk of early-exiting launches instead of an extra host round trip on every call
    k, done = 0, False
    while not done and steps < max_iters:
        nst = min(int(chunk), max_iters - steps)
        last_partial = final and not read
        x64_updated = False
        if fused and last_partial and x64 is not None:
            N.call("ck_cg_iterate_final_f64", aop[0], pop[0] if pop else None, B, N.ptr(st.x), N.ptr(st.r),
                   N.ptr(st.p), N.ptr(st.Ap), N.ptr(st.z), ld, N.ptr(st.state), N.ptr(st.partials), N.ptr(work), wb,
                   nst, N.ptr(x64), x64.stride(0), s)
            x64_updated = True
        elif fused:
            N.call("ck_cg_iterate_final" if last_partial else "ck_cg_iterate", aop[0], pop[0] if pop else None, B,
                   N.ptr(st.x), N.ptr(st.r), N.ptr(st.p), N.ptr(st.Ap), N.ptr(st.z), ld, N.ptr(st.state),
                   N.ptr(st.partials), N.ptr(work), wb, nst, s)
        else:
            for i in range(nst):
                apply_blocks(A, aop, st.p, st.Ap, n, B, work, wb)
                N.call("ck_cg_step_xr", n, B, N.ptr(st.p), N.ptr(st.Ap), N.ptr(st.x), N.ptr(st.r), ld,
                       N.ptr(st.state), N.ptr(st.partials), s)
                if last_partial and i + 1 == nst:
                    break
                if P is not None:
                    apply_blocks(P, pop, st.r, st.z, n, B, work, wb)
                N.call("ck_cg_step_p", n, B, N.ptr(st.r), N.ptr(st.z), N.ptr(st.p), ld, N.ptr(st.state),
                       N.ptr(st.partials), s)
        steps += nst
        if not read:
            if finish and not x64_updated:
                N.call("ck_cg_finish", n, B, N.ptr(st.x), N.ptr(st.r), ld, N.ptr(st.state), s)
            return None, None, {"x64_updated": x64_updated}
        k, done, _, _, _, status = read_ints(st.state, 6, dev)
    if steps == 0:
        k, done, _, _, _, status = read_ints(st.state, 6, dev)
    if finish:
        N.call("ck_cg_finish", n, B, N.ptr(st.x), N.ptr(st.r), ld, N.ptr(st.state), s)
    if not want_info:
        return k, done, {"iterations": k + 1}
    errs = st.floats(N.CK_CG_ERRORS, k + 1).cpu().numpy().astype(np.float64)
    elapsed = time.time() - t0
    # torch_tqdm.py:38-62: one error per cond evaluation (k+1) plus the final one, first two dropped
    info = {"iterations": k + 1, "errors": np.concatenate([errs, errs[-1:]])[2:],
            "iteration_time": elapsed / (k + 1), "converged": bool(done and k < max_iters) or _all_below(st, B),
            "status": _status_name(status), "fused": False}
    return k, done, info


def _all_below(st, B):
    rn = st.floats(N.CK_CG_RNORM, B).cpu()
    tl = st.floats(N.CK_CG_TOL, B).cpu()
    return bool((rn <= tl).all())
