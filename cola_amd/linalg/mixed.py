"""Mixed-precision CG: FP64-residual iterative refinement around FP32 inner CG.

No reference counterpart -- cola's CG runs in the operator's dtype
(cola/linalg/inverse/cg.py:94-119).  ``HipMixedCG`` keeps the inner solver the
engine's CG (same kernels, optional preconditioner ``P``) and wraps it in

    r = b - A x   (FP64, the matrix as stored)      x += CG_fp32(A, r)

until every column satisfies ||b - A x|| / ||b|| <= tol in FP64.  The FP32
recurrence's attainable accuracy limit therefore no longer bounds the answer.

``KroneckerEigInverse`` is the structured preconditioner for K1 x K2 x K3 + s I
with 64x64 symmetric factors: (Q1 x Q2 x Q3) diag(1/(lambda + s)) (Q1 x Q2 x Q3)^T
applied with the same MFMA Kronecker passes as the operator (CK_OP_KRON_EIG_F32).
"""
import ctypes
import threading
import time
from dataclasses import dataclass

import numpy as np

import torch

from .. import _native as N
from .._blocks import from_blocks, ld_of, to_blocks
from ..annotations import PSD
from ..ops import Dense, Identity, Kronecker, LinearOperator, Product, ScalarMul, Sum, _op
from ._engine import lower, workspace
from .algorithm_base import Algorithm
from .cg import CGState, cg_blocks


def kron_structure(A):
    """(factors, scale, shift) when A is scale*(K1 x K2 x K3) + shift*I with dense factors, else None."""
    got = A.hip_op()
    if got is None or got[0].kind != N.CK_OP_KRON_F32:
        return None
    op = got[0]
    facs = [M for M in _find_kron(A).Ms]
    return [M.A for M in facs], float(op.scale), float(op.shift)


def _find_kron(A):
    if isinstance(A, Kronecker):
        return A
    if isinstance(A, (Sum, Product)):
        for M in A.Ms:
            k = _find_kron(M)
            if k is not None:
                return k
    return None


class KroneckerEigInverse(LinearOperator):
    """(scale*(K1 x K2 x K3) + shift*I)^{-1} through the factors' eigendecompositions.

    Built once (three 64x64 symmetric eigensolves in FP64); applied as two MFMA
    Kronecker passes pairs and a diagonal scaling."""

    def __init__(self, factors, scale=1.0, shift=0.0):
        dev = factors[0].device
        self.device = dev
        n = 1
        for f in factors:
            n *= f.shape[0]
        super().__init__(torch.float32, (n, n), annotations={PSD})
        lams, Qs = [], []
        for f in factors:
            w, Q = torch.linalg.eigh(f.double())
            lams.append(w)
            Qs.append(Q)
        lam = lams[0]
        for w in lams[1:]:
            lam = (lam[:, None] * w[None, :]).reshape(-1)
        self.lam = scale * lam + shift
        self.dinv = (1.0 / self.lam).to(torch.float32).contiguous()
        self.Q = [Q.to(torch.float32).contiguous() for Q in Qs]
        self.Qt = [Q.T.contiguous() for Q in self.Q]

    def hip_op(self):
        if len(self.Q) != 3 or any(Q.shape[0] != 64 for Q in self.Q):
            return None
        op = _op(N.CK_OP_KRON_EIG_F32, self.shape[0])
        op.nfactors = 3
        for i in range(3):
            op.fdim[i] = 64
            op.factor[i] = self.Q[i].data_ptr()
            op.factor2[i] = self.Qt[i].data_ptr()
        op.aux = self.dinv.data_ptr()
        return op, self.Q + self.Qt + [self.dinv]

    def _matmat(self, X):
        if self._use_hip(X):
            return self._hip_apply(X)
        raise NotImplementedError("KroneckerEigInverse runs on the engine (HIP device, 64x64x64)")


class _F64Apply:
    """FP64 application of A for refinement residuals (the matrix as stored):
    64^3 Kronecker operators through ck_kron_apply_f64, dense fp32 matrices (with
    scale and shift) through ck_dense_apply_f64acc, CSR matrices through
    ck_csr_apply_f64acc."""

    def __init__(self, A, nrhs):
        ks = kron_structure(A)
        self.kind = None
        got = A.hip_op()
        if ks is not None and all(f.shape[0] == 64 for f in ks[0]) and len(ks[0]) == 3:
            self.kind = "kron64"
            self.f64 = [f.double().contiguous() for f in ks[0]]
            self.scale, self.shift = ks[1], ks[2]
            self.dims = torch.tensor([64, 64, 64], dtype=torch.int32)
            wbytes = max(N.q("ck_kron_residual_workspace_bytes", nrhs), nrhs * 64**3 * 8)
            self.work = torch.empty(wbytes // 8, dtype=torch.float64, device=A.device)
        elif got is not None and got[0].kind == N.CK_OP_DENSE_F32:
            self.kind = "dense"
            self.op, self.keep = got
        elif got is not None and got[0].kind == N.CK_OP_CSR_F32:
            self.kind = "csr"
            self.op, self.keep = got
        else:
            raise N.UnsupportedOperator("mixed-precision refinement", N.CK_ERR_UNSUPPORTED,
                                        "FP64 residuals need a dense or CSR fp32 matrix or a 64^3 Kronecker "
                                        "operator (plus shift)")

    def residual(self, X, Bv, R32, norms, nrhs):
        """R32 = float(Bv - A X), norms[2c] = ||Bv_c - A X_c||, norms[2c+1] = ||Bv_c|| in one
        FP64 pass pair (64^3 Kronecker operators); returns False for other operators."""
        if self.kind != "kron64":
            return False
        N.call("ck_kron_residual_f64", N.ptr(self.f64[0]), N.ptr(self.f64[1]), N.ptr(self.f64[2]), self.scale,
               self.shift, nrhs, N.ptr(X), X.stride(0), N.ptr(Bv), Bv.stride(0), N.ptr(R32), R32.stride(0),
               N.ptr(norms), N.ptr(self.work), self.work.numel() * 8, N.stream_ptr(X.device))
        return True

    def __call__(self, X, Y, nrhs):
        ld = X.stride(0)
        if self.kind == "dense":
            op = self.op
            N.call("ck_dense_apply_f64acc", op.n, ctypes.c_void_p(op.a), op.lda, float(op.scale), float(op.shift),
                   nrhs, N.ptr(X), ld, N.ptr(Y), Y.stride(0), N.stream_ptr(X.device))
            return
        if self.kind == "csr":
            op = self.op
            N.call("ck_csr_apply_f64acc", op.n, ctypes.c_void_p(op.rowptr), ctypes.c_void_p(op.colind),
                   ctypes.c_void_p(op.values), float(op.scale), float(op.shift), nrhs, N.ptr(X), ld, N.ptr(Y),
                   Y.stride(0), N.stream_ptr(X.device))
            return
        dims = (ctypes_int32_array(self.dims))
        N.call("ck_kron_apply_f64", 3, dims, N.ptr(self.f64[0]), N.ptr(self.f64[1]), N.ptr(self.f64[2]),
               self.scale, self.shift, nrhs, N.ptr(X), ld, N.ptr(Y), Y.stride(0), N.ptr(self.work),
               self.work.numel() * 8, N.stream_ptr(X.device))


def ctypes_int32_array(t):
    vals = [int(v) for v in t.tolist()]
    return (ctypes.c_int32 * len(vals))(*vals)


@dataclass
class HipMixedCG(Algorithm):
    """Iterative refinement with FP64 residuals around the engine's FP32 (P)CG.

    The solution is returned as an FP64 tensor (the refined iterate).

    Args:
        tol: target FP64 relative residual ||b - A x|| / ||b|| for every column.
        inner_tol: CG tolerance of each correction solve (cg.py:101 semantics).
        inner_max_iters: step cap of each correction solve.
        max_refinements: outer step cap.
        P: preconditioner of the inner CG (approximation of A^-1).
        chunk: CG steps between reads of the device loop flag.
        inner: operator the correction solves use in place of A (e.g. a bf16 copy of a
            dense matrix, see ``bf16_inner``); the residuals always use A as stored.
        contraction: with a lower-precision ``inner``, a refinement that does not cut
            the FP64 residual by this factor switches the remaining corrections to A
            itself (the perturbation A - inner is then too large for refinement to
            converge on this system).
        graph: replay a repeated solve's schedule (refinements x inner steps, taken from
            the previous call on the same solver) as one captured HIP graph instead of
            launch by launch.  Every captured launch is a kernel (no memset or host
            read), and each inner step still stops on the device flag.
    """
    tol: float = 1e-6
    inner_tol: float = 1e-4
    inner_max_iters: int = 1000
    max_refinements: int = 20
    P: object = None
    chunk: int = 8
    inner: object = None
    contraction: float = 0.05
    graph: bool = False

    def __call__(self, A, b):
        is_vector = b.dim() == 1
        x, info = mixed_cg(A, b[:, None] if is_vector else b, self)
        return (x[:, 0] if is_vector else x), info


class MixedSolver:
    """Reusable buffers for repeated refinement solves with the same operator and block shape."""

    def __init__(self, A, nrhs, alg):
        self.A, self.alg = A, alg
        n = A.shape[0]
        self.n, self.B = n, nrhs
        dev = A.device
        self.ld = ld_of(n)
        self.f64 = _F64Apply(A, nrhs)
        self.st = CGState(n, nrhs, alg.inner_max_iters, dev, alg.P is not None)
        self.lowered_for = {}
        self.inner = alg.inner if alg.inner is not None else A
        for op in {id(A): A, id(self.inner): self.inner}.values():
            aop = lower(op)
            pop = lower(alg.P) if alg.P is not None else None
            work, wb = workspace([aop[0] if aop else None, pop[0] if pop else None], nrhs, dev)
            self.lowered_for[id(op)] = (aop, pop, work, wb)
        z64 = lambda: torch.zeros((nrhs, self.ld), dtype=torch.float64, device=dev)  # noqa: E731
        self.b64, self.x64, self.Ax64 = z64(), z64(), z64()
        # the inner CG's column norms: written by ck_ir_inner_start, read by the scaled update
        self.mult = self.st.floats(N.CK_CG_MULT, nrhs)
        self.r32 = torch.zeros((nrhs, self.ld), dtype=torch.float32, device=dev)
        self.norms = torch.zeros(2 * nrhs, dtype=torch.float64, device=dev)
        wbytes = N.q("ck_ir_workspace_bytes", nrhs)
        self.irwork = torch.empty(wbytes // 8, dtype=torch.float64, device=dev)

    def residual(self, with_ax=True, read=True):
        n, B, ld = self.n, self.B, self.ld
        s = N.stream_ptr(self.b64.device)
        # 64^3 Kronecker operators: one fused FP64 pass pair; otherwise A x, then the residual pass
        if not (with_ax and self.f64.residual(self.x64, self.b64, self.r32, self.norms, B)):
            if with_ax:
                self.f64(self.x64, self.Ax64, B)
            N.call("ck_ir_residual_f64", n, B, N.ptr(self.b64), ld, N.ptr(self.Ax64) if with_ax else None, ld,
                   None, ld, N.ptr(self.r32), ld, N.ptr(self.norms), N.ptr(self.irwork), self.irwork.numel() * 8, s)
        if not read:
            return None
        nr = self.norms.view(B, 2).cpu()
        return (nr[:, 0] / torch.clamp(nr[:, 1], min=1e-300)).numpy()

    def correction(self, inner, max_steps, chunk, read=True):
        """One inner correction solve on the current residual and its FP64 update of x:
        the inner CG starts from the residual pass's own norms (ck_ir_inner_start) and
        its un-normalisation is folded into the update (ck_ir_update_scaled_f64)."""
        n, B, ld, alg = self.n, self.B, self.ld, self.alg
        s = N.stream_ptr(self.b64.device)
        if not read and max_steps == 1 and self._step1_ok(inner):
            # one PCG step from x = 0 is scale invariant: no normalisation, no CG state (ck_pcg_step1_f64)
            aop, pop, work, wb = self.lowered_for[id(inner)]
            N.call("ck_pcg_step1_f64", aop[0], pop[0], B, N.ptr(self.r32), N.ptr(self.st.z), N.ptr(self.st.Ap), ld,
                   N.ptr(self.x64), ld, N.ptr(self.st.partials), N.ptr(work), wb, s)
            return 1
        N.call("ck_ir_inner_start", n, B, N.ptr(self.r32), ld, N.ptr(self.norms), N.ptr(self.st.r), N.ptr(self.st.x),
               ld, N.ptr(self.mult), s)
        # a replayed correction runs exactly max_steps steps and only its x is used: its last step
        # skips the preconditioner and the next search direction
        k, _, info = cg_blocks(inner, None, self.st, None, max_steps, alg.inner_tol, alg.P, chunk,
                               self.lowered_for[id(inner)], want_info=False, read=read, prepared=True, finish=False,
                               final=not read, x64=self.x64)
        if not info.get("x64_updated"):
            N.call("ck_ir_update_scaled_f64", n, B, N.ptr(self.x64), ld, N.ptr(self.st.x), ld, N.ptr(self.mult), s)
        return k

    def _step1_ok(self, inner):
        """Whether ck_pcg_step1_f64 covers (inner, P): a lowered 64^3 Kronecker operator with the
        Kronecker eigenbasis preconditioner (checked once per operator)."""
        cache = self.__dict__.setdefault("_step1", {})
        if id(inner) not in cache:
            aop, pop, work, wb = self.lowered_for[id(inner)]
            cache[id(inner)] = (aop is not None and pop is not None and aop[0].kind == N.CK_OP_KRON_F32
                                and pop[0].kind == N.CK_OP_KRON_EIG_F32 and self.n == 64**3)
        return cache[id(inner)]

    def _schedule(self, plan):
        """One correction solve of k steps and one FP64 residual per plan entry (from x = 0, set
        by the rhs load): launches only (no host reads), so it can be captured as a graph.
        One-step entries on a 64^3 Kronecker system with its eigenbasis preconditioner run as one
        ck_ir_kron_step1_f64 call (the x update inside the residual's first pass); only the last
        residual's norms are formed (the host reads them once)."""
        for i, k_steps in enumerate(plan):
            last = i + 1 == len(plan)
            if k_steps == 1 and self.f64.kind == "kron64" and self._step1_ok(self.A):
                self._kron_step1(accumulate=i > 0, norms=last)
            else:
                self.correction(self.A, k_steps, k_steps, read=False)
                self.residual(with_ax=True, read=False)

    def _kron_step1(self, accumulate, norms):
        n, B, ld = self.n, self.B, self.ld
        aop, pop, work, wb = self.lowered_for[id(self.A)]
        f = self.f64
        N.call("ck_ir_kron_step1_f64", aop[0], pop[0], N.ptr(f.f64[0]), N.ptr(f.f64[1]), N.ptr(f.f64[2]), f.scale,
               f.shift, B, N.ptr(self.r32), N.ptr(self.st.z), N.ptr(self.st.Ap), ld, N.ptr(self.x64), N.ptr(self.b64),
               1 if accumulate else 0, N.ptr(self.norms) if norms else None, N.ptr(self.st.partials), N.ptr(work), wb,
               N.ptr(f.work), f.work.numel() * 8, N.stream_ptr(self.b64.device))

    def solve(self, b):
        alg = self.alg
        n, B, ld = self.n, self.B, self.ld
        t0 = time.time()
        # b64 = b, r = b, x = 0 and the norms of b in one pass over the caller's b (any strides; the
        # residual and update passes touch columns [:n] only, so the blocks' padding is never read).
        # Relative residual 1 for every column (0 for a zero column, which the first correction
        # solve leaves at x = 0): no host read is needed before the first solve.
        bf = b if b.dtype == torch.float32 else b.float()
        plan = getattr(self, "_plan", None)
        replay = bool(plan) and self.inner is self.A and len(plan) <= alg.max_refinements
        # a replay that starts with a one-step Kronecker refinement sets x itself and never reads
        # the initial norms: the load then writes b64 and r32 only
        lean = replay and plan[0] == 1 and self.f64.kind == "kron64" and self._step1_ok(self.A)
        N.call("ck_ir_load_rhs", n, B, N.ptr(bf), bf.stride(0), bf.stride(1), N.ptr(self.b64), N.ptr(self.r32),
               None if lean else N.ptr(self.x64), ld, None if lean else N.ptr(self.norms), N.ptr(self.irwork),
               self.irwork.numel() * 8, N.stream_ptr(b.device))
        rel = np.ones(B)
        history = [1.0]
        inner_steps = []
        it = 0
        # the first correction solve usually takes as many steps as in the previous call on this
        # solver: enqueue that many before the first read (steps past convergence still cost launches)
        chunk = max(1, min(alg.chunk, getattr(self, "_first_steps", alg.chunk)))
        inner = self.inner
        fallback_at = None
        if replay:
            # replay the previous call's schedule (refinements x inner steps) with no host read in
            # between, then read the FP64 residual once; if that is not yet enough, the loop below
            # continues from the refined iterate
            if alg.graph:
                if getattr(self, "_graph_plan", None) != plan:
                    self._graph = None
                    torch.cuda.synchronize(b.device)
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g):
                        self._schedule(plan)
                    self._graph, self._graph_plan = g, list(plan)
                self._graph.replay()
            else:
                self._schedule(plan)
            nr = self.norms.view(B, 2).cpu()
            rel = (nr[:, 0] / torch.clamp(nr[:, 1], min=1e-300)).numpy()
            it = len(plan)
            inner_steps = list(plan)
            history.append(float(rel.max()))
        while not (np.all(np.isfinite(rel)) and rel.max() <= alg.tol) and it < alg.max_refinements:
            if not np.all(np.isfinite(rel)):
                raise FloatingPointError("non-finite FP64 residual in iterative refinement")
            keep = self.x64.clone() if inner is not self.A else None
            k = self.correction(inner, alg.inner_max_iters, chunk)
            inner_steps.append(int(k))
            if it == 0:
                self._first_steps = max(1, int(k))
            # the next correction solve usually needs as many steps: enqueue that many before the first check
            chunk = max(1, min(alg.chunk, int(k)))
            prev = rel.max()
            rel = self.residual(with_ax=True)
            it += 1
            if inner is not self.A and not np.all(np.isfinite(rel)):
                # the lower-precision correction solve broke down: drop it, go on with A itself
                self.x64.copy_(keep)
                rel = self.residual(with_ax=True)
                inner = self.A
                fallback_at = it
            elif inner is not self.A and not (rel.max() <= alg.contraction * prev):
                inner = self.A
                fallback_at = it
            history.append(float(rel.max()))
        if fallback_at is None and inner is self.A:
            self._plan = [max(1, int(k)) for k in inner_steps]
        # the refined iterate is returned in FP64: rounding it to FP32 would cost ~kappa * 6e-8 of residual
        x = self.x64[:, :n].T
        info = {"refinements": it, "inner_iterations": inner_steps, "rel_residual_f64": rel,
                "precision_fallback_at": fallback_at,
                "history": history, "converged": bool(np.all(np.isfinite(rel)) and rel.max() <= alg.tol), "time": time.time() - t0}
        return x, info


def mixed_cg(A, b, alg):
    N.require_device(b, "right-hand side")
    solver = MixedSolver(A, b.shape[1], alg)
    return solver.solve(b)


_STREAMS = {}


def solve_concurrent(solvers, rhs, streams=2):
    """``[s.solve(b) for s, b in zip(solvers, rhs)]`` for independent systems, on ``streams``
    HIP streams at once: one host thread per stream takes the systems k, k + streams, ... in
    order.  A refinement step of one system is a large matrix product followed by a chain of
    small latency-bound launches (vector updates, the preconditioner's core, step scalars); with
    two systems in flight the small launches of one run beside the other's product, whose
    bandwidth use they barely touch.  Every solver must already have been called once (the
    first call records its refinement schedule as a graph, which is not done concurrently).
    The streams keep the default priority and are reused across calls; the caller's stream
    waits for all of them before this returns.  Returns the list of (x, info)."""
    if len(solvers) != len(rhs):
        raise ValueError("one right-hand side block per solver")
    if streams <= 1 or len(solvers) <= 1:
        return [sv.solve(b) for sv, b in zip(solvers, rhs)]
    dev = rhs[0].device
    key = (str(dev), streams)
    if key not in _STREAMS:
        _STREAMS[key] = [torch.cuda.Stream(device=dev) for _ in range(streams)]
    ss = _STREAMS[key]
    cur = torch.cuda.current_stream(dev)
    for st in ss:
        st.wait_stream(cur)
    out = [None] * len(solvers)
    errors = []

    def run(k):
        try:
            with torch.cuda.stream(ss[k]):
                for i in range(k, len(solvers), streams):
                    out[i] = solvers[i].solve(rhs[i])
        except BaseException as e:  # re-raised in the caller's thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(streams)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for st in ss:
        cur.wait_stream(st)
    if errors:
        raise errors[0]
    return out


_ = (from_blocks, to_blocks, Dense, Identity, ScalarMul)
