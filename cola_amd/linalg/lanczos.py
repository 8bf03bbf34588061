"""HipLanczos: batched Lanczos with full double Gram-Schmidt on the engine.

Drop-in for cola's ``Lanczos`` (cola/linalg/decompositions/decompositions.py:116-144)
and ``lanczos``/``lanczos_eigs`` (cola/linalg/decompositions/lanczos.py:34-61,185-232):
same fields and the same (Q, T, info) results, the torch branch's trimming to
the executed iterations (lanczos.py:222-224), and the same stopping rule.
The tridiagonal eigensolve runs in FP64 on the device (the reference uses A's
dtype, lanczos.py:57); Ritz values are returned in A's dtype with the FP64
values in ``info['eigvals_f64']``.

Default start vector: a Gaussian probe of the engine's counter stream seeded
with ``PRNGKey(42)`` (the reference draws torch.randn with that seed on the
target device, lanczos.py:211-213, so its CPU and GPU vectors already differ).
"""
import time
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np
import torch

from .. import _native as N
from .._blocks import ld_of, to_blocks
from ..ops import Dense, Tridiagonal
from ._engine import lower, probes, read_ints, workspace
from .algorithm_base import Algorithm
from .keys import PRNGKey


@dataclass
class HipLanczos(Algorithm):
    """Lanczos decomposition A Q = Q T (+ residual) on the HIP engine.

    Args:
        start_vector: (n,) or (n, b) start vector(s).
        max_iters: steps (capped at n).
        tol: stop once every probe's beta_{i-1} <= tol * beta_1 (lanczos.py:263-268).
        pbar: accepted for compatibility.
        key: seed of the default start vector.
        chunk: steps enqueued between reads of the device loop flag.
        reorth: "full" (the reference's re-orthogonalisation of every new vector against the
            whole basis) or "block" (engine option: against its own block of ``block``
            columns each step, and each completed block against all older columns at once on
            FP32 MFMA; ``ck_lanczos_iterate_block``).
        block: block size of the "block" mode (2..16).
    """
    start_vector: Any = None
    max_iters: int = 1_000
    tol: float = 1e-6
    pbar: bool = False
    key: Optional[int] = None
    chunk: int = 32
    reorth: str = "full"
    block: int = 16

    def __call__(self, A):
        return lanczos(A, self.start_vector, self.max_iters, self.tol, self.pbar, self.key, chunk=self.chunk,
                       reorth=self.reorth, block=self.block)


class LanczosRun:
    """Device buffers and results of one batched Lanczos run."""

    def __init__(self, n, nb, m, device):
        self.n, self.nb, self.m, self.device = n, nb, m, device
        self.ld = ld_of(n)
        self.ldb = (m + 2) * self.ld
        self.V = torch.zeros((nb, m + 2, self.ld), dtype=torch.float32, device=device)
        self.diag = torch.zeros((nb, m), dtype=torch.float32, device=device)
        self.subdiag = torch.zeros((nb, m + 1), dtype=torch.float32, device=device)
        self.state = torch.zeros(N.q("ck_lanczos_state_bytes", nb, m), dtype=torch.uint8, device=device)
        sb = N.q("ck_lanczos_scratch_bytes", n, nb, m)
        self.scratch = torch.empty((sb + 3) // 4, dtype=torch.float32, device=device)
        self.iters = 0
        self.info = {}

    def basis(self, b=0):
        """(n, iters) view of probe b's Lanczos vectors."""
        return self.V[b, 1:self.iters + 1, :self.n].T

    def alpha(self):
        return self.subdiag[:, 1:self.iters]

    def beta(self):
        return self.diag[:, :self.iters]


def default_start(A, key, nb=1):
    seed = PRNGKey(42) if key is None else int(key)
    Z, _ = probes("normal", A.shape[0], nb, seed, A.device)
    return Z[:, :A.shape[0]].T


def run_lanczos(A, start, max_iters, tol, chunk=32, reorth="full", block=16):
    """Batched Lanczos on the device; start is (n, b). Returns a LanczosRun."""
    if reorth not in ("full", "block"):
        raise ValueError(f"reorth must be 'full' or 'block', got {reorth!r}")
    N.require_device(start, "start vector")
    n = A.shape[0]
    m = min(int(max_iters), n)
    if start.dim() == 1:
        start = start[:, None]
    nb = start.shape[1]
    dev = start.device
    run = LanczosRun(n, nb, m, dev)
    aop = lower(A)
    work, wb = workspace([aop[0] if aop else None], nb, dev)
    s = N.stream_ptr(dev)
    bscr, bsb = None, 0
    if reorth == "block":
        if aop is None:
            raise N.UnsupportedOperator("block re-orthogonalisation", N.CK_ERR_UNSUPPORTED,
                                        "the block mode runs with library operators only (hip_op)")
        bsb = N.q("ck_lanczos_block_scratch_bytes", n, nb, int(block))
        if bsb == 0:
            raise ValueError("block must lie in 2..16")
        bscr = torch.empty((bsb + 3) // 4, dtype=torch.float32, device=dev)
    t0 = time.time()
    S = to_blocks(start, run.ld)
    N.call("ck_lanczos_start", n, nb, N.ptr(S), run.ld, N.ptr(run.V), run.ld, run.ldb, N.ptr(run.diag),
           N.ptr(run.subdiag), N.ptr(run.state), m, N.ptr(run.scratch), s)
    step = 0
    i, done = read_ints(run.state, 2, dev)
    while not done and step < m:
        nst = min(int(chunk), m - step)
        if bscr is not None:
            N.call("ck_lanczos_iterate_block", aop[0], n, nb, m, N.ptr(run.V), run.ld, run.ldb, N.ptr(run.diag),
                   N.ptr(run.subdiag), N.ptr(run.state), float(tol), step, nst, int(block), N.ptr(run.scratch),
                   N.ptr(bscr), bsb, N.ptr(work), wb, s)
        elif aop is not None:
            N.call("ck_lanczos_iterate", aop[0], n, nb, m, N.ptr(run.V), run.ld, run.ldb, N.ptr(run.diag),
                   N.ptr(run.subdiag), N.ptr(run.state), float(tol), step, nst, N.ptr(run.scratch), N.ptr(work), wb, s)
        else:
            for st in range(step, step + nst):
                N.call("ck_lanczos_pre", n, nb, m, N.ptr(run.V), run.ld, run.ldb, N.ptr(run.subdiag),
                       N.ptr(run.state), st, N.ptr(run.scratch), s)
                vi = torch.as_strided(run.V, (n, nb), (1, run.ldb), (st + 1) * run.ld)
                W = A._matmat(vi)
                torch.as_strided(run.V, (nb, n), (run.ldb, 1), (st + 2) * run.ld).copy_(W.T)
                N.call("ck_lanczos_post", n, nb, m, N.ptr(run.V), run.ld, run.ldb, N.ptr(run.diag),
                       N.ptr(run.subdiag), N.ptr(run.state), float(tol), st, N.ptr(run.scratch), s)
        step += nst
        i, done = read_ints(run.state, 2, dev)
    run.iters = i - 1
    # cola/utils/torch_tqdm.py: errors at each cond evaluation plus the final one, first two dropped
    errs = run.state[N.CTRL_BYTES:N.CTRL_BYTES + 4 * i].view(torch.float32).cpu().numpy().astype(np.float64)
    run.info = {"iterations": i, "errors": np.concatenate([errs, errs[-1:]])[2:],
                "iteration_time": (time.time() - t0) / max(i, 1)}
    return run


def lanczos(A, start_vector=None, max_iters=100, tol=1e-7, pbar=False, key=None, chunk=32, reorth="full", block=16):
    """lanczos.py:185-232: returns (Q, T, info) with Q (n, k) [or (b, n, k)] and
    T = Tridiagonal(alpha, beta, alpha)."""
    if start_vector is None:
        start_vector = default_start(A, key)[:, 0]
    run = run_lanczos(A, start_vector, max_iters, tol, chunk, reorth, block)
    alpha, beta = run.alpha(), run.beta()
    if start_vector.dim() == 1:
        T = Tridiagonal(alpha[0], beta[0], alpha[0])
        return Dense(run.basis(0)), T, run.info
    Q = torch.stack([run.basis(b) for b in range(run.nb)])
    return Q, Tridiagonal(alpha, beta, alpha), run.info


# clusters (eigenvalues closer than 1e-3 ||T||) larger than this send the automatic choice to QL
CLUSTER_LIMIT = 32


def largest_cluster(ev):
    """Size of the largest run of ascending eigenvalues (rows of ev) whose neighbours are closer
    than 1e-3 max(|first|, |last|) -- the clusters ck_tridiag_eigh_bisect_f64 orthogonalises
    in one thread."""
    ev = np.atleast_2d(np.asarray(ev, dtype=np.float64))
    best = 1 if ev.shape[1] else 0
    for row in ev:
        if row.size < 2 or not np.all(np.isfinite(row)):
            continue
        tol = 1e-3 * max(abs(row[0]), abs(row[-1]))
        run = 1
        for close in np.diff(row) <= tol:
            run = run + 1 if close else 1
            best = max(best, run)
    return best


def tridiag_eigh_batch(d, e, vectors=True, method=None):
    """FP64 eigen-decompositions of nb tridiagonals: d (nb, k) and e (nb, >= k-1) fp32 rows.

    ``method`` "bisect" (Sturm multisection + inverse iteration, ck_tridiag_eigh_bisect_f64,
    the default for up to 64 problems of order up to 1024: its eigenvalues and vectors are
    computed in parallel, 0.1 ms for 32 problems of order 30 and 0.5 ms for one of order 120)
    or "ql" (implicit QL, ck_tridiag_eigh_f64, whose rotations are serial: 0.45 and 4.9 ms;
    the default for larger batches).  Inverse iteration orthogonalises the vectors of a
    cluster (eigenvalues closer than 1e-3 ||T||) serially in one thread, O(c^2 m) for a
    cluster of c: spectra with most values in one cluster (kernel matrices) are faster on QL,
    so the automatic choice computes the eigenvalues first when vectors are asked for and
    takes QL if a cluster holds more than CLUSTER_LIMIT of them.
    Returns (evals (nb, k) f64 ascending, evecs (nb, k, k) f64 or None, first_row (nb, k) f64,
    status (nb,) int32)."""
    nb, k = d.shape
    dev = d.device
    d = d if d.stride(-1) == 1 else d.contiguous()
    e = e if e.stride(-1) == 1 else e.contiguous()
    # a single row's stride(0) is not a leading dimension (torch may report 0 or 1 for it)
    ldd = d.stride(0) if nb > 1 else k
    lde = e.stride(0) if nb > 1 else max(e.shape[1], 1)
    ev = torch.empty((nb, k), dtype=torch.float64, device=dev)
    vecs = torch.empty((nb, k, k), dtype=torch.float64, device=dev) if vectors else None
    fr = torch.empty((nb, k), dtype=torch.float64, device=dev)
    st = torch.empty(nb, dtype=torch.int32, device=dev)
    if method is None:
        method = "bisect" if k <= 1024 and nb <= 64 else "ql"  # the bisection kernel holds T in LDS
        if method == "bisect" and vectors and k > CLUSTER_LIMIT:
            # eigenvalues first (one multisection pass): a spectrum with a large cluster goes to QL
            N.call("ck_tridiag_eigh_bisect_f64", nb, k, N.ptr(d), ldd, N.ptr(e), lde, N.ptr(ev), None, None,
                   N.ptr(st), None, 0, N.stream_ptr(dev))
            if largest_cluster(ev.cpu().numpy()) > CLUSTER_LIMIT:
                method = "ql"
    if method == "bisect":
        tmp = vecs if vectors else torch.empty((nb, k, k), dtype=torch.float64, device=dev)
        # the inverse iteration's factors: LDS for a few problems, this workspace for many
        wb = N.q("ck_tridiag_eigh_bisect_workspace_bytes", nb, k) if nb > 1 and nb * k > 512 else 0
        work = torch.empty(wb, dtype=torch.uint8, device=dev) if wb else None
        N.call("ck_tridiag_eigh_bisect_f64", nb, k, N.ptr(d), ldd, N.ptr(e), lde, N.ptr(ev), N.ptr(tmp), N.ptr(fr),
               N.ptr(st), N.ptr(work), wb, N.stream_ptr(dev))
    elif method == "ql":
        N.call("ck_tridiag_eigh_f64", nb, k, N.ptr(d), ldd, N.ptr(e), lde, N.ptr(ev), N.ptr(vecs), N.ptr(fr),
               N.ptr(st), N.stream_ptr(dev))
    else:
        raise ValueError(f"unknown tridiagonal eigensolver {method!r}")
    return ev, vecs, fr, st


def tridiag_eigh(run, vectors=True, method=None):
    """FP64 eigen-decomposition of each probe's T on the device (tridiag_eigh_batch)."""
    k = run.iters
    d = run.diag[:, :k].contiguous()
    e = run.subdiag[:, 1:max(k, 2)].contiguous()
    return tridiag_eigh_batch(d, e, vectors, method)


def ritz_vectors(run, S, b=0):
    """(n, c) Ritz vectors Q_b S for S (k, c) FP64 on the device (lanczos.py:59)."""
    k, c = S.shape
    out = torch.empty((c, run.ld), dtype=torch.float32, device=run.device)
    S = S.contiguous()
    N.call("ck_basis_combine", run.n, k, c, N.ptr(run.V[b, 1]), run.ld, N.ptr(S), S.stride(0), N.ptr(out), run.ld,
           N.stream_ptr(run.device))
    return out[:, :run.n].T


def lanczos_eigs(A, start_vector=None, max_iters=100, tol=1e-7, pbar=False, key=None, eig_slice=None, chunk=32,
                 reorth="full", block=16):
    """lanczos.py:34-61: ascending Ritz values and (lazy in the reference) Ritz vectors.

    ``eig_slice`` limits which Ritz vectors are formed."""
    if start_vector is None:
        start_vector = default_start(A, key)[:, 0]
    run = run_lanczos(A, start_vector, max_iters, tol, chunk, reorth, block)
    ev, vecs, _, st = tridiag_eigh(run, vectors=True)
    if int(st.max()) != 0:
        raise RuntimeError("tridiagonal eigensolve did not converge")
    sl = eig_slice if eig_slice is not None else slice(None)
    S = vecs[0][:, sl]
    V = ritz_vectors(run, S)
    info = dict(run.info)
    info["eigvals_f64"] = ev[0, sl]
    return ev[0, sl].to(A.dtype), Dense(V), info
