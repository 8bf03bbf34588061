"""Learned low-rank preconditioner for kernel systems (K + noise I) x = b.

No reference counterpart: the result is an ordinary preconditioner ``P`` for cola's
``CG(P=...)`` (cola/linalg/inverse/cg.py:39-69; applied as z = P r at cg.py:150).

* ``PointMLP`` -- a per-point network din -> 128 -> 128 -> k (GELU) with forward,
  backward and Adam on the engine (``ck_mlp_forward``, ``ck_mlp_backward``,
  ``ck_adam_step``).  Its output G (n x k) is a *sketch*: the preconditioner uses
  its Nystrom completion F = K G (G^T K G)^-1/2, so P^-1 = noise I + F F^T with
  F F^T = K G (G^T K G)^-1 G^T K <= K, applied by the Woodbury identity.
* Training target: the top eigenvectors U (weights w = eigenvalues) of seeded
  instances, computed by the engine's own Lanczos.  The loss is the share of the
  weighted energy the sketch misses, L = 1 - tr((G^T G)^-1 G^T U W U^T G) / sum w,
  which is invariant to G -> G R, so the network only has to learn a subspace.
* ``PivotedCholeskyPreconditioner`` -- the rank-r pivoted Cholesky comparator
  (Harbrecht, Peters & Schneider 2012; Gardner et al. 2018), P^-1 = noise I + L L^T.
* ``select_preconditioner`` -- builds both, keeps the one whose Nystrom error
  trace tr(K - F F^T) is smaller, and returns no preconditioner (plain CG) when even
  that error exceeds the largest captured eigenvalue, i.e. when the bound
  kappa(P A) <= 1 + tr(E) / noise promises no gain over kappa(A).
"""
import ctypes
import math
import time

import numpy as np
import torch

from .. import _native as N
from .._blocks import ld_of
from ..ops import Dense, LinearOperator, WoodburyInverse
from ._engine import apply_blocks, lower, workspace
from .lanczos import default_start, run_lanczos, tridiag_eigh, tridiag_eigh_batch

HIDDEN = 128


def _init_params(din, dout, seed):
    """torch.nn.Linear's default ranges U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a seeded NumPy stream."""
    rs = np.random.RandomState(seed)
    parts = [(din * HIDDEN, din), (HIDDEN, din), (HIDDEN * HIDDEN, HIDDEN), (HIDDEN, HIDDEN),
             (HIDDEN * dout, HIDDEN), (dout, HIDDEN)]
    out = []
    for cnt, fan in parts:
        b = 1.0 / math.sqrt(fan)
        out.append(rs.uniform(-b, b, size=cnt).astype(np.float32))
    return np.concatenate(out)


class PointMLP:
    """Per-point MLP din -> 128 -> 128 -> dout with exact GELU, parameters in one flat fp32 tensor
    (W1, b1, W2, b2, W3, b3; y = x W + b) and Adam state, all on the device."""

    def __init__(self, din=4, dout=64, device="cuda", seed=0, params=None):
        self.din, self.dout = int(din), int(dout)
        self.count = int(N.q("ck_mlp_param_count", self.din, self.dout))
        if self.count < 0:
            raise ValueError("din must be in [1, 8] and dout in [1, 128]")
        init = _init_params(self.din, self.dout, seed) if params is None else np.asarray(params, np.float32)
        if init.size != self.count:
            raise ValueError(f"expected {self.count} parameters, got {init.size}")
        self.params = torch.from_numpy(init.copy()).to(device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.steps = 0
        self.device = self.params.device

    def forward(self, X, save=False):
        """G (dout, ld) feature-major block for points X (n, din); with save, also Z1, Z2 (128, ld)."""
        N.require_device(X, "points")
        X = X.to(torch.float32).contiguous()
        n = X.shape[0]
        ld = ld_of(n)
        G = torch.zeros((self.dout, ld), dtype=torch.float32, device=X.device)
        Z1 = Z2 = None
        if save:
            Z1 = torch.empty((HIDDEN, ld), dtype=torch.float32, device=X.device)
            Z2 = torch.empty_like(Z1)
        N.call("ck_mlp_forward", n, self.din, self.dout, N.ptr(X), X.stride(0), N.ptr(self.params), N.ptr(G), 1, ld,
               N.ptr(Z1), N.ptr(Z2), ld, N.stream_ptr(X.device))
        return (G, Z1, Z2) if save else G

    def backward(self, X, Z1, Z2, dG, scale=1.0):
        """Flat gradient of sum(G * dG) (times scale) with respect to the parameters."""
        X = X.to(torch.float32).contiguous()
        n = X.shape[0]
        wb = N.q("ck_mlp_grad_workspace_bytes", n, self.din, self.dout)
        work = torch.empty((wb + 3) // 4, dtype=torch.float32, device=X.device)
        grad = torch.empty_like(self.params)
        N.call("ck_mlp_backward", n, self.din, self.dout, N.ptr(X), X.stride(0), N.ptr(self.params), N.ptr(Z1),
               N.ptr(Z2), Z1.stride(0), N.ptr(dG), 1, dG.stride(0), float(scale), N.ptr(grad), N.ptr(work), wb,
               N.stream_ptr(X.device))
        return grad

    def adam(self, grad, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.steps += 1
        N.call("ck_adam_step", self.count, N.ptr(self.params), N.ptr(grad), N.ptr(self.m), N.ptr(self.v), float(lr),
               float(beta1), float(beta2), float(eps), self.steps, N.stream_ptr(self.device))

    def state_dict(self):
        return {"din": self.din, "dout": self.dout, "params": self.params.cpu().numpy()}

    @classmethod
    def from_state_dict(cls, sd, device="cuda"):
        return cls(sd["din"], sd["dout"], device=device, params=sd["params"])


# ----------------------------------------------------------------------------- small algebra
def gram(X, xsr, xsc, Y, ysr, ysc, n, ka, kb):
    """C = X^T Y (ka x kb, FP64) for strided tall-skinny fp32 operands."""
    dev = X.device
    wb = N.q("ck_gram_workspace_bytes", n, ka, kb)
    work = torch.empty(wb // 8, dtype=torch.float64, device=dev)
    C = torch.empty((ka, kb), dtype=torch.float64, device=dev)
    N.call("ck_gram_f64", n, ka, kb, N.ptr(X), xsr, xsc, N.ptr(Y), ysr, ysc, N.ptr(C), N.ptr(work), wb,
           N.stream_ptr(dev))
    return C


def gram_blocks(Xb, Yb, n):
    """X^T Y for feature-major blocks Xb (ka, ld), Yb (kb, ld)."""
    return gram(Xb, 1, Xb.stride(0), Yb, 1, Yb.stride(0), n, Xb.shape[0], Yb.shape[0])


def rows_gemm(X, xsr, xsc, W, Y, ysr, ysc, n, alpha=1.0, beta=0.0):
    """Y = alpha X W + beta Y with X (n x kin) strided, W (kin x kout) fp32 row-major."""
    W = W.to(torch.float32).contiguous()
    N.call("ck_rows_gemm_f32", n, W.shape[0], W.shape[1], N.ptr(X), xsr, xsc, N.ptr(W), float(alpha), float(beta),
           N.ptr(Y), ysr, ysc, N.stream_ptr(X.device))


def spd_inverse(C, jitter_rel=0.0, jitter_abs=0.0, inv64=False, inv32=False, linvT32=False, check=True):
    """Cholesky-based inverse pieces of C + (jitter_abs + jitter_rel tr(C)/k) I (k <= 64)."""
    k = C.shape[0]
    dev = C.device
    C = C.contiguous()
    out = {"linv": torch.empty((k, k), dtype=torch.float64, device=dev),
           "status": torch.zeros(1, dtype=torch.int32, device=dev)}
    if inv64:
        out["inv64"] = torch.empty((k, k), dtype=torch.float64, device=dev)
    if inv32:
        out["inv32"] = torch.empty((k, k), dtype=torch.float32, device=dev)
    if linvT32:
        out["linvT32"] = torch.empty((k, k), dtype=torch.float32, device=dev)
    N.call("ck_spd_inverse_f64", 1, k, N.ptr(C), k, k * k, float(jitter_rel), float(jitter_abs), N.ptr(out["linv"]),
           N.ptr(out.get("inv64")), N.ptr(out.get("inv32")), N.ptr(out.get("linvT32")), N.ptr(out["status"]),
           N.stream_ptr(dev))
    if check:
        st = int(out["status"].item())
        if st != 0:
            raise np.linalg.LinAlgError(f"small SPD factorisation failed at pivot {st - 1}")
    return out


def energy_loss_grad(G, U, w, n, ridge_rel=1e-6):
    """Captured-energy loss of features G (k, ld) against targets U (m, ld) with weights w (m,)
    FP64; returns (loss as a 1-element FP64 device tensor, dL/dG (k, ld))."""
    k, m = G.shape[0], U.shape[0]
    dev = G.device
    S = gram_blocks(G, G, n)
    T = gram_blocks(U, G, n)
    inv = spd_inverse(S, jitter_rel=ridge_rel, inv64=True, check=False)
    C1 = torch.empty((m, k), dtype=torch.float32, device=dev)
    C2 = torch.empty((k, k), dtype=torch.float32, device=dev)
    loss = torch.empty(1, dtype=torch.float64, device=dev)
    N.call("ck_energy_loss_f64", k, m, N.ptr(inv["inv64"]), N.ptr(T), N.ptr(w), N.ptr(C1), N.ptr(C2), N.ptr(loss),
           N.stream_ptr(dev))
    dG = torch.empty_like(G)
    ld = G.stride(0)
    rows_gemm(U, 1, U.stride(0), C1, dG, 1, ld, n, 1.0, 0.0)
    rows_gemm(G, 1, ld, C2, dG, 1, ld, n, 1.0, 1.0)
    if ld > n:
        dG[:, n:].zero_()
    return loss, dG, inv["status"]


# ----------------------------------------------------------------------------- instances
def rbf_matrix(points, lengthscale, bf16=False):
    """Dense RBF kernel matrix of the points (n, d) on the device: fp32 (n, n) [and a bf16 copy]."""
    N.require_device(points, "points")
    P = points.to(torch.float32).contiguous()
    n, d = P.shape
    K = torch.empty((n, n), dtype=torch.float32, device=P.device)
    Kb = torch.empty((n, n), dtype=torch.bfloat16, device=P.device) if bf16 else None
    N.call("ck_rbf_dense_f32", n, d, N.ptr(P), P.stride(0), float(lengthscale), N.ptr(K), n, N.ptr(Kb), n,
           N.stream_ptr(P.device))
    return (K, Kb) if bf16 else K


def sample_instance(seed, n, din=4, ell_range=(0.05, 1.0), noise_range=(1e-4, 1e-1)):
    """Seeded points U[0,1]^din, log-uniform lengthscale and noise (CPU generator seeded with `seed`)."""
    g = torch.Generator().manual_seed(int(seed))
    X = torch.rand(n, din, generator=g)
    u = torch.rand(2, generator=g, dtype=torch.float64)
    ell = math.exp(math.log(ell_range[0]) + float(u[0]) * (math.log(ell_range[1]) - math.log(ell_range[0])))
    noise = math.exp(math.log(noise_range[0]) + float(u[1]) * (math.log(noise_range[1]) - math.log(noise_range[0])))
    return X, ell, noise


class TrainingInstance:
    """Points and the top-m eigenpairs of their kernel matrix (from the engine's Lanczos)."""

    def __init__(self, X, U, w, ell, n):
        self.X, self.U, self.w, self.ell, self.n = X, U, w, ell, n


def top_eigenvectors(K, m, steps, key):
    """Top-m Ritz pairs of the dense symmetric K (n, n) from the engine's Lanczos (CGS2):
    U (m', ld) feature-major and eigenvalues (m',) FP64, m' <= m when Lanczos stops early."""
    n = K.shape[0]
    A = Dense(K)
    start = default_start(A, key)[:, 0]
    run = run_lanczos(A, start, min(steps, n), 1e-7)
    ev, vecs, _, st = tridiag_eigh(run, vectors=True)
    if int(st.max()) != 0:
        raise RuntimeError("tridiagonal eigensolve did not converge")
    k = run.iters
    mm = min(m, k)
    S = vecs[0][:, k - mm:].flip(1).contiguous()  # descending
    lam = ev[0, k - mm:].flip(0).contiguous()
    U = torch.zeros((mm, run.ld), dtype=torch.float32, device=K.device)
    N.call("ck_basis_combine", n, k, mm, N.ptr(run.V[0, 1]), run.ld, N.ptr(S), S.stride(0), N.ptr(U), run.ld,
           N.stream_ptr(K.device))
    return U, lam


class _InstanceStack(LinearOperator):
    """Engine-only operator for a batched Lanczos run: probe b is multiplied by its own
    matrix K[b] (CK_OP_DENSE_BATCHED_F32), so many training instances share every launch."""

    def __init__(self, K):
        self.K = K  # (batch, n, n) fp32
        self.device = K.device
        super().__init__(torch.float32, K.shape[1:])

    def hip_op(self):
        from ..ops import _op
        n = self.K.shape[1]
        op = _op(N.CK_OP_DENSE_BATCHED_F32, n)
        op.a, op.lda, op.nnz = self.K.data_ptr(), self.K.stride(1), self.K.stride(0)
        return op, [self.K]

    def _matmat(self, X):
        raise NotImplementedError("per-probe matrices exist only on the engine")


def _lanczos_instances(seeds, n, din, ell_range, steps, device):
    """One batched Lanczos run over the instances of `seeds` (all of order n): every probe
    starts from its instance's default start vector and multiplies by its own kernel matrix."""
    Xs, ells = [], []
    K = torch.empty((len(seeds), n, n), dtype=torch.float32, device=device)
    starts = torch.empty((n, len(seeds)), dtype=torch.float32, device=device)
    for j, seed in enumerate(seeds):
        X, ell, _ = sample_instance(seed, n, din, ell_range)
        Xd = X.to(device)
        N.call("ck_rbf_dense_f32", n, din, N.ptr(Xd), Xd.stride(0), float(ell), N.ptr(K[j]), n, None, n,
               N.stream_ptr(device))
        starts[:, j] = default_start(Dense(K[j]), seed)[:, 0]
        Xs.append(Xd)
        ells.append(ell)
    run = run_lanczos(_InstanceStack(K), starts, min(steps, n), 1e-7)
    return Xs, ells, run


def make_training_set(seeds, n, m=64, din=4, ell_range=(0.05, 1.0), steps=None, device="cuda", batch=256):
    """Training instances for many seeds: batches of up to ``batch`` instances run one
    Lanczos together (CK_OP_DENSE_BATCHED_F32, every launch serves all of them), and their
    tridiagonal eigenproblems are solved in one ck_tridiag_eigh_bisect_f64 launch."""
    steps = steps or (2 * m + 32)
    out = []
    seeds = list(seeds)
    batched_ok = n % 4 == 0
    for c0 in range(0, len(seeds), batch):
        chunk = seeds[c0:c0 + batch]
        runs = []
        if batched_ok:
            Xs, ells, run = _lanczos_instances(chunk, n, din, ell_range, steps, device)
            finite = torch.isfinite(run.diag[:, :run.iters]).all(1) & torch.isfinite(run.subdiag[:, 1:run.iters]).all(1)
            finite = finite.cpu().numpy()
            for j, seed in enumerate(chunk):
                if finite[j]:
                    runs.append((seed, Xs[j], ells[j], run, j))
            redo = [sd for j, sd in enumerate(chunk) if not finite[j]]
        else:
            redo = chunk
        for seed in redo:  # a probe that broke down in the shared run is redone on its own
            X, ell, _ = sample_instance(seed, n, din, ell_range)
            Xd = X.to(device)
            K = rbf_matrix(Xd, ell)
            A = Dense(K)
            run1 = run_lanczos(A, default_start(A, seed)[:, 0], min(steps, n), 1e-7)
            del K, A
            runs.append((seed, Xd, ell, run1, 0))
        groups = {}
        for item in runs:
            groups.setdefault(item[3].iters, []).append(item)
        for k, grp in groups.items():
            nb = len(grp)
            d = torch.stack([it[3].diag[it[4], :k] for it in grp]).contiguous()
            e = torch.stack([it[3].subdiag[it[4], 1:max(k, 2)] for it in grp]).contiguous()
            # QL: RBF spectra put most Ritz values in one cluster (gaps below 1e-3 ||T||), which
            # inverse iteration orthogonalises serially in one thread
            ev, vecs, _, st = tridiag_eigh_batch(d, e, vectors=True, method="ql")
            if int(st.max()) != 0:
                raise RuntimeError("tridiagonal eigensolve did not converge")
            mm = min(m, k)
            for j, (seed, Xd, ell, run, b) in enumerate(grp):
                S = vecs[j][:, k - mm:].flip(1).contiguous()
                lam = ev[j, k - mm:].flip(0).contiguous()
                U = torch.zeros((mm, run.ld), dtype=torch.float32, device=device)
                N.call("ck_basis_combine", n, k, mm, N.ptr(run.V[b, 1]), run.ld, N.ptr(S), S.stride(0), N.ptr(U),
                       run.ld, N.stream_ptr(device))
                out.append((seed, TrainingInstance(Xd, U, torch.clamp(lam, min=0.0).contiguous(), ell, n)))
        del runs, groups
    order = {sd: i for i, sd in enumerate(seeds)}
    out.sort(key=lambda t: order[t[0]])
    return [inst for _, inst in out]


def make_training_instance(seed, n, m=64, din=4, ell_range=(0.05, 1.0), steps=None, device="cuda"):
    X, ell, _ = sample_instance(seed, n, din, ell_range)
    Xd = X.to(device)
    K = rbf_matrix(Xd, ell)
    U, lam = top_eigenvectors(K, m, steps or (2 * m + 32), seed)
    del K
    return TrainingInstance(Xd, U, torch.clamp(lam, min=0.0).contiguous(), ell, n)


def train(net, instances, epochs=1, lr=2e-3, ridge_rel=1e-6, seed=0, log_every=0):
    """Adam on the captured-energy loss, one instance per step in a seeded shuffled order.
    Returns the per-step losses (read back once at the end)."""
    rs = np.random.RandomState(seed)
    losses = []
    for ep in range(epochs):
        for idx in rs.permutation(len(instances)):
            inst = instances[idx]
            G, Z1, Z2 = net.forward(inst.X, save=True)
            loss, dG, _ = energy_loss_grad(G, inst.U, inst.w, inst.n, ridge_rel)
            grad = net.backward(inst.X, Z1, Z2, dG)
            net.adam(grad, lr)
            losses.append(loss)
    return torch.cat(losses).cpu().numpy() if losses else np.zeros(0)


def mean_loss(net, instances, ridge_rel=1e-6):
    vals = []
    for inst in instances:
        G = net.forward(inst.X)
        loss, _, _ = energy_loss_grad(G, inst.U, inst.w, inst.n, ridge_rel)
        vals.append(loss)
    return float(torch.cat(vals).mean())


# ----------------------------------------------------------------------------- preconditioners
def dense_sketch_bf16(lowered, G, n, split=True):
    """Y = A G on BF16 MFMA (``ck_sketch_bf16``) for a lowered dense fp32 operator A and a
    feature-major sketch block G (k, ld) with k = 32 or 64; G is first rounded to bf16, A to
    a bf16 pair hi + lo (split, the default) or to one bf16 value.
    Returns (Y, G~) with G~ the rounded sketch as fp32 (the one the product used, so that a
    Gram matrix G~^T Y stays consistent with Y), or None when the shapes fall outside the
    kernel (then the caller's FP32 path runs)."""
    if lowered is None or lowered[0].kind != N.CK_OP_DENSE_F32:
        return None
    op = lowered[0]
    k, ld = G.shape
    dev = G.device
    G16 = torch.empty((k, ld), dtype=torch.int16, device=dev)
    if k not in (32, 64) or not N.q("ck_sketch_bf16_supported", n, k, op.lda, ctypes.c_void_p(op.a), ld, N.ptr(G16)):
        return None
    Gr = torch.zeros((k, ld), dtype=torch.float32, device=dev)
    Y = torch.zeros((k, ld), dtype=torch.float32, device=dev)
    s = N.stream_ptr(dev)
    N.call("ck_round_bf16", n, k, N.ptr(G), ld, N.ptr(G16), ld, N.ptr(Gr), ld, s)
    N.call("ck_sketch_bf16", n, k, ctypes.c_void_p(op.a), op.lda, float(op.scale), float(op.shift), N.ptr(G16), ld,
           N.ptr(Gr), ld, N.ptr(Y), ld, 1 if split else 0, s)
    return Y, Gr


def woodbury_core(F, noise):
    """(noise I + F^T F)^-1 (FP64) for a row-major factor F (n, k), with tr(F^T F) and
    max diag(F^T F), all formed on the engine."""
    n, k = F.shape
    FtF = gram(F, k, 1, F, k, 1, n, k, k)
    core = spd_inverse(FtF, jitter_abs=noise, inv64=True)
    diag = torch.diagonal(FtF)
    return core["inv64"], float(diag.sum()), float(diag.max())


def _dense_matrix(K):
    if isinstance(K, Dense):
        K = K.A
    if not (torch.is_tensor(K) and K.dim() == 2 and K.dtype == torch.float32 and K.is_cuda):
        raise TypeError("the kernel matrix must be a dense fp32 device tensor (or cola_amd.Dense of one)")
    return K


class LearnedPreconditioner(WoodburyInverse):
    """P = (noise I + F F^T)^-1 with F the Nystrom completion of the network's sketch of K."""

    def __init__(self, K, points, net, noise, jitter_rel=1e-6, bf16_sketch=True):
        t0 = time.time()
        K = _dense_matrix(K)
        n = K.shape[0]
        G = net.forward(points)
        k = G.shape[0]
        ld = G.stride(0)
        Kop = Dense(K)
        aop = lower(Kop)
        # the sketch product K G: one BF16 MFMA pass over K where the shapes allow (the
        # rounding perturbs the preconditioner only), else the FP32 apply
        got = dense_sketch_bf16(aop, G, n) if bf16_sketch else None
        self.sketch = "bf16_mfma" if got is not None else "fp32"
        if got is not None:
            Y, G = got
        else:
            work, wb = workspace([aop[0]], k, K.device)
            Y = torch.zeros_like(G)
            apply_blocks(Kop, aop, G, Y, n, k, work, wb)
        C = gram_blocks(G, Y, n)
        fac = spd_inverse(C, jitter_rel=jitter_rel, linvT32=True)
        F = torch.empty((n, k), dtype=torch.float32, device=K.device)
        rows_gemm(Y, 1, ld, fac["linvT32"], F, k, 1, n)
        core, captured, top = woodbury_core(F, noise)
        super().__init__(F, noise, core_inv=core)
        self.trace_K = float(torch.diagonal(K).sum())
        self.captured_trace = captured
        self.residual_trace = self.trace_K - captured
        self.top_captured = top
        self.build_time = time.time() - t0


class PivotedCholeskyPreconditioner(WoodburyInverse):
    """P = (noise I + L L^T)^-1 with L the rank-r pivoted Cholesky factor of K."""

    def __init__(self, K, rank, noise):
        t0 = time.time()
        K = _dense_matrix(K)
        n = K.shape[0]
        ld = ld_of(n)
        dev = K.device
        L = torch.zeros((rank, ld), dtype=torch.float32, device=dev)
        piv = torch.empty(rank, dtype=torch.int32, device=dev)
        dpiv = torch.empty(rank, dtype=torch.float32, device=dev)
        wb = N.q("ck_pivchol_workspace_bytes", n)
        work = torch.empty((wb + 3) // 4, dtype=torch.float32, device=dev)
        s = N.stream_ptr(dev)
        N.call("ck_pivchol_dense_f32", n, N.ptr(K), K.stride(0), rank, N.ptr(L), ld, N.ptr(piv), N.ptr(dpiv),
               N.ptr(work), wb, s)
        F = torch.empty((n, rank), dtype=torch.float32, device=dev)
        N.call("ck_transpose_f32", n, rank, N.ptr(L), ld, N.ptr(F), rank, 0, s)
        core, captured, top = woodbury_core(F, noise)
        super().__init__(F, noise, core_inv=core)
        self.pivots, self.pivot_diag = piv, dpiv
        self.trace_K = float(torch.diagonal(K).sum())
        self.captured_trace = captured
        self.residual_trace = self.trace_K - captured
        self.top_captured = top
        self.build_time = time.time() - t0


def select_preconditioner(K, points, noise, net=None, rank=64, candidates=None, A=None, probe_steps=24):
    """Pick plain CG, the learned or the pivoted-Cholesky preconditioner for A = K + noise I.

    With the system operator ``A``, each choice runs ``probe_steps`` CG steps on one
    fixed Gaussian probe (the engine's counter stream, seed 0) and the one with the
    smallest residual wins, so a preconditioner is kept only when it measurably cuts
    iterations.  Without ``A`` the rule is the bound kappa(P A) <= 1 + tr(E)/noise:
    the candidate with the smaller Nystrom error trace tr(E) = tr(K - F F^T) is kept
    unless tr(E) exceeds its largest captured diagonal.  ``candidates`` may pass
    prebuilt preconditioners by name.  Returns (P, report)."""
    cands = dict(candidates or {})
    if net is not None and "learned" not in cands:
        cands["learned"] = LearnedPreconditioner(K, points, net, noise)
    if "pivoted_cholesky" not in cands:
        cands["pivoted_cholesky"] = PivotedCholeskyPreconditioner(K, rank, noise)
    report = {name: {"residual_trace": P.residual_trace, "captured_trace": P.captured_trace,
                     "build_s": P.build_time} for name, P in cands.items()}
    if A is not None:
        from ._engine import probes
        from .cg import cg
        n = A.shape[0]
        Z, _ = probes("normal", n, 1, 0, K.device)
        z = Z[0, :n].contiguous()
        scores = {}
        for name, P in [("none", None)] + list(cands.items()):
            _, info = cg(A, z, P=P, tol=0.0, max_iters=int(probe_steps), chunk=int(probe_steps))
            errs = info["errors"]
            scores[name] = float(errs[-1]) if len(errs) else float("inf")
            if name in report:
                report[name]["probe_residual"] = scores[name]
        report["none"] = {"probe_residual": scores["none"]}
        best = min(scores, key=scores.get)
        report["choice"] = best
        return (None if best == "none" else cands[best]), report
    best = min(cands, key=lambda c: cands[c].residual_trace)
    P = cands[best]
    if P.residual_trace >= P.top_captured:
        report["choice"] = "none"
        return None, report
    report["choice"] = best
    return P, report
