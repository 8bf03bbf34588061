"""Learned low-rank preconditioner, pivoted Cholesky and their small algebra (CPU, FP64).

TEST INFRASTRUCTURE ONLY (see the package docstring).  These parts have no
counterpart in wilson-labs/cola; they enter its solver as ``CG(P=...)``
(cola/linalg/inverse/cg.py:39-69, z = P r at cg.py:150).  Each function states the
mathematics it restates, and tests/test_oracle_learned.py pins it:

* ``mlp_forward`` / ``mlp_grads`` -- gradients from torch.autograd in FP64, checked
  against central finite differences;
* ``energy_loss`` -- L = 0 when span(G) contains the targets, L = 1 when G is
  orthogonal to them, invariance under G -> G R, gradient vs finite differences;
* ``adam_step`` -- against torch.optim.Adam;
* ``pivoted_cholesky`` -- against LAPACK's dpstrf (same greedy max-diagonal rule),
  exact recovery of rank-r matrices, L L^T = P A P^T at full rank;
* ``nystrom_factor`` / ``woodbury_apply`` -- K_hat = K for a sketch spanning range(K),
  Woodbury against a dense inverse;
* ``rbf_dense`` -- against scipy's cdist.
"""
import numpy as np
import torch
import torch.nn.functional as Fn

HIDDEN = 128


def param_layout(din, dout, hidden=HIDDEN):
    """Offsets of W1 (din x h), b1, W2 (h x h), b2, W3 (h x dout), b3 in the flat vector."""
    sizes = [("W1", (din, hidden)), ("b1", (hidden, )), ("W2", (hidden, hidden)), ("b2", (hidden, )),
             ("W3", (hidden, dout)), ("b3", (dout, ))]
    out, off = {}, 0
    for name, shp in sizes:
        cnt = int(np.prod(shp))
        out[name] = (off, shp)
        off += cnt
    return out, off


def unpack(params, din, dout):
    lay, _ = param_layout(din, dout)
    return {k: params[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items()}


def init_params(din, dout, seed):
    """torch.nn.Linear's default ranges: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases."""
    rs = np.random.RandomState(seed)
    lay, total = param_layout(din, dout)
    out = np.zeros(total, dtype=np.float32)
    fan = {"W1": din, "b1": din, "W2": HIDDEN, "b2": HIDDEN, "W3": HIDDEN, "b3": HIDDEN}
    for k, (o, s) in lay.items():
        bound = 1.0 / np.sqrt(fan[k])
        out[o:o + int(np.prod(s))] = rs.uniform(-bound, bound, size=int(np.prod(s))).astype(np.float32)
    return out


def mlp_forward(params, X, din, dout):
    """G = gelu(gelu(X W1 + b1) W2 + b2) W3 + b3 with the exact (erf) GELU; also Z1, Z2."""
    p = unpack(params, din, dout)
    Z1 = X @ p["W1"] + p["b1"]
    Z2 = Fn.gelu(Z1) @ p["W2"] + p["b2"]
    G = Fn.gelu(Z2) @ p["W3"] + p["b3"]
    return G, Z1, Z2


def mlp_grads(params, X, dG, din, dout):
    """d(sum(G * dG))/d params by torch.autograd (FP64)."""
    prm = params.detach().clone().double().requires_grad_(True)
    G, _, _ = mlp_forward(prm, X.double(), din, dout)
    (G * dG.double()).sum().backward()
    return prm.grad.detach()


def energy_loss(G, U, w, ridge_rel=1e-6):
    """L = 1 - tr(S^-1 G^T U diag(w) U^T G) / sum(w), S = G^T G + delta I with
    delta = ridge_rel * tr(G^T G) / k held constant (no gradient through delta)."""
    k = G.shape[1]
    S0 = G.T @ G
    delta = (ridge_rel * torch.trace(S0) / k).detach()
    S = S0 + delta * torch.eye(k, dtype=G.dtype)
    T = U.T @ G
    E = T.T @ (w[:, None] * T)
    return 1.0 - torch.trace(torch.linalg.solve(S, E)) / w.sum()


def energy_loss_grad(G, U, w, ridge_rel=1e-6):
    g = G.detach().clone().double().requires_grad_(True)
    L = energy_loss(g, U.double(), w.double(), ridge_rel)
    L.backward()
    return float(L.detach()), g.grad.detach()


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """torch.optim.Adam (no weight decay, no amsgrad), one step; returns new (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    bc1 = 1 - beta1**step
    bc2 = 1 - beta2**step
    denom = np.sqrt(v) / np.sqrt(bc2) + eps
    return p - (lr / bc1) * m / denom, m, v


def pivoted_cholesky(A, rank):
    """Greedy pivoted Cholesky (Harbrecht, Peters & Schneider 2012): pivot = argmax of the
    residual diagonal (first index on ties); L[:, j] = (A[:, p] - L[:, :j] L[p, :j]^T) / sqrt(d_p).
    Returns L (n x rank), pivots, pivot diagonals, residual diagonal."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    d = np.diag(A).copy()
    L = np.zeros((n, rank))
    piv = np.zeros(rank, dtype=np.int64)
    dpiv = np.zeros(rank)
    for j in range(rank):
        p = int(np.argmax(d))
        piv[j], dpiv[j] = p, d[p]
        if d[p] <= 0:
            continue
        col = (A[:, p] - L[:, :j] @ L[p, :j]) / np.sqrt(d[p])
        L[:, j] = col
        d = d - col * col
        d[p] = 0.0
    return L, piv, dpiv, d


def nystrom_factor(K, G, jitter_rel=1e-6):
    """F = K G L^-T with G^T K G + jitter = L L^T, so F F^T = K G (G^T K G)^-1 G^T K (Nystrom)."""
    K = np.asarray(K, dtype=np.float64)
    G = np.asarray(G, dtype=np.float64)
    Y = K @ G
    C = G.T @ Y
    k = C.shape[0]
    C = C + jitter_rel * np.trace(C) / k * np.eye(k)
    Lc = np.linalg.cholesky(C)
    return np.linalg.solve(Lc, Y.T).T


def woodbury_apply(F, s, V):
    """(s I + F F^T)^-1 V = (V - F (s I + F^T F)^-1 F^T V) / s."""
    F = np.asarray(F, dtype=np.float64)
    core = s * np.eye(F.shape[1]) + F.T @ F
    return (V - F @ np.linalg.solve(core, F.T @ V)) / s


def rbf_dense(P, ell):
    """K_ij = exp(-|p_i - p_j|^2 / (2 l^2)) in FP64."""
    P = np.asarray(P, dtype=np.float64)
    D2 = ((P[:, None, :] - P[None, :, :])**2).sum(-1)
    return np.exp(-D2 / (2 * ell * ell))
