"""Hutchinson diagonal / trace estimation, restated from
cola/linalg/trace/diagonal_estimation.py:158-210 and diag_trace.py:139-142.

``hutchinson_diag_estimate`` keeps the reference's semantics (blocks of
min(100, n) probes drawn from the SHA-256 key chain, adaptive standard-error
stop, at least one block).  ``hutch_with_probes`` is the same estimator for an
explicit probe block, which is what the GPU parity tests use.
"""
import torch

from .cg import while_loop_winfo
from .rng import PRNGKey, next_key, torch_randn


def hutchinson_diag_estimate(matmat, n, dtype=torch.float32, k=0, tol=3e-2, max_iters=10000, rand='normal',
                             key=None, probe_fn=None):
    """diagonal_estimation.py:158-210 (k = 0 handled fully; roll/slice as in :191-194).

    probe_fn(key, n, bs) may replace the keyed torch randn (diagonal_estimation.py:188)."""
    bs = min(100, n)
    assert tol > 1e-3
    assert rand in ['normal', 'rademacher']
    key = PRNGKey(42) if key is None else key

    def body(state):
        i, diag_sum, diag_sumsq, key = state
        key = next_key(key)
        if probe_fn is None:
            z = torch_randn(n, bs, dtype=dtype, key=key)
        else:
            z = probe_fn(key, n, bs)
        if rand == 'rademacher':
            z = torch.sign(z)
        z2 = torch.roll(z, -k, 0)
        if k <= 0:
            z2[0:abs(k)] = 0
        else:
            z2[-abs(k):] = 0
        slc = slice(abs(k), None) if -k > 0 else slice(None, -abs(k) or None)
        estimator = (matmat(z) * z2)[slc]
        return i + 1, diag_sum + estimator.sum(-1), diag_sumsq + (estimator**2).sum(-1), key

    def err(state):
        i, diag_sum, diag_sumsq, _ = state
        mean = diag_sum / (i * bs)
        stderr = torch.sqrt((diag_sumsq / (i * bs) - mean**2) / (i * bs))
        return torch.mean(stderr / torch.maximum(mean.abs(), .1 * torch.ones_like(mean)))

    def cond(state):
        return (state[0] == 0) or ((state[0] < max_iters) and bool(err(state) > tol))

    zeros = torch.zeros((n - abs(k), ), dtype=dtype)

    def errorfn(state):
        return err(state) if state[0] > 0 else torch.tensor(float('nan'))

    (nblk, diag_sum, *_), info = while_loop_winfo(errorfn, cond, body, (0, zeros, zeros, key))
    return diag_sum / (nblk * bs), info


def hutch_with_probes(matmat, Z):
    """One block of the estimator for explicit probes Z (n, B): returns
    (diag_sum, diag_sumsq) with diag_sum[i] = sum_b z_b[i] (A z_b)[i]."""
    est = matmat(Z) * Z
    return est.sum(-1), (est**2).sum(-1)


def trace_from_probes(matmat, Z):
    """(1/B) sum_b z_b^T A z_b = diag(A) estimate summed (diag_trace.py:139-142)."""
    s, _ = hutch_with_probes(matmat, Z)
    return s.sum() / Z.shape[1]
