"""The fused PCG step for a dense operator with a Woodbury preconditioner (dense matvec with
p.Ap partials in its epilogue, then wbf_xr / wbf_core / wbf_z, cg.hip), taken for 3..8 rhs:
against the unfused path (taken for 1-2 rhs) on the same columns, and against the oracle's
restatement of run_batched_cg (cg.py:94-178) with the same preconditioner."""
import numpy as np
import pytest
import torch

import cola_amd as ca
from cola_amd.linalg.cg import cg, run_batched_cg
from oracle import cg as ocg
from oracle import learned as ol

pytestmark = pytest.mark.gpu


def rbf_system(n, ell, seed):
    X = torch.rand(n, 3, generator=torch.Generator().manual_seed(seed))
    return torch.exp(-torch.cdist(X, X)**2 / (2 * ell * ell))


def low_rank_factor(K, k, seed):
    """F = K G (G^T K G)^-1/2 for a Gaussian sketch G: a rank-k Nystrom factor, FP64 on the host."""
    G = torch.randn(K.shape[0], k, generator=torch.Generator().manual_seed(seed), dtype=torch.float64)
    Y = K.double() @ G
    L = torch.linalg.cholesky(G.T @ Y + 1e-10 * torch.eye(k, dtype=torch.float64))
    return torch.linalg.solve_triangular(L, Y.T, upper=False).T.float()


@pytest.mark.parametrize("n,k,B", [(2048, 64, 4), (1000, 17, 8), (4100, 64, 3)])
def test_fused_woodbury_pcg_matches_unfused_and_oracle(dev, n, k, B):
    noise = 0.1
    K = rbf_system(n, 0.5, n + k)
    F = low_rank_factor(K, k, 7)
    P = ca.WoodburyInverse(F.to(dev), noise)
    A = ca.PSD(ca.Dense(K.to(dev)) + noise * ca.I_like(ca.Dense(K.to(dev))))
    b = torch.randn(n, B, generator=torch.Generator().manual_seed(B))
    # fused (B >= 3 rhs) against unfused (two rhs) on the first two columns
    xf, inf_f = cg(A, b.to(dev), P=P, tol=1e-6, max_iters=500)
    xu, inf_u = cg(A, b[:, :2].contiguous().to(dev), P=P, tol=1e-6, max_iters=500)
    s32 = float(np.float32(noise))
    A64 = K.double() + s32 * torch.eye(n, dtype=torch.float64)
    def res(x, bb):
        return (A64 @ x.double().cpu() - bb.double()).norm(dim=0) / bb.double().norm(dim=0)

    rf, ru = res(xf[:, :2], b[:, :2]), res(xu, b[:, :2])
    d = float((xf[:, :2] - xu).norm() / xu.norm())
    print(n, k, B, "steps fused", inf_f["iterations"], "unfused(2)", inf_u["iterations"], "x diff", d,
          "fp64 residual fused", rf.tolist(), "unfused", ru.tolist(), flush=True)
    # the same columns to FP32 rounding: the step counts (the fused run also waits for its other
    # columns), the solutions, and the true residuals FP32 CG attains at this conditioning
    assert abs(inf_f["iterations"] - inf_u["iterations"]) <= max(3, 0.05 * inf_u["iterations"])
    assert d < 1e-4
    assert float(rf.max()) <= 3 * float(ru.max()) + 1e-6
    # the residual history of 25 steps against the oracle with the same (FP64-applied) Woodbury P:
    # held to the oracle's FP32 run while that agrees with its FP64 run, and to the FP64 run
    # as closely as the FP32 oracle stays to it after that
    xg, _, kg, ig = run_batched_cg(A, b.to(dev), None, 25, 0.0, P)
    Fd = F.double().numpy()
    Pw = lambda V: torch.from_numpy(ol.woodbury_apply(Fd, noise, V.double().numpy())).to(V.dtype)  # noqa: E731
    xo, _, ko, io = ocg.run_batched_cg(lambda V: K @ V + s32 * V, b, None, 25, 0.0, Pw)
    _, _, _, i64 = ocg.run_batched_cg(lambda V: A64 @ V, b.double(), None, 25, 0.0, Pw)
    ge, oe, e64 = np.log10(ig["errors"]), np.log10(io["errors"]), np.log10(i64["errors"])
    spread = np.abs(oe - e64)
    K32 = int(np.argmax(spread > 0.005)) if (spread > 0.005).any() else len(oe)
    print("oracle FP32 tracks FP64 for", K32, "steps; engine", ig["errors"][::6], "oracle", io["errors"][::6],
          flush=True)
    assert kg == ko == 25 and K32 >= 4
    assert np.allclose(ge[:K32], oe[:K32], atol=0.01)
    assert np.abs(ge - e64).max() <= 3 * spread.max() + 0.02
