"""The learned-preconditioner and batched workloads at their full sizes.

* Dense RBF system on 16384 points in [0,1]^4 with 8 rhs (lengthscale and noise drawn as in the
  bench): the learned preconditioner inside FP64-residual refinement reaches an FP64-checked
  relative residual of 1e-6 for every rhs; in FP32 the step counts order learned < pivoted
  Cholesky < plain; and the engine's first 20 preconditioned steps match the oracle's
  restatement of cola's run_batched_cg (cg.py:94-178) with the same preconditioner.
* 512 uneven RBF systems, n ~ U{256..4096} in [0,1]^3, noise 1e-2: FP64-residual refinement
  around batched PCG (rank-64 pivoted Cholesky, bf16 storage where the screen admits it)
  reaches 1e-8 for every system, checked per system in FP64, and 10 plain FP32 batched CG
  steps match the oracle's CG on three systems.
"""
import math

import numpy as np
import pytest
import torch

import cola_amd as ca
from cola_amd.linalg import learned as L
from cola_amd.linalg.batched import BatchedCG, BatchedMixedSolver, BatchedPrecond, BatchedSystems
from cola_amd.linalg.cg import cg, run_batched_cg
from cola_amd.linalg.mixed import HipMixedCG, MixedSolver
from oracle import cg as ocg
from oracle import learned as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).norm() / max(b.norm(), 1e-30))


@pytest.fixture(scope="module")
def trained_net(dev):
    """The bench's network: 2000 seeded 1024-point instances, top-64 Lanczos targets, 4 Adam epochs."""
    insts = L.make_training_set(range(2000), 1024, m=64, ell_range=(0.05, 1.0), device=dev)
    net = L.PointMLP(4, 64, device=dev, seed=0)
    L.train(net, insts, epochs=4, lr=2e-3, seed=0)
    del insts
    return net


def test_learned_config_full_size(dev, trained_net):
    n, B = 16384, 8
    X, ell, noise = L.sample_instance(100003, n)  # lengthscale 0.419, noise 0.0676
    Xd = X.to(dev)
    K = L.rbf_matrix(Xd, ell)
    A = ca.PSD(ca.Dense(K) + noise * ca.I_like(ca.Dense(K)))
    b = torch.randn(n, B, generator=torch.Generator().manual_seed(200003)).to(dev)
    s32 = float(np.float32(noise))
    K64 = K.double()

    def fp64_res(x):
        return ((K64 @ x.double() + s32 * x.double()) - b.double()).norm(dim=0) / b.double().norm(dim=0)

    Pl = L.LearnedPreconditioner(K, Xd, trained_net, noise)
    Pc = L.PivotedCholeskyPreconditioner(K, 64, noise)
    # FP32 step counts to the bench's recursive target: learned < pivoted Cholesky < plain
    steps = {}
    for name, P in (("learned", Pl), ("pivoted_cholesky", Pc), ("plain", None)):
        _, info = cg(A, b, P=P, tol=1e-7, max_iters=1000, chunk=16)
        steps[name] = info["iterations"] - 1
    print(f"ell {ell:.3f} noise {noise:.4f}: FP32 steps {steps}", flush=True)
    assert steps["learned"] < steps["pivoted_cholesky"] < steps["plain"]
    # FP64-residual refinement around the learned PCG: every rhs to 1e-6
    x, info = MixedSolver(A, B, HipMixedCG(tol=1e-6, inner_tol=1e-4, inner_max_iters=1000, max_refinements=20,
                                           P=Pl)).solve(b)
    res = fp64_res(x)
    print(f"refinements {info['refinements']}, inner {info['inner_iterations']}, max FP64 residual "
          f"{float(res.max()):.2e}", flush=True)
    assert info["converged"] and float(res.max()) <= 1e-6
    # the first 20 preconditioned FP32 steps against the oracle's restatement on the host
    xg, _, kg, ig = run_batched_cg(A, b, None, 20, 0.0, Pl)
    Kc, bc = K.cpu(), b.cpu()
    F = Pl.F.cpu().double().numpy()
    xo, _, ko, io = ocg.run_batched_cg(lambda V: Kc @ V + s32 * V, bc, None, 20, 0.0,
                                       lambda V: torch.from_numpy(ol.woodbury_apply(F, Pl.s, V.double().numpy())).float())
    assert kg == ko == 20
    ge, oe = np.log10(ig["errors"]), np.log10(io["errors"])
    print("residual history engine", ig["errors"][:5], "oracle", io["errors"][:5], flush=True)
    assert np.allclose(ge, oe, atol=0.01)
    assert rel(xg, xo) < 1e-3


def gp_systems(nsys):
    """The batched workload: sizes ~ U{256..4096} (seed 0); system i from manual_seed(i): points
    U[0,1]^3, lengthscale LogU[0.05,0.5], one N(0,1) rhs."""
    sizes = torch.randint(256, 4097, (nsys, ), generator=torch.Generator().manual_seed(0)).tolist()
    pts, ells, rhs = [], [], []
    for i, n in enumerate(sizes):
        g = torch.Generator().manual_seed(i)
        pts.append(torch.rand(n, 3, generator=g))
        u = float(torch.rand(1, generator=g, dtype=torch.float64))
        ells.append(math.exp(math.log(0.05) + u * (math.log(0.5) - math.log(0.05))))
        rhs.append(torch.randn(n, generator=g))
    return sizes, pts, ells, rhs


def test_batched_config_full_512_systems(dev):
    sizes, pts, ells, rhs = gp_systems(512)
    sy = BatchedSystems.rbf(pts, ells, 1e-2, dev, bf16=True)
    b = sy.concat([r.to(dev) for r in rhs])
    pcc = BatchedPrecond.pivoted_cholesky(sy, 64)
    solver = BatchedMixedSolver(sy, rows_per_tile=16, inner_max_iters=1000, bf16=True, precond=pcc)
    x, info = solver.solve(b, tol=1e-8, inner_tol=3e-5, chunk=16)
    rel_all = solver.residual(np.ones(len(sizes), dtype=bool))
    print(f"512 systems (sum n^2 = {sum(n * n for n in sizes):.3e}): refinements {info['refinements']}, "
          f"inner steps {int(info['inner_iterations'].sum())}, worst FP64 residual {rel_all.max():.2e}", flush=True)
    assert info["converged"] and float(rel_all.max()) <= 1e-8
    s32 = float(np.float32(1e-2))
    for s in (0, 255, 511):  # independent dense FP64 products
        n = sizes[s]
        Ad = sy.matrix(s).double() + s32 * torch.eye(n, dtype=torch.float64, device=dev)
        bs = rhs[s].double().to(dev)
        assert float((Ad @ sy.segment(x, s).double() - bs).norm() / bs.norm()) <= 1.2e-8
    # 10 plain FP32 batched CG steps against the oracle's CG on the three smallest systems; the bar is
    # the spread FP32 rounding alone gives these ill-conditioned systems (the oracle in FP32 vs FP64)
    bcg = BatchedCG(sy, 16, max_iters=10)
    xp, ks = bcg.solve(b, tol_rel=0.0, tol_abs=0.0, chunk=10)
    for s in np.argsort(sizes)[:3]:
        n = sizes[s]
        Af = sy.matrix(int(s)).cpu() + s32 * torch.eye(n)
        xo, _, ko, _ = ocg.run_batched_cg(lambda V: Af @ V, rhs[s][:, None], None, 10, 0.0, None)
        A64 = Af.double()
        x64, _, _, _ = ocg.run_batched_cg(lambda V: A64 @ V, rhs[s][:, None].double(), None, 10, 0.0, None)
        spread = rel(xo, x64)
        d = rel(sy.segment(xp, int(s)), xo[:, 0])
        print(f"system {s} (n={n}): engine {int(ks[s])} steps, oracle {ko}, rel diff {d:.2e} (fp32 spread {spread:.2e})",
              flush=True)
        assert int(ks[s]) == ko == 10
        assert d < max(1e-5, 5 * spread)
