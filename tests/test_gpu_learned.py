"""Learned preconditioner on the engine: MLP forward/backward and Adam against the
oracle's torch-autograd restatement, the small FP64 algebra, pivoted Cholesky
against the oracle (itself pinned to LAPACK dpstrf), the dense RBF builder, FP64
dense residuals, and an ablation showing what the trained network gains."""
import numpy as np
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd.linalg import learned as L
from cola_amd.linalg.cg import cg
from cola_amd.linalg.mixed import HipMixedCG, MixedSolver
from oracle import learned as ol

pytestmark = pytest.mark.gpu


# dout = 7 exercises the padded W3 rows (dout not a multiple of 4) and layer 3's uneven
# split over the four waves; (8, 128) are the largest supported dimensions
@pytest.mark.parametrize("n,din,dout", [(1000, 4, 64), (64, 4, 64), (77, 3, 40), (150, 2, 7), (65, 8, 128)])
def test_mlp_forward_matches_oracle(dev, n, din, dout):
    g = torch.Generator().manual_seed(n)
    X = torch.rand(n, din, generator=g)
    net = L.PointMLP(din, dout, device=dev, seed=5)
    G, Z1, Z2 = net.forward(X.to(dev), save=True)
    p = torch.from_numpy(ol.init_params(din, dout, 5)).double()
    assert torch.equal(net.params.cpu(), p.float())
    Gr, Z1r, Z2r = ol.mlp_forward(p, X.double(), din, dout)
    assert torch.allclose(G[:, :n].T.cpu().double(), Gr, rtol=1e-5, atol=1e-5)
    assert torch.allclose(Z1[:, :n].T.cpu().double(), Z1r, rtol=1e-5, atol=1e-6)
    assert torch.allclose(Z2[:, :n].T.cpu().double(), Z2r, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("n,din,dout", [(300, 4, 64), (130, 3, 40), (150, 2, 7), (70, 8, 128)])
def test_mlp_backward_matches_autograd(dev, n, din, dout):
    g = torch.Generator().manual_seed(7)
    X = torch.rand(n, din, generator=g)
    dG = torch.randn(n, dout, generator=g)
    net = L.PointMLP(din, dout, device=dev, seed=2)
    _, Z1, Z2 = net.forward(X.to(dev), save=True)
    ld = Z1.shape[1]
    dGb = torch.zeros((dout, ld), device=dev)
    dGb[:, :n] = dG.T.to(dev)
    grad = net.backward(X.to(dev), Z1, Z2, dGb, scale=0.5).cpu().double()
    ref = 0.5 * ol.mlp_grads(torch.from_numpy(ol.init_params(din, dout, 2)), X, dG, din, dout)
    err = (grad - ref).abs().max() / ref.abs().max()
    assert err < 2e-5, err
    # repeatable bit for bit (fixed-order reductions)
    again = net.backward(X.to(dev), Z1, Z2, dGb, scale=0.5).cpu().double()
    assert torch.equal(grad, again)


def test_adam_matches_oracle(dev):
    net = L.PointMLP(4, 64, device=dev, seed=1)
    p = net.params.cpu().double().numpy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    g = torch.Generator().manual_seed(3)
    for t in range(1, 4):
        gr = torch.randn(net.count, generator=g)
        net.adam(gr.to(dev), 1e-3)
        p, m, v = ol.adam_step(p, gr.double().numpy(), m, v, t, 1e-3)
    assert np.allclose(net.params.cpu().double().numpy(), p, rtol=0, atol=2e-6)


def test_gram_rows_gemm_spd(dev):
    g = torch.Generator().manual_seed(4)
    n, ka, kb = 3000, 40, 64
    X = torch.randn(n, ka, generator=g)
    Y = torch.randn(n, kb, generator=g)
    Xb = X.T.contiguous().to(dev)  # feature-major
    Yd = Y.contiguous().to(dev)  # row-major
    C = L.gram(Xb, 1, n, Yd, kb, 1, n, ka, kb).cpu()
    assert torch.allclose(C, X.double().T @ Y.double(), rtol=1e-12, atol=1e-10)
    W = torch.randn(ka, kb, generator=g)
    out = torch.ones((n, kb), device=dev)
    L.rows_gemm(Xb, 1, n, W.to(dev), out, kb, 1, n, alpha=2.0, beta=-1.0)
    assert torch.allclose(out.cpu().double(), 2 * X.double() @ W.double() - 1, rtol=1e-5, atol=1e-4)
    A = (X.double().T @ X.double()) / n
    res = L.spd_inverse(A.to(dev), jitter_abs=0.25, inv64=True, inv32=True, linvT32=True)
    ref = torch.linalg.inv(A + 0.25 * torch.eye(ka, dtype=torch.float64))
    assert torch.allclose(res["inv64"].cpu(), ref, rtol=1e-10, atol=1e-12)
    Lc = torch.linalg.cholesky(A + 0.25 * torch.eye(ka, dtype=torch.float64))
    assert torch.allclose(res["linv"].cpu(), torch.linalg.inv(Lc), rtol=1e-10, atol=1e-12)
    assert torch.allclose(res["linvT32"].cpu().double(), torch.linalg.inv(Lc).T, rtol=1e-5, atol=1e-6)
    bad = -torch.eye(5, dtype=torch.float64)
    with pytest.raises(np.linalg.LinAlgError):
        L.spd_inverse(bad.to(dev))


def test_gram_many_chunks_and_rows_gemm_remainder(dev):
    # n > 1024 chunks of 64 rows: each partial covers 128 rows; kin = 37 leaves a remainder
    # after the eight-column groups of rows_gemm, kout = 45 a partial second column slab
    g = torch.Generator().manual_seed(11)
    n, ka, kb = 70000, 33, 17
    X = torch.randn(n, ka, generator=g)
    Y = torch.randn(n, kb, generator=g)
    assert N.q("ck_gram_workspace_bytes", n, ka, kb) == 8 * 547 * ka * kb
    C = L.gram(X.T.contiguous().to(dev), 1, n, Y.T.contiguous().to(dev), 1, n, n, ka, kb).cpu()
    assert torch.allclose(C, X.double().T @ Y.double(), rtol=1e-12, atol=1e-9)
    kin, kout, m = 37, 45, 999
    Xs = torch.randn(m, kin, generator=g)
    W = torch.randn(kin, kout, generator=g)
    out = torch.zeros((m, kout), device=dev)
    L.rows_gemm(Xs.T.contiguous().to(dev), 1, m, W.to(dev), out, kout, 1, m)
    assert torch.allclose(out.cpu().double(), Xs.double() @ W.double(), rtol=1e-5, atol=1e-4)


def test_energy_loss_and_grad_match_autograd(dev):
    g = torch.Generator().manual_seed(5)
    n, m, k = 2000, 20, 32
    U, _ = torch.linalg.qr(torch.randn(n, m, generator=g, dtype=torch.float64))
    w = torch.linspace(5, 1, m, dtype=torch.float64)
    G = torch.randn(n, k, generator=g)
    ld = 2048
    Gb = torch.zeros((k, ld), device=dev)
    Gb[:, :n] = G.T.to(dev)
    Ub = torch.zeros((m, ld), device=dev)
    Ub[:, :n] = U.T.float().to(dev)
    loss, dG, st = L.energy_loss_grad(Gb, Ub, w.to(dev), n, 1e-6)
    ref_loss, ref_grad = ol.energy_loss_grad(G.double(), U.float().double(), w, 1e-6)
    assert int(st.item()) == 0
    assert abs(float(loss.item()) - ref_loss) < 1e-6
    got = dG[:, :n].T.cpu().double()
    assert (got - ref_grad).abs().max() / ref_grad.abs().max() < 1e-4


@pytest.mark.parametrize("n,rank", [(700, 24), (3000, 64)])
def test_pivoted_cholesky_matches_oracle(dev, n, rank):
    rs = np.random.RandomState(n)
    B = rs.normal(size=(n, 40)).astype(np.float32)
    A = (B @ B.T / 40 + np.diag(rs.uniform(0.5, 4.0, size=n))).astype(np.float32)
    Ad = torch.from_numpy(A).to(dev)
    P = L.PivotedCholeskyPreconditioner(Ad, rank, 0.1)
    Lr, piv, dpiv, d = ol.pivoted_cholesky(A.astype(np.float64), rank)
    piv_gpu = P.pivots.cpu().numpy()
    # pivots agree until FP32 rounding could reorder a near tie; here the diagonal is well spread
    assert np.array_equal(piv_gpu[:8], piv[:8])
    if np.array_equal(piv_gpu, piv):
        assert np.allclose(P.F.cpu().numpy(), Lr, atol=2e-4 * np.abs(Lr).max())
    # low-rank exactness on a rank-12 matrix
    V = rs.normal(size=(n, 12)).astype(np.float32)
    A2 = torch.from_numpy(V @ V.T).to(dev)
    P2 = L.PivotedCholeskyPreconditioner(A2, 12, 0.1)
    F2 = P2.F.cpu().double()
    K2 = torch.from_numpy(V).double() @ torch.from_numpy(V).double().T
    assert (F2 @ F2.T - K2).abs().max() < 2e-3 * K2.abs().max()
    assert P2.residual_trace < 1e-3 * P2.trace_K


def test_rbf_dense_matches_oracle(dev):
    g = torch.Generator().manual_seed(8)
    P = torch.rand(513, 4, generator=g)
    K, Kb = L.rbf_matrix(P.to(dev), 0.3, bf16=True)
    Kc = K.cpu()
    assert torch.equal(Kc, Kc.T)
    ref = torch.from_numpy(ol.rbf_dense(P.numpy(), 0.3))
    assert torch.allclose(Kc.double(), ref, rtol=2e-6, atol=1e-7)
    assert torch.equal(Kb.cpu(), Kc.to(torch.bfloat16))


@pytest.mark.parametrize("n,B,ld", [(1001, 11, 1024), (1024, 11, 1024), (1028, 8, 1032), (2048, 3, 2048),
                                    (516, 1, 520)])
def test_dense_f64acc(dev, n, B, ld):
    # n % 4 != 0 takes the row-per-lane kernel; the rest the LDS-staged one (chunks of 512
    # columns, ragged last chunk and last block of rows included)
    g = torch.Generator().manual_seed(9 + n + B)
    A = torch.randn(n, n, generator=g)
    X = torch.randn(B, ld, generator=g, dtype=torch.float64)
    Y = torch.zeros(B, ld, dtype=torch.float64, device=dev)
    Ad, Xd = A.to(dev), X.to(dev)
    N.call("ck_dense_apply_f64acc", n, N.ptr(Ad), n, 0.5, 0.25, B, N.ptr(Xd), ld, N.ptr(Y), ld,
           N.stream_ptr(dev))
    ref = 0.5 * (A.double() @ X[:, :n].T).T + 0.25 * X[:, :n]
    assert torch.allclose(Y[:, :n].cpu(), ref, rtol=1e-12, atol=1e-11)


def test_woodbury_accurate_when_ill_conditioned(dev):
    """lambda(F F^T) / s up to 1e7: the FP64 application keeps the result to fp32 rounding."""
    g = torch.Generator().manual_seed(10)
    n, k, s = 4000, 48, 1e-3
    Q, _ = torch.linalg.qr(torch.randn(n, k, generator=g, dtype=torch.float64))
    lam = torch.logspace(4, -2, k, dtype=torch.float64)
    F = (Q * lam.sqrt()).float()
    X = torch.randn(n, 3, generator=g)
    out = ca.WoodburyInverse(F.to(dev), s) @ X.to(dev)
    Fd = F.double()
    ref = torch.linalg.solve(s * torch.eye(n, dtype=torch.float64) + Fd @ Fd.T, X.double())
    err = ((out.cpu().double() - ref).norm(dim=0) / ref.norm(dim=0)).max()
    assert err < 1e-6, err


def test_training_set_batched_matches_single(dev):
    """The batched path (one Lanczos over several instances, one eigensolve launch) gives the
    targets of one instance at a time, up to the fp32 summation order of the two matvec kernels."""
    batch = L.make_training_set([3, 4, 5], 300, m=16, steps=60, device=dev, batch=2)
    for seed, inst in zip([3, 4, 5], batch):
        one = L.make_training_instance(seed, 300, m=16, steps=60, device=dev)
        assert torch.allclose(inst.w, one.w, rtol=1e-4, atol=1e-6 * float(one.w.max()))
        # leading Ritz vectors up to sign
        dots = (inst.U[:8, :300] * one.U[:8, :300]).sum(1).abs()
        assert torch.allclose(dots, torch.ones_like(dots), atol=1e-3)


def test_dense_batched_operator(dev):
    """CK_OP_DENSE_BATCHED_F32: vector b multiplied by its own matrix."""
    from cola_amd.ops import _op
    g = torch.Generator().manual_seed(12)
    nb, n = 5, 260
    K = torch.randn(nb, n, n, generator=g)
    X = torch.randn(nb, n, generator=g)
    Kd, Xd = K.to(dev), X.to(dev)
    Y = torch.empty_like(Xd)
    op = _op(N.CK_OP_DENSE_BATCHED_F32, n, scale=0.5, shift=0.25)
    op.a, op.lda, op.nnz = Kd.data_ptr(), n, n * n
    N.call("ck_op_apply", op, nb, N.ptr(Xd), n, N.ptr(Y), n, None, 0, N.stream_ptr(dev))
    ref = 0.5 * torch.einsum("bij,bj->bi", K.double(), X.double()) + 0.25 * X.double()
    assert torch.allclose(Y.cpu().double(), ref, rtol=1e-5, atol=1e-4)


def _iters(A, b, P=None):
    _, info = cg(A, b, P=P, tol=1e-6, max_iters=600)
    return info["iterations"]


def test_learned_preconditioner_ablation(dev):
    """Training cuts the captured-energy loss, and the trained network's preconditioner
    needs fewer CG steps than the untrained network's and than plain CG."""
    n_tr, n_ev, m = 1024, 2048, 64
    insts = [L.make_training_instance(100 + i, n_tr, m=m, ell_range=(0.2, 1.0), device=dev) for i in range(24)]
    net = L.PointMLP(4, 64, device=dev, seed=0)
    before = L.mean_loss(net, insts)
    losses = L.train(net, insts, epochs=8, lr=2e-3, seed=0)
    after = L.mean_loss(net, insts)
    print(f"energy loss before {before:.4f} after {after:.4f} (last step {losses[-1]:.4f})", flush=True)
    assert np.all(np.isfinite(losses))
    assert after < 0.7 * before
    X, _, _ = L.sample_instance(999, n_ev)
    ell, noise = 0.5, 1e-3
    Xd = X.to(dev)
    K = L.rbf_matrix(Xd, ell)
    A = ca.PSD(ca.Dense(K) + noise * ca.I_like(ca.Dense(K)))
    b = torch.randn(n_ev, 4, generator=torch.Generator().manual_seed(1)).to(dev)
    it_plain = _iters(A, b)
    P0 = L.LearnedPreconditioner(K, Xd, L.PointMLP(4, 64, device=dev, seed=0), noise)
    P1 = L.LearnedPreconditioner(K, Xd, net, noise)
    Pc = L.PivotedCholeskyPreconditioner(K, 64, noise)
    it0, it1, itc = _iters(A, b, P0), _iters(A, b, P1), _iters(A, b, Pc)
    print(f"CG steps: plain {it_plain}, untrained {it0}, trained {it1}, pivoted Cholesky {itc}; "
          f"residual traces untrained {P0.residual_trace:.3f} trained {P1.residual_trace:.3f} "
          f"pivchol {Pc.residual_trace:.3f}", flush=True)
    assert P1.residual_trace < P0.residual_trace
    assert it1 < it0 and it1 < it_plain
    Psel, rep = L.select_preconditioner(K, Xd, noise, net=net, rank=64)
    print("selection", rep, flush=True)
    assert rep["choice"] in ("learned", "pivoted_cholesky")


def test_mixed_dense_bf16_refinement(dev):
    """BF16 inner solves with FP64 residuals reach 1e-8 on a well-conditioned kernel
    system; on an ill-conditioned one the solver falls back to FP32 corrections."""
    g = torch.Generator().manual_seed(3)
    n = 2048
    X = torch.rand(n, 3, generator=g).to(dev)
    b = torch.randn(n, 2, generator=g).to(dev)
    for ell, noise, expect_fallback in [(0.05, 5e-2, False), (0.5, 1e-3, True)]:
        K, Kb = L.rbf_matrix(X, ell, bf16=True)
        A = ca.PSD(ca.Dense(K) + noise * ca.I_like(ca.Dense(K)))
        Ab = ca.DenseBF16(K) + noise * ca.I_like(ca.Dense(K))
        Pc = L.PivotedCholeskyPreconditioner(K, 64, noise)
        alg = HipMixedCG(tol=1e-8, inner_tol=1e-3, inner_max_iters=2000, max_refinements=30, P=Pc, inner=Ab)
        x, info = MixedSolver(A, 2, alg).solve(b)
        Kd = K.double()
        s32 = float(np.float32(noise))  # the operator's shift as stored (ScalarMul keeps c in fp32)
        res = ((Kd @ x + s32 * x) - b.double()).norm(dim=0) / b.double().norm(dim=0)
        print(f"ell {ell} noise {noise}: refinements {info['refinements']}, inner {info['inner_iterations']}, "
              f"fallback at {info['precision_fallback_at']}, residual {res.max().item():.2e}", flush=True)
        assert info["converged"] and res.max().item() < 1.2e-8
        if not expect_fallback:  # bf16 corrections carry at least the first refinement
            assert info["precision_fallback_at"] is None or info["precision_fallback_at"] >= 2
