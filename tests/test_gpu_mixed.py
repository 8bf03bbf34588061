"""Mixed precision and the structured Kronecker preconditioner: FP64 Kronecker
matvec, eigenbasis inverse, refinement kernels, and HipMixedCG reaching an
FP64-checked 1e-6 residual on the 64^3 Kronecker system (closed-form FP64
solution from the factors' eigendecompositions as the oracle)."""
import ctypes

import numpy as np
import pytest
import torch

import cola_amd as ca
import cola_amd._native as N
from cola_amd.linalg import inv
from cola_amd.linalg.mixed import HipMixedCG, KroneckerEigInverse
from oracle import operators as oop

pytestmark = pytest.mark.gpu


def rbf(t, ell):
    return torch.exp(-(t[:, None] - t[None, :])**2 / (2 * ell**2))


def kron_system(dev, nrhs, seed=0):
    g = torch.Generator().manual_seed(seed)
    facs = [rbf(torch.sort(torch.rand(64, generator=g)).values, 0.1) for _ in range(3)]
    b = torch.randn(64**3, nrhs, generator=g)
    K = ca.Kronecker(*[ca.Dense(f.to(dev)) for f in facs])
    return ca.PSD(K + 0.01 * ca.I_like(K)), facs, b


def closed_form_solve(facs, b, shift=0.01):
    """x = (Q1 x Q2 x Q3)(L + s)^-1 (Q1 x Q2 x Q3)^T b in FP64 (oracle)."""
    ws, Qs = zip(*[torch.linalg.eigh(f.double()) for f in facs])
    lam = (ws[0][:, None, None] * ws[1][None, :, None] * ws[2][None, None, :]).reshape(-1) + shift
    y = oop.kron_matmat([Q.T for Q in Qs], b.double())
    return oop.kron_matmat(list(Qs), y / lam[:, None])


def test_kron_f64_matches_oracle(dev):
    g = torch.Generator().manual_seed(3)
    facs = [torch.randn(64, 64, generator=g, dtype=torch.float64) for _ in range(3)]
    X = torch.randn(64**3, 2, generator=g, dtype=torch.float64)
    ref = oop.kron_matmat(facs, X) + 0.5 * X
    Xb = X.T.contiguous().to(dev)
    Yb = torch.empty_like(Xb)
    work = torch.empty(2 * 64**3, dtype=torch.float64, device=dev)
    f = [t.contiguous().to(dev) for t in facs]
    dims = (ctypes.c_int32 * 3)(64, 64, 64)
    N.call("ck_kron_apply_f64", 3, dims, N.ptr(f[0]), N.ptr(f[1]), N.ptr(f[2]), 1.0, 0.5, 2, N.ptr(Xb), 64**3,
           N.ptr(Yb), 64**3, N.ptr(work), work.numel() * 8, N.stream_ptr(dev))
    err = float((Yb.T.cpu() - ref).norm() / ref.norm())
    assert err < 1e-13


def test_kron_eig_inverse(dev):
    A, facs, b = kron_system(dev, 3, seed=1)
    P = KroneckerEigInverse([f.to(dev) for f in facs], 1.0, 0.01)
    x = P @ b.to(dev)
    ref = closed_form_solve(facs, b)
    assert float((x.cpu().double() - ref).norm() / ref.norm()) < 1e-3


def test_ir_residual_kernel(dev):
    n, B = 1000, 3
    b = torch.randn(B, 1024, dtype=torch.float64, device=dev)
    ax = torch.randn(B, 1024, dtype=torch.float64, device=dev)
    r = torch.empty_like(b)
    r32 = torch.empty(B, 1024, device=dev)
    norms = torch.empty(2 * B, dtype=torch.float64, device=dev)
    w = torch.empty(N.q("ck_ir_workspace_bytes", B) // 8, dtype=torch.float64, device=dev)
    N.call("ck_ir_residual_f64", n, B, N.ptr(b), 1024, N.ptr(ax), 1024, N.ptr(r), 1024, N.ptr(r32), 1024,
           N.ptr(norms), N.ptr(w), w.numel() * 8, N.stream_ptr(dev))
    ref = (b - ax)[:, :n]
    assert torch.allclose(r[:, :n], ref)
    assert torch.allclose(norms.view(B, 2)[:, 0], ref.norm(dim=1))
    assert torch.allclose(norms.view(B, 2)[:, 1], b[:, :n].norm(dim=1))


def test_kron_residual_matches_apply_then_residual(dev):
    """The fused FP64 residual (A x never stored) against ck_kron_apply_f64 + ck_ir_residual_f64
    and against the oracle's FP64 Kronecker product."""
    g = torch.Generator().manual_seed(4)
    facs = [torch.randn(64, 64, generator=g, dtype=torch.float64) for _ in range(3)]
    B, n = 3, 64**3
    X = torch.randn(n, B, generator=g, dtype=torch.float64)
    Bv = oop.kron_matmat(facs, X) * 0.7 + 0.25 * X + 1e-3 * torch.randn(n, B, generator=g, dtype=torch.float64)
    f = [t.contiguous().to(dev) for t in facs]
    Xb, Bb = X.T.contiguous().to(dev), Bv.T.contiguous().to(dev)
    s = N.stream_ptr(dev)
    wb = N.q("ck_kron_residual_workspace_bytes", B)
    work = torch.empty(wb // 8, dtype=torch.float64, device=dev)
    r32 = torch.empty(B, n, device=dev)
    norms = torch.empty(2 * B, dtype=torch.float64, device=dev)
    N.call("ck_kron_residual_f64", N.ptr(f[0]), N.ptr(f[1]), N.ptr(f[2]), 0.7, 0.25, B, N.ptr(Xb), n, N.ptr(Bb), n,
           N.ptr(r32), n, N.ptr(norms), N.ptr(work), wb, s)
    # two-pass path
    Ax = torch.empty_like(Xb)
    dims = (ctypes.c_int32 * 3)(64, 64, 64)
    N.call("ck_kron_apply_f64", 3, dims, N.ptr(f[0]), N.ptr(f[1]), N.ptr(f[2]), 0.7, 0.25, B, N.ptr(Xb), n, N.ptr(Ax), n,
           N.ptr(work), wb, s)
    r32b = torch.empty_like(r32)
    norms2 = torch.empty_like(norms)
    w2 = torch.empty(N.q("ck_ir_workspace_bytes", B) // 8, dtype=torch.float64, device=dev)
    N.call("ck_ir_residual_f64", n, B, N.ptr(Bb), n, N.ptr(Ax), n, None, n, N.ptr(r32b), n, N.ptr(norms2), N.ptr(w2),
           w2.numel() * 8, s)
    assert torch.equal(r32, r32b)
    assert torch.allclose(norms, norms2, rtol=1e-12, atol=0)
    ref = Bv - (0.7 * oop.kron_matmat(facs, X) + 0.25 * X)
    assert torch.allclose(norms.view(B, 2)[:, 0].cpu(), ref.norm(dim=0), rtol=1e-9)
    assert torch.allclose(norms.view(B, 2)[:, 1].cpu(), Bv.norm(dim=0), rtol=1e-12)


@pytest.mark.parametrize("B,transposed", [(3, False), (8, False), (16, False), (20, False), (5, True)])
def test_ir_load_rhs(dev, B, transposed):
    """b64 = b, r32 = b, x64 = 0 and both norms = ||b|| from an (n, B) view of any strides;
    more than 16 columns take two passes."""
    n, ld = 1000, 1024
    g = torch.Generator().manual_seed(B)
    base = torch.randn(B, n, generator=g) if transposed else torch.randn(n, B, generator=g)
    b = (base.T if transposed else base).to(dev)
    b64 = torch.full((B, ld), 3.0, dtype=torch.float64, device=dev)
    x64 = torch.full((B, ld), 3.0, dtype=torch.float64, device=dev)
    r32 = torch.full((B, ld), 3.0, device=dev)
    norms = torch.empty(2 * B, dtype=torch.float64, device=dev)
    w = torch.empty(N.q("ck_ir_workspace_bytes", B) // 8, dtype=torch.float64, device=dev)
    N.call("ck_ir_load_rhs", n, B, N.ptr(b), b.stride(0), b.stride(1), N.ptr(b64), N.ptr(r32), N.ptr(x64), ld,
           N.ptr(norms), N.ptr(w), w.numel() * 8, N.stream_ptr(dev))
    bt = b.T.cpu()
    assert torch.equal(b64[:, :n].cpu(), bt.double()) and torch.equal(r32[:, :n].cpu(), bt)
    assert torch.all(x64[:, :n] == 0) and torch.all(x64[:, n:] == 3.0)
    nb = bt.double().norm(dim=1)
    assert torch.allclose(norms.view(B, 2)[:, 0].cpu(), nb, rtol=1e-13)
    assert torch.allclose(norms.view(B, 2)[:, 1].cpu(), nb, rtol=1e-13)
    # lean form: no x64 and no norms (nothing else written)
    b64.fill_(3.0), x64.fill_(3.0), r32.fill_(3.0), norms.fill_(-1.0)
    N.call("ck_ir_load_rhs", n, B, N.ptr(b), b.stride(0), b.stride(1), N.ptr(b64), N.ptr(r32), None, ld, None, None, 0,
           N.stream_ptr(dev))
    assert torch.equal(b64[:, :n].cpu(), bt.double()) and torch.equal(r32[:, :n].cpu(), bt)
    assert torch.all(x64 == 3.0) and torch.all(norms == -1.0)


def test_ir_inner_start_and_scaled_update(dev):
    """The inner solve's normalisation folded into the refinement: r = r32 / ||r|| (fp32
    division, a zero column stays zero), x = 0, mult = float(||r||); then x64 += d * mult
    with the product rounded in fp32 and the sum in fp64."""
    n, B, ld = 1000, 3, 1024
    g = torch.Generator().manual_seed(5)
    r32 = torch.randn(B, ld, generator=g)
    r32[2] = 0.0
    norms = torch.zeros(2 * B, dtype=torch.float64)
    norms[0::2] = r32[:, :n].double().norm(dim=1)
    r32d, normsd = r32.to(dev), norms.to(dev)
    r = torch.full((B, ld), 7.0, device=dev)
    x = torch.full((B, ld), 7.0, device=dev)
    mult = torch.full((B, ), -1.0, device=dev)
    N.call("ck_ir_inner_start", n, B, N.ptr(r32d), ld, N.ptr(normsd), N.ptr(r), N.ptr(x), ld, N.ptr(mult),
           N.stream_ptr(dev))
    m32 = norms[0::2].float()
    assert torch.equal(mult.cpu(), m32)
    safe = torch.where(m32.abs() < 1e-40, torch.full_like(m32, 1e-40), m32)
    assert torch.equal(r[:, :n].cpu(), r32[:, :n] / safe[:, None])
    assert torch.all(x[:, :n] == 0) and torch.all(r[:, n:] == 7.0)
    x64 = torch.randn(B, ld, generator=g, dtype=torch.float64)
    d = torch.randn(B, ld, generator=g)
    xd = x64.to(dev)
    N.call("ck_ir_update_scaled_f64", n, B, N.ptr(xd), ld, N.ptr(d.to(dev)), ld, N.ptr(mult), N.stream_ptr(dev))
    ref = x64[:, :n] + (d[:, :n] * m32[:, None]).double()
    assert torch.equal(xd[:, :n].cpu(), ref)
    assert torch.equal(xd[:, n:].cpu(), x64[:, n:])


def test_mixed_cg_kron_reaches_fp64_tolerance(dev):
    A, facs, b = kron_system(dev, 4, seed=0)
    P = KroneckerEigInverse([f.to(dev) for f in facs], 1.0, 0.01)
    solver = inv(A, HipMixedCG(tol=1e-6, inner_tol=1e-4, P=P))
    x = solver @ b.to(dev)
    info = solver.info
    print("refinements", info["refinements"], "inner", info["inner_iterations"], "res", info["rel_residual_f64"],
          flush=True)
    assert info["converged"] and float(info["rel_residual_f64"].max()) <= 1e-6
    ref = closed_form_solve(facs, b)
    assert x.dtype == torch.float64
    # independent FP64 residual of the returned solution (oracle matvec on the CPU)
    Ax = oop.kron_matmat([f.double() for f in facs], x.cpu()) + 0.01 * x.cpu()
    res = (Ax - b.double()).norm(dim=0) / b.double().norm(dim=0)
    print("returned x residual", res, "x err", float((x.cpu() - ref).norm() / ref.norm()), flush=True)
    assert float(res.max()) <= 1e-6
    assert float((x.cpu() - ref).norm() / ref.norm()) < 1e-4


def test_mixed_cg_graph_replays_match_eager(dev):
    """A repeated solve replays the previous schedule; captured as a HIP graph it must give
    the eager replay's bits on every replay, with eager work and syncs in between, and on a
    new right-hand side (the graph reads the solver's own rhs buffer)."""
    from cola_amd.linalg.mixed import MixedSolver
    A, facs, b = kron_system(dev, 4, seed=2)
    P = KroneckerEigInverse([f.to(dev) for f in facs], 1.0, 0.01)
    bd = b.to(dev)
    eager = MixedSolver(A, 4, HipMixedCG(tol=1e-6, inner_tol=1e-4, P=P, graph=False))
    graphed = MixedSolver(A, 4, HipMixedCG(tol=1e-6, inner_tol=1e-4, P=P, graph=True))
    x_first, _ = eager.solve(bd)  # records the schedule (host-checked loop, full CG steps)
    x_first = x_first.clone()
    graphed.solve(bd)
    x_ref, info_ref = eager.solve(bd)
    x_ref = x_ref.clone()
    # one-step corrections replay as ck_pcg_step1_f64 (no normalisation): equal to fp32 rounding
    assert info_ref["converged"]
    assert float((x_ref - x_first).norm() / x_first.norm()) < 1e-5
    for rep in range(3):
        junk = torch.randn(1 << 20, device=dev).sum()  # eager work between replays
        torch.cuda.synchronize()
        x, info = graphed.solve(bd)
        assert info["converged"] and info["refinements"] == info_ref["refinements"], (rep, info)
        assert torch.equal(x, x_ref), rep
        del junk
    b2 = torch.randn(64**3, 4, generator=torch.Generator().manual_seed(9)).to(dev)
    x2_ref, _ = eager.solve(b2)
    x2, info2 = graphed.solve(b2)
    assert info2["converged"]
    assert torch.equal(x2, x2_ref)


def test_pcg_step1_matches_one_cg_step(dev):
    """ck_pcg_step1_f64 against the generic path (normalise, start, one final CG step into
    the FP64 accumulator) on the 64^3 Kronecker system with its eigenbasis preconditioner;
    a zero column stays zero."""
    from cola_amd.linalg.cg import CGState, cg_blocks
    from cola_amd.linalg._engine import lower, workspace
    A, facs, b = kron_system(dev, 3, seed=5)
    P = KroneckerEigInverse([f.to(dev) for f in facs], 1.0, 0.01)
    n, B = 64**3, 3
    r = b.T.contiguous().to(dev) * torch.tensor([1.0, 1e-3, 0.0], device=dev)[:, None]
    aop, pop = lower(A), lower(P)
    work, wb = workspace([aop[0], pop[0]], B, dev)
    st = CGState(n, B, 10, dev, True)
    x1 = torch.zeros(B, n, dtype=torch.float64, device=dev)
    N.call("ck_pcg_step1_f64", aop[0], pop[0], B, N.ptr(r), N.ptr(st.z), N.ptr(st.Ap), n, N.ptr(x1), n,
           N.ptr(st.partials), N.ptr(work), wb, N.stream_ptr(dev))
    x2 = torch.zeros(B, n, dtype=torch.float64, device=dev)
    st2 = CGState(n, B, 10, dev, True)
    cg_blocks(A, r, st2, None, 1, 1e-3, P, 1, (aop, pop, work, wb), want_info=False, read=False, finish=False,
              final=True, x64=x2)
    # (the generic final step already adds mult * x to x2: both are un-normalised)
    for c in range(2):
        assert float((x1[c] - x2[c]).norm() / x2[c].norm()) < 1e-5, c
    assert torch.all(x1[2] == 0)
    # x1 = alpha z with alpha = r.z / z.Az (the one-step minimiser), from the kernel's own z and A z
    z, az, rd = st.z[:, :n].double(), st.Ap[:, :n].double(), r.double()
    for c in range(2):
        alpha = float((rd[c] * z[c]).sum() / (z[c] * az[c]).sum())
        assert float((x1[c] - alpha * z[c]).norm() / x1[c].norm()) < 1e-6, c


def test_ir_kron_step1_matches_two_calls(dev):
    """ck_ir_kron_step1_f64 (x update inside the FP64 residual's first pass) against
    ck_pcg_step1_f64 followed by ck_kron_residual_f64: the same bits for x64, r32 and the
    norms, for a first (x64 = alpha z) and an accumulating step."""
    from cola_amd.linalg.cg import CGState
    from cola_amd.linalg._engine import lower, workspace
    A, facs, b = kron_system(dev, 2, seed=7)
    P = KroneckerEigInverse([f.to(dev) for f in facs], 1.0, 0.01)
    n, B = 64**3, 2
    f64 = [f.double().contiguous().to(dev) for f in facs]
    s32 = float(np.float32(0.01))
    aop, pop = lower(A), lower(P)
    work, wb = workspace([aop[0], pop[0]], B, dev)
    w64b = N.q("ck_kron_residual_workspace_bytes", B)
    s = N.stream_ptr(dev)
    b64 = b.T.contiguous().double().to(dev)
    outs = []
    for fused in (True, False):
        st = CGState(n, B, 10, dev, True)
        w64 = torch.empty(w64b // 8, dtype=torch.float64, device=dev)
        r32 = b.T.contiguous().to(dev)
        x64 = torch.full((B, n), 5.0, dtype=torch.float64, device=dev)
        norms = torch.empty(2 * B, dtype=torch.float64, device=dev)
        for step in range(2):
            if fused:
                N.call("ck_ir_kron_step1_f64", aop[0], pop[0], N.ptr(f64[0]), N.ptr(f64[1]), N.ptr(f64[2]), 1.0, s32,
                       B, N.ptr(r32), N.ptr(st.z), N.ptr(st.Ap), n, N.ptr(x64), N.ptr(b64), step, N.ptr(norms),
                       N.ptr(st.partials), N.ptr(work), wb, N.ptr(w64), w64b, s)
            else:
                if step == 0:
                    x64.zero_()
                N.call("ck_pcg_step1_f64", aop[0], pop[0], B, N.ptr(r32), N.ptr(st.z), N.ptr(st.Ap), n, N.ptr(x64), n,
                       N.ptr(st.partials), N.ptr(work), wb, s)
                N.call("ck_kron_residual_f64", N.ptr(f64[0]), N.ptr(f64[1]), N.ptr(f64[2]), 1.0, s32, B, N.ptr(x64),
                       n, N.ptr(b64), n, N.ptr(r32), n, N.ptr(norms), N.ptr(w64), w64b, s)
        outs.append((x64.clone(), r32.clone(), norms.clone()))
    for u, v in zip(outs[0], outs[1]):
        assert torch.equal(u, v)
    rel = outs[0][2].view(B, 2)
    assert float((rel[:, 0] / rel[:, 1]).max()) < 1e-6


def test_solve_concurrent_matches_sequential(dev):
    """Independent refinement solves on two streams at once (solve_concurrent): every system's
    FP64 iterate and refinement history equal, bit for bit, those of the same solvers run one
    after another; three dense RBF systems with Woodbury preconditioners (the fused PCG step)
    and one without, 8 rhs each, graphs replayed."""
    from cola_amd.linalg.mixed import MixedSolver, solve_concurrent
    solvers, rhs = [], []
    for j, (ell, noise, rank) in enumerate([(0.4, 0.05, 32), (0.2, 0.1, 16), (0.7, 0.02, 32), (0.3, 0.05, 0)]):
        g = torch.Generator().manual_seed(40 + j)
        X = torch.rand(1536, 3, generator=g)
        K = torch.exp(-torch.cdist(X, X, compute_mode="donot_use_mm_for_euclid_dist")**2 / (2 * ell * ell))
        Kd = K.to(dev)
        A = ca.PSD(ca.Dense(Kd) + noise * ca.I_like(ca.Dense(Kd)))
        P = None
        if rank:
            G = torch.randn(1536, rank, generator=g, dtype=torch.float64)
            Y = K.double() @ G
            L = torch.linalg.cholesky(G.T @ Y + 1e-10 * torch.eye(rank, dtype=torch.float64))
            F = torch.linalg.solve_triangular(L, Y.T, upper=False).T.float()
            P = ca.WoodburyInverse(F.to(dev), noise)
        solvers.append(MixedSolver(A, 8, HipMixedCG(tol=1e-6, inner_tol=1e-3, P=P, max_refinements=20)))
        rhs.append(torch.randn(1536, 8, generator=g).to(dev))
    for sv, b in zip(solvers, rhs):
        sv.solve(b)  # records each schedule
    seq = [(x.clone(), info) for x, info in (sv.solve(b) for sv, b in zip(solvers, rhs))]
    for rep in range(2):
        con = solve_concurrent(solvers, rhs, streams=2)
        torch.cuda.synchronize()
        for j, ((xs, is_), (xc, ic)) in enumerate(zip(seq, con)):
            assert is_["converged"] and ic["converged"], (rep, j)
            assert ic["inner_iterations"] == is_["inner_iterations"], (rep, j)
            assert torch.equal(xc, xs), (rep, j)
