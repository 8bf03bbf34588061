"""Host logic that needs no GPU: the dispatch registry, annotations, operator
algebra and lowering decisions, and the engine's refusal of CPU tensors."""
import numpy as np
import pytest
import torch

import cola_amd as ca
from cola_amd.dispatch import dispatch
from cola_amd.linalg import Auto, HipCG, HipHutch, HipLanczos, inv
from cola_amd.linalg.algorithm_base import Algorithm, IterativeOperatorWInfo


class _Base:
    pass


class _Child(_Base):
    pass


@dispatch.abstract
def _route(a: _Base, b: int = 0):
    pass


@dispatch(precedence=-1)
def _route(a: _Base, b: int):
    return "base"


@dispatch
def _route(a: _Child, b: int):
    return "child"


@dispatch(cond=lambda a, b: b > 10)
def _route(a: _Base, b: int):
    return "cond"


def test_dispatch_specificity_precedence_cond():
    assert _route(_Base(), 1) == "base"
    assert _route(_Child(), 1) == "child"
    assert _route(_Base(), 20) == "cond"  # equal signatures: precedence 0 beats -1
    assert _route(_Base()) == "base"  # default argument from the abstract signature
    with pytest.raises(NotImplementedError):
        _route("x", 1)


def test_annotations_inference():
    A = ca.PSD(ca.Dense(torch.eye(3)))
    assert A.isa(ca.PSD) and A.isa(ca.SelfAdjoint)
    B = ca.Dense(torch.eye(3))
    assert not B.isa(ca.PSD)
    K = ca.Kronecker(A, A)
    assert K.isa(ca.PSD)
    S = K + 0.1 * ca.I_like(K)
    assert S.isa(ca.PSD) and not S.isa(ca.Unitary)
    assert ca.I_like(B).isa(ca.Unitary)


def test_operator_algebra_cpu():
    g = torch.Generator().manual_seed(0)
    A = torch.randn(6, 6, generator=g)
    B = torch.randn(6, 6, generator=g)
    X = torch.randn(6, 2, generator=g)
    opA, opB = ca.Dense(A), ca.Dense(B)
    assert torch.allclose((opA + opB) @ X, (A + B) @ X, atol=1e-6)
    assert torch.allclose((opA @ opB) @ X, A @ B @ X, atol=1e-5)
    assert torch.allclose((2.0 * opA) @ X, 2 * A @ X, atol=1e-6)
    assert torch.allclose((opA - opB) @ X, (A - B) @ X, atol=1e-6)
    K = ca.kron(ca.Dense(A[:2, :2]), ca.Dense(B[:3, :3]))
    assert torch.allclose(K.to_dense(), torch.kron(A[:2, :2], B[:3, :3]), atol=1e-6)
    x1 = torch.randn(6, generator=g)
    assert (opA @ x1).shape == (6, )


def test_sparse_cpu_and_validation():
    data = torch.tensor([4., 5., 6., 1., 2., 3.])
    rows = torch.tensor([2, 2, 2, 0, 0, 1])
    cols = torch.tensor([0, 1, 2, 1, 3, 3])
    A = ca.Sparse(data, rows, cols, (3, 4))
    assert torch.allclose(A @ torch.tensor([1., 0., -1., 1.]), torch.tensor([2., 3., -2.]))
    assert A.indptr.tolist() == [0, 2, 3, 6] and A.indices.dtype == torch.int32
    with pytest.raises(ValueError):
        ca.Sparse(data, rows, cols + 10, (3, 4))
    with pytest.raises(ValueError):
        ca.Sparse.from_csr(torch.tensor([0, 2, 1, 6]), A.indices, A.data, (3, 4))


def test_lowering_needs_device_tensors():
    A = ca.Dense(torch.eye(4))
    assert A.hip_op() is None  # CPU storage never lowers to a kernel
    assert ca.Sparse(torch.ones(2), torch.tensor([0, 1]), torch.tensor([0, 1]), (2, 2)).hip_op() is None


def test_inv_routing_and_cpu_refusal():
    A = ca.PSD(ca.Dense(torch.eye(4) * 2))
    Ainv = inv(A, HipCG())
    assert isinstance(Ainv, IterativeOperatorWInfo)
    # Auto (inv.py:73-92): N^2 <= 1e6 -> dense Cholesky (PSD) / LU, larger -> CG / GMRES
    big = ca.PSD(ca.Dense(torch.eye(1001) * 2))
    assert isinstance(inv(big, Auto(tol=1e-3)).alg, HipCG) and inv(big, Auto(tol=1e-3)).alg.tol == 1e-3
    assert torch.allclose(inv(A, Auto()) @ torch.ones(4), 0.5 * torch.ones(4))
    with pytest.raises(ValueError, match="HIP"):
        Ainv @ torch.ones(4)
    with pytest.raises(AssertionError):
        inv(ca.Dense(torch.eye(4)), HipCG())
    # structural rules win over the algorithm base case (inv.py:139-146)
    D = ca.Diagonal(torch.tensor([2., 4.]))
    assert torch.allclose(inv(D, HipCG()).diag, torch.tensor([0.5, 0.25]))
    K = ca.Kronecker(ca.Diagonal(torch.tensor([2., 4.])), ca.Diagonal(torch.tensor([1., 2.])))
    assert isinstance(inv(K, HipCG()), ca.Kronecker)


def test_algorithm_fields_match_reference_defaults():
    cg = HipCG()
    assert (cg.tol, cg.max_iters, cg.pbar, cg.x0, cg.P) == (1e-6, 1000, False, None, None)
    lz = HipLanczos()
    assert (lz.start_vector, lz.max_iters, lz.tol, lz.key) == (None, 1000, 1e-6, None)
    h = HipHutch()
    assert (h.tol, h.max_iters, h.bs, h.rand) == (3e-2, 10_000, 100, "normal")
    assert issubclass(HipCG, Algorithm) and ca.linalg.CG is HipCG and ca.linalg.Lanczos is HipLanczos


def test_keys_match_reference_chain():
    from cola_amd.linalg.keys import PRNGKey, next_key
    assert PRNGKey(42) == 3300923923 and next_key(3300923923) == 1419105812


def test_flatten_unflatten_round_trip():
    """operator_base.py:89-95: flatten lists every tensor of a composite operator; unflatten builds
    the same structure around new tensors (here on the host, where operators use torch arithmetic)."""
    g = torch.Generator().manual_seed(0)
    M = torch.randn(5, 5, generator=g)
    c = torch.tensor(0.3)
    d = torch.rand(5, generator=g) + 1.0
    A = ca.Dense(M) + ca.ScalarMul(c, (5, 5)) + ca.Diagonal(d)
    leaves, unflatten = A.flatten()
    assert any(t is M for t in leaves) and any(t is c for t in leaves) and any(t is d for t in leaves)
    X = torch.randn(5, 2, generator=g)
    B = unflatten([2 * t if t.is_floating_point() else t for t in leaves])
    assert type(B) is type(A)
    assert torch.allclose(B @ X, 2 * (A @ X), atol=1e-5)
    # the original is untouched
    assert torch.allclose(A @ X, M @ X + 0.3 * X + d[:, None] * X, atol=1e-5)


def test_vjp_derivs_of_operator_products():
    """d/dtheta <A(theta) X, D> for A = Dense(M) + c I: D X^T for M and <X, D> for c."""
    from cola_amd.autodiff import vjp_derivs
    g = torch.Generator().manual_seed(1)
    M = torch.randn(4, 4, generator=g)
    c = torch.tensor(0.7)
    A = ca.Dense(M) + ca.ScalarMul(c, (4, 4))
    leaves, unflatten = A.flatten()
    X = torch.randn(4, 3, generator=g)
    D = torch.randn(4, 3, generator=g)
    grads = vjp_derivs(lambda *th: unflatten(th) @ X, leaves, D)
    byid = {id(t): gr for t, gr in zip(leaves, grads)}
    assert torch.allclose(byid[id(M)], D @ X.T, atol=1e-5)
    assert torch.allclose(byid[id(c)], (X * D).sum(), atol=1e-5)


def test_stationary_kernel_fns_closed_forms():
    """The kernel callables of Kernel(x1, x2, fn): value 1 at r = 0 and the closed forms at r = l."""
    import math
    x = torch.tensor([[0.0, 0.0], [0.3, 0.4]], dtype=torch.float64)  # |x1 - x0| = 0.5
    ell = 0.5
    want = {0: math.exp(-0.5), 1: math.exp(-1.0), 2: (1 + math.sqrt(3)) * math.exp(-math.sqrt(3)),
            3: (1 + math.sqrt(5) + 5.0 / 3.0) * math.exp(-math.sqrt(5))}
    for fam, fn in ((0, ca.rbf(ell)), (1, ca.matern12(ell)), (2, ca.matern32(ell)), (3, ca.matern52(ell))):
        K = fn(x, x)
        assert abs(float(K[0, 0]) - 1.0) < 1e-12 and abs(float(K[0, 1]) - want[fam]) < 1e-12
    with pytest.raises(ValueError):
        ca.StationaryKernelFn(7, 1.0)


def test_kernel_blocks_match_dense_on_cpu():
    """Kernel's block loop (operators.py:631-643) on the CPU equals fn on the whole point set; a CPU
    operator is not lowered to the engine."""
    x = torch.rand(50, 3, generator=torch.Generator().manual_seed(0))
    v = torch.randn(50, 2, generator=torch.Generator().manual_seed(1))
    for fn in (ca.matern32(0.3),
               lambda a, b: torch.exp(-torch.cdist(a, b, compute_mode="donot_use_mm_for_euclid_dist"))):
        K = ca.Kernel(x, x, fn, 16, 16)
        assert K.hip_op() is None
        assert torch.allclose(K @ v, fn(x, x) @ v, rtol=1e-5, atol=1e-4)


def test_structural_rules_triangular_permutation_kronsum_blockdiag():
    """inv/slogdet/diag rules of inv.py:123-165, logdet.py:161-184 and diag_trace.py:87-118 on the
    CPU against dense torch results."""
    from cola_amd.linalg import diag, slogdet
    g = torch.Generator().manual_seed(0)
    L = torch.tril(torch.randn(6, 6, generator=g, dtype=torch.float64)) + 4 * torch.eye(6, dtype=torch.float64)
    T = ca.Triangular(L, lower=True)
    Ti = inv(T)
    v = torch.randn(6, 2, generator=g, dtype=torch.float64)
    assert torch.allclose(L @ (Ti @ v), v)
    U = ca.Triangular(L.T.contiguous(), lower=False)
    assert torch.allclose(L.T @ (inv(U) @ v), v)
    s, ld = slogdet(T)
    s0, ld0 = torch.linalg.slogdet(L)
    assert torch.allclose(ld, ld0) and float(s) == float(s0)
    perm = torch.tensor([2, 0, 3, 1])
    P = ca.Permutation(perm, torch.float64)
    w = torch.randn(4, 3, generator=g, dtype=torch.float64)
    assert torch.equal(P @ w, w[perm]) and torch.equal(inv(P) @ (P @ w), w)
    assert float(slogdet(P)[1]) == 0.0
    A = torch.randn(3, 3, generator=g, dtype=torch.float64)
    B = torch.randn(4, 4, generator=g, dtype=torch.float64)
    KS = ca.KronSum(ca.Dense(A), ca.Dense(B))
    dense = torch.kron(A, torch.eye(4, dtype=torch.float64)) + torch.kron(torch.eye(3, dtype=torch.float64), B)
    x = torch.randn(12, 2, generator=g, dtype=torch.float64)
    assert torch.allclose(KS @ x, dense @ x) and torch.allclose(KS.to_dense(), dense)
    assert torch.allclose(diag(KS), torch.diagonal(dense))
    M1 = torch.randn(2, 2, generator=g, dtype=torch.float64) + 3 * torch.eye(2, dtype=torch.float64)
    M2 = torch.randn(3, 3, generator=g, dtype=torch.float64) + 3 * torch.eye(3, dtype=torch.float64)
    BD = ca.BlockDiag(ca.Dense(M1), ca.Dense(M2), multiplicities=[2, 1])
    D = torch.block_diag(M1, M1, M2)
    assert torch.allclose(diag(BD), torch.diagonal(D))
    d1 = torch.tensor([2.0, -3.0], dtype=torch.float64)
    d2 = torch.tensor([0.5, 4.0, 1.5], dtype=torch.float64)
    BDd = ca.BlockDiag(ca.Diagonal(d1), ca.Diagonal(d2), multiplicities=[2, 1])
    sb, lb = slogdet(BDd)
    s_ref, l_ref = torch.linalg.slogdet(torch.diag(torch.cat([d1, d1, d2])))
    assert abs(float(lb) - float(l_ref)) < 1e-12 and float(sb) == float(s_ref)


def test_auto_small_operators_use_dense_factorizations():
    """inv.py:73-104 and logdet.py:82-108: below 1e6 entries Auto factorizes densely, Cholesky for
    PSD operators and P L U otherwise; solves and log-determinants against torch in FP64, the
    structured factors, and the reference's (1, 0) for the permutation (its parity is not in the
    sign: an LU whose pivoting swaps an odd number of rows reports the opposite sign)."""
    from cola_amd.linalg import LU, Cholesky, cholesky, plu, slogdet
    g = torch.Generator().manual_seed(3)
    M = torch.randn(60, 60, generator=g, dtype=torch.float64)
    S = M @ M.T + 60 * torch.eye(60, dtype=torch.float64)
    b = torch.randn(60, 2, generator=g, dtype=torch.float64)
    A = ca.PSD(ca.Dense(S))
    L = cholesky(A)
    assert isinstance(L, ca.Triangular) and L.lower and torch.allclose(L.A @ L.A.T, S)
    assert torch.allclose(inv(A) @ b, torch.linalg.solve(S, b))
    assert torch.allclose(inv(A, Cholesky()) @ b, torch.linalg.solve(S, b))
    s, ld = slogdet(A)
    s0, ld0 = torch.linalg.slogdet(S)
    assert float(s) == float(s0) and abs(float(ld) - float(ld0)) < 1e-10
    G = ca.Dense(M)
    P, Lg, U = plu(G)
    assert isinstance(P, ca.Permutation) and isinstance(Lg, ca.Triangular) and not U.lower
    assert torch.allclose((P @ Lg @ U).to_dense(), M)
    assert torch.allclose(inv(G) @ b, torch.linalg.solve(M, b))
    assert torch.allclose(inv(G, LU()) @ b, torch.linalg.solve(M, b))
    s, ld = slogdet(G)
    s0, ld0 = torch.linalg.slogdet(M)
    assert abs(float(ld) - float(ld0)) < 1e-10
    Pm = P.to_dense()
    parity = float(torch.linalg.det(Pm))
    assert float(s) == float(s0) * parity
    # structured factors: Diagonal, Kronecker
    D = ca.PSD(ca.Diagonal(torch.tensor([4., 9.], dtype=torch.float64)))
    assert torch.allclose(cholesky(D).diag, torch.tensor([2., 3.], dtype=torch.float64))
    K = ca.Kronecker(ca.PSD(ca.Dense(S[:5, :5])), ca.PSD(ca.Dense(S[5:9, 5:9])))
    CK = cholesky(K)
    assert isinstance(CK, ca.Kronecker)
    assert torch.allclose((CK @ CK.to_dense().T), K.to_dense())


def test_largest_cluster_rule():
    # the bisection solver's cluster rule: neighbours closer than 1e-3 max(|first|, |last|)
    from cola_amd.linalg.lanczos import largest_cluster
    assert largest_cluster(np.array([[0.0, 1.0, 2.0, 3.0]])) == 1
    ev = np.linspace(0.0, 1000.0, 53)  # gaps 19.2 > 1e-3 * 1000
    assert largest_cluster(ev[None, :]) == 1
    near = np.concatenate([np.linspace(0.0, 0.5, 40), [10.0, 100.0, 1000.0]])  # gaps 0.0128 < 1.0
    assert largest_cluster(near[None, :]) == 40
    tail = np.concatenate([1e-6 * np.arange(50), [3.0, 700.0, 1000.0]])  # 50 values within 1e-3 * 1000
    assert largest_cluster(tail[None, :]) == 50
    assert largest_cluster(np.stack([ev, tail])) == 50
