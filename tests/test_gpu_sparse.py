"""COO -> CSR on the device (ck_coo_to_csr_*, replacing the host construction of
cola/ops/operators.py:64-75) and the Matrix Market fixture 1138_bus.mtx (the reference's
tests/data) loaded, multiplied, solved and its spectrum bounded on the engine."""
import os
import zlib

import numpy as np
import pytest
import scipy.io
import scipy.sparse
import scipy.sparse.linalg
import torch

import cola_amd as ca
from cola_amd.io import load_mtx
from cola_amd.linalg.cg import run_batched_cg
from cola_amd.linalg.mixed import HipMixedCG, MixedSolver
from oracle import cg as ocg
from oracle import operators as oop

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _case(name):
    rs = np.random.RandomState(zlib.crc32(name.encode()))
    if name == "random_pairs":
        # 200k entries, each (row, col) at most twice: scipy's duplicate sums are then order-free
        n, m, base = 5000, 7000, 150000
        r = rs.randint(0, n, base)
        c = rs.randint(0, m, base)
        key = np.unique(r.astype(np.int64) * m + c)
        r, c = key // m, key % m
        dup = rs.choice(key.size, 50000, replace=False)
        rows = np.concatenate([r, r[dup]])
        cols = np.concatenate([c, c[dup]])
        data = rs.normal(size=rows.size).astype(np.float32)
        perm = rs.permutation(rows.size)
        return rows[perm], cols[perm], data[perm], (n, m)
    if name == "row_lengths":
        # every sorting tier: empty rows, one wave (<= 64), LDS (<= 4096), global (> 4096); many
        # duplicates with small-integer values, whose sums are exact in any order
        lens = [0, 1, 2, 63, 64, 65, 0, 700, 4095, 4096, 4097, 20000, 3, 0]
        m = 5000
        rows = np.concatenate([np.full(L, i) for i, L in enumerate(lens)])
        cols = rs.randint(0, m, rows.size)
        cols[rows == 11] = rs.randint(0, 300, (rows == 11).sum())  # the long row: heavy duplication
        data = rs.randint(-8, 9, rows.size).astype(np.float32)
        perm = rs.permutation(rows.size)
        return rows[perm], cols[perm], data[perm], (len(lens), m)
    if name == "one_entry_repeated":
        rows = np.full(10000, 3)
        cols = np.full(10000, 5)
        data = rs.randint(-4, 5, 10000).astype(np.float32)
        return rows, cols, data, (7, 9)
    if name == "f64":
        n = 3000
        rows = rs.randint(0, n, 40000)
        cols = rs.randint(0, n, 40000)
        key = np.unique(rows.astype(np.int64) * n + cols)
        rows, cols = key // n, key % n
        perm = rs.permutation(rows.size)
        return rows[perm], cols[perm], rs.normal(size=rows.size), (n, n)
    if name == "empty":
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32), (5, 5)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["random_pairs", "row_lengths", "one_entry_repeated", "f64", "empty"])
def test_coo_to_csr_matches_oracle(dev, name):
    """Bit-exact indptr, indices and values against the oracle's restatement of the reference's
    construction, and the same arrays on a second run (atomics order entries differently)."""
    rows, cols, data, shape = _case(name)
    print(name, "nnz", data.size, "shape", shape, flush=True)
    ip, ix, v = oop.coo_to_csr(data, rows, cols, shape)
    d = torch.from_numpy(data).to(dev)
    r = torch.from_numpy(rows.astype(np.int64)).to(dev)
    c = torch.from_numpy(cols.astype(np.int64)).to(dev)
    A = ca.Sparse(d, r, c, shape)
    assert A.indptr.dtype == torch.int32 and A.indices.dtype == torch.int32
    assert np.array_equal(A.indptr.cpu().numpy(), ip)
    assert np.array_equal(A.indices.cpu().numpy(), ix)
    assert np.array_equal(A.data.cpu().numpy(), v.astype(data.dtype))
    B = ca.Sparse(d, r, c, shape)
    assert torch.equal(A.indptr, B.indptr) and torch.equal(A.indices, B.indices) and torch.equal(A.data, B.data)


def test_coo_to_csr_rejects_out_of_range(dev):
    d = torch.ones(3, device=dev)
    r = torch.tensor([0, 4, 1], device=dev)
    c = torch.tensor([0, 1, 2], device=dev)
    with pytest.raises(ValueError):
        ca.Sparse(d, r, c, (4, 4))
    with pytest.raises(ValueError):
        ca.Sparse(d, torch.tensor([0, -1, 1], device=dev), c, (4, 4))


@pytest.fixture(scope="module")
def bus(dev):
    path = os.path.join(GOLD, "1138_bus.mtx")
    return load_mtx(path, device=dev), scipy.io.mmread(path).tocsr()


def test_1138_bus_csr_and_matvec(dev, bus):
    """load_mtx on the device: the CSR scipy builds (values rounded to fp32), and the engine's CSR
    product against scipy's FP64 product."""
    A, M = bus
    assert np.array_equal(A.indptr.cpu().numpy(), M.indptr)
    assert np.array_equal(A.indices.cpu().numpy(), M.indices)
    assert np.array_equal(A.data.cpu().numpy(), M.data.astype(np.float32))
    X = torch.randn(M.shape[0], 3, generator=torch.Generator().manual_seed(0))
    Y = (A @ X.to(dev)).double().cpu().numpy()
    M32 = scipy.sparse.csr_matrix((M.data.astype(np.float32).astype(np.float64), M.indices, M.indptr), M.shape)
    ref = M32 @ X.double().numpy()
    assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-6


def test_1138_bus_cg_against_oracle(dev, bus):
    """40 FP32 CG steps on 1138_bus (SPD, condition ~1e7): the engine's residual history against
    the oracle's restatement of run_batched_cg (cg.py:94-178) on the same fp32 CSR matrix.  At
    this conditioning any two FP32 evaluation orders part after a dozen steps, so the history is
    held to the oracle while the oracle's own FP32 and FP64 runs agree, and after that to the
    FP64 run as closely as the FP32 oracle stays to it."""
    A, M = bus
    n = M.shape[0]
    b = torch.randn(n, 2, generator=torch.Generator().manual_seed(5))
    Ad = ca.PSD(A)
    _, _, kg, ig = run_batched_cg(Ad, b.to(dev), None, 40, 0.0, None)
    ip, ix, v = A.indptr.cpu().numpy(), A.indices.cpu().numpy(), A.data.cpu()
    _, _, ko, io = ocg.run_batched_cg(lambda V: oop.csr_matmat(ip, ix, v, V), b, None, 40, 0.0, None)
    _, _, _, i64 = ocg.run_batched_cg(lambda V: oop.csr_matmat(ip, ix, v.double(), V), b.double(), None, 40, 0.0, None)
    ge, oe, e64 = np.log10(ig["errors"]), np.log10(io["errors"]), np.log10(i64["errors"])
    spread = np.abs(oe - e64)
    K = int(np.argmax(spread > 0.005)) if (spread > 0.005).any() else len(oe)
    print("steps where the FP32 oracle tracks FP64:", K, "engine", ig["errors"][::10], "oracle", io["errors"][::10],
          flush=True)
    assert kg == ko == 40 and K >= 8
    assert np.allclose(ge[:K], oe[:K], atol=0.01)
    assert np.abs(ge - e64).max() <= 3 * spread.max() + 0.02


def test_1138_bus_mixed_refinement_and_lanczos(dev, bus):
    """FP64-residual refinement around FP32 CG (residuals through ck_csr_apply_f64acc) solves
    1138_bus to 1e-10 in FP64 -- below what FP32 CG alone reaches at this conditioning -- and
    Lanczos brackets the largest eigenvalue scipy's eigsh finds (FP64)."""
    A, M = bus
    n = M.shape[0]
    b = torch.randn(n, 2, generator=torch.Generator().manual_seed(6))
    Ad = ca.PSD(A)
    x, info = MixedSolver(Ad, 2, HipMixedCG(tol=1e-10, inner_tol=1e-3, inner_max_iters=3000,
                                            max_refinements=40)).solve(b.to(dev))
    M32 = scipy.sparse.csr_matrix((M.data.astype(np.float32).astype(np.float64), M.indices, M.indptr), M.shape)
    res = np.linalg.norm(M32 @ x.double().cpu().numpy() - b.double().numpy(), axis=0) / np.linalg.norm(
        b.double().numpy(), axis=0)
    print("refinements", info["refinements"], "FP64 residual", res, flush=True)
    assert info["converged"] and res.max() <= 1e-10
    lam_max = scipy.sparse.linalg.eigsh(M32, k=1, which="LA", return_eigenvectors=False)[0]
    got = float(ca.linalg.eigmax(Ad, ca.linalg.Lanczos(max_iters=200, tol=1e-6)))
    print("eigmax engine", got, "eigsh", lam_max, flush=True)
    assert abs(got - lam_max) / lam_max < 1e-5
