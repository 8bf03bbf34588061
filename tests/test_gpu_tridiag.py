"""Device FP64 tridiagonal eigensolvers against LAPACK (scipy): implicit QL (ck_tridiag_eigh_f64) and
Sturm bisection + inverse iteration (ck_tridiag_eigh_bisect_f64)."""
import numpy as np
import pytest
import torch

import cola_amd._native as N
from oracle.tridiag import eigh_tridiag

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,nb", [(1, 1), (2, 3), (10, 4), (30, 32), (120, 1), (200, 2)])
def test_tridiag_eigh(dev, m, nb):
    rs = np.random.RandomState(m + nb)
    d = rs.normal(size=(nb, m)).astype(np.float32)
    e = rs.uniform(0.1, 1.0, size=(nb, max(m - 1, 1))).astype(np.float32)
    dd, ee = torch.from_numpy(d).to(dev), torch.from_numpy(e).to(dev)
    ev = torch.empty(nb, m, dtype=torch.float64, device=dev)
    vecs = torch.empty(nb, m, m, dtype=torch.float64, device=dev)
    fr = torch.empty(nb, m, dtype=torch.float64, device=dev)
    st = torch.empty(nb, dtype=torch.int32, device=dev)
    N.call("ck_tridiag_eigh_f64", nb, m, N.ptr(dd), m, N.ptr(ee), max(m - 1, 1), N.ptr(ev), N.ptr(vecs), N.ptr(fr),
           N.ptr(st), N.stream_ptr(dev))
    fr2 = torch.empty_like(fr)
    ev2 = torch.empty_like(ev)
    N.call("ck_tridiag_eigh_f64", nb, m, N.ptr(dd), m, N.ptr(ee), max(m - 1, 1), N.ptr(ev2), None, N.ptr(fr2),
           N.ptr(st), N.stream_ptr(dev))
    torch.cuda.synchronize()
    assert (st.cpu() == 0).all()
    # the eigenvector kernel (overlapped sweeps) and the first-row kernel run the same recurrence
    assert torch.equal(ev, ev2)
    for b in range(nb):
        w, v = eigh_tridiag(d[b], e[b, :m - 1])
        assert np.allclose(ev[b].cpu().numpy(), w, rtol=1e-12, atol=1e-12)
        assert np.allclose(ev2[b].cpu().numpy(), w, rtol=1e-12, atol=1e-12)
        V = vecs[b].cpu().numpy()
        T = np.diag(d[b].astype(np.float64)) + np.diag(e[b, :m - 1].astype(np.float64), 1) + \
            np.diag(e[b, :m - 1].astype(np.float64), -1)
        assert np.abs(T @ V - V * w[None, :]).max() < 1e-11
        assert np.abs(V.T @ V - np.eye(m)).max() < 1e-12
        assert np.allclose(np.abs(fr[b].cpu().numpy()), np.abs(v[0]), atol=1e-10)
        assert np.allclose(np.abs(fr2[b].cpu().numpy()), np.abs(v[0]), atol=1e-10)


@pytest.mark.parametrize("m", [64, 100, 127])
def test_tridiag_eigh_split_blocks(dev, m):
    """Off-diagonals with exact zeros every 17 entries and tiny ones every 13: the matrix splits
    into blocks, sweeps start at every block and deflate early; vectors against LAPACK."""
    rs = np.random.RandomState(m)
    d = rs.normal(size=(1, m)).astype(np.float32)
    e = rs.uniform(0.1, 1.0, size=(1, m - 1)).astype(np.float32)
    e[0, ::17] = 0.0
    e[0, 5::13] = 1e-30
    dd, ee = torch.from_numpy(d).to(dev), torch.from_numpy(e).to(dev)
    ev = torch.empty(1, m, dtype=torch.float64, device=dev)
    vecs = torch.empty(1, m, m, dtype=torch.float64, device=dev)
    st = torch.empty(1, dtype=torch.int32, device=dev)
    N.call("ck_tridiag_eigh_f64", 1, m, N.ptr(dd), m, N.ptr(ee), m - 1, N.ptr(ev), N.ptr(vecs), None, N.ptr(st),
           N.stream_ptr(dev))
    torch.cuda.synchronize()
    assert int(st.cpu()[0]) == 0
    w, _ = eigh_tridiag(d[0], e[0])
    assert np.allclose(ev[0].cpu().numpy(), w, rtol=1e-12, atol=1e-12)
    V = vecs[0].cpu().numpy()
    T = np.diag(d[0].astype(np.float64)) + np.diag(e[0].astype(np.float64), 1) + np.diag(e[0].astype(np.float64), -1)
    assert np.abs(T @ V - V * w[None, :]).max() < 1e-11
    assert np.abs(V.T @ V - np.eye(m)).max() < 1e-12


def _bisect(dev, d, e, use_work=True):
    """ck_tridiag_eigh_bisect_f64 on fp32 rows d (nb, m), e (nb, m-1); returns evals, evecs, first_row, status.
    With a workspace, batches of more than 512 LDS blocks take the workspace kernel."""
    nb, m = d.shape
    dd, ee = torch.from_numpy(d).to(dev), torch.from_numpy(np.ascontiguousarray(e)).to(dev)
    ev = torch.empty(nb, m, dtype=torch.float64, device=dev)
    vecs = torch.empty(nb, m, m, dtype=torch.float64, device=dev)
    fr = torch.empty(nb, m, dtype=torch.float64, device=dev)
    st = torch.full((nb,), -1, dtype=torch.int32, device=dev)
    wb = N.q("ck_tridiag_eigh_bisect_workspace_bytes", nb, m) if use_work else 0
    work = torch.empty(wb, dtype=torch.uint8, device=dev) if wb else None
    N.call("ck_tridiag_eigh_bisect_f64", nb, m, N.ptr(dd), m, N.ptr(ee), max(e.shape[1], 1), N.ptr(ev), N.ptr(vecs),
           N.ptr(fr), N.ptr(st), N.ptr(work), wb, N.stream_ptr(dev))
    torch.cuda.synchronize()
    return ev.cpu().numpy(), vecs.cpu().numpy(), fr.cpu().numpy(), st.cpu().numpy()


def _check_decomposition(d, e, w, V, fr, what):
    """Eigenvalues against LAPACK to 1e-12 of ||T||, residual ||T v - lambda v|| and orthogonality
    to 1e-11 of ||T||, the sign rule (largest-magnitude component positive), first_row = V[0]."""
    m = d.shape[0]
    d64, e64 = d.astype(np.float64), e.astype(np.float64)
    T = np.diag(d64) + np.diag(e64, 1) + np.diag(e64, -1)
    tn = max(np.abs(np.linalg.eigvalsh(T)).max(), 1e-300)
    wl, _ = eigh_tridiag(d, e)
    res = np.abs(T @ V - V * w[None, :]).max() / tn
    orth = np.abs(V.T @ V - np.eye(m)).max()
    print(f"{what}: m={m} |w - LAPACK|/||T|| {np.abs(w - wl).max() / tn:.1e} residual {res:.1e} "
          f"orthogonality {orth:.1e}", flush=True)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - wl).max() <= 1e-12 * tn
    assert res < 1e-11 and orth < 1e-11
    # sign rule, up to rounding ties (symmetric vectors have equal-magnitude components)
    assert np.all(V.max(axis=0) >= np.abs(V).max(axis=0) - 1e-9)
    assert np.array_equal(fr, V[0])


@pytest.mark.parametrize("m,nb", [(1, 1), (2, 3), (10, 4), (30, 32), (120, 1), (120, 6), (200, 2), (700, 1), (1024, 1),
                                  (30, 600), (120, 64)])
def test_tridiag_bisect_random(dev, m, nb):
    """Bisection + inverse iteration on random tridiagonals: against LAPACK, residual and
    orthogonality bars, and the same eigenvalues as the QL kernel to 1e-12."""
    rs = np.random.RandomState(100 + m + nb)
    d = rs.normal(size=(nb, m)).astype(np.float32)
    e = rs.uniform(0.1, 1.0, size=(nb, max(m - 1, 1))).astype(np.float32)[:, :m - 1]
    w, V, fr, st = _bisect(dev, d, e if m > 1 else np.zeros((nb, 1), np.float32))
    assert (st == 0).all()
    for b in range(nb):
        _check_decomposition(d[b], e[b], w[b], V[b], fr[b], "random")


def _wilkinson(m):
    """W_m^+: d = |i - (m-1)/2|, e = 1; its largest eigenvalues come in pairs that agree to ~1e-14."""
    return np.abs(np.arange(m) - (m - 1) / 2.0).astype(np.float32), np.ones(m - 1, np.float32)


def _repeated(m):
    """Diagonal with each value three times and zero off-diagonals: exact triple eigenvalues."""
    rs = np.random.RandomState(m)
    return np.repeat(rs.normal(size=(m + 2) // 3), 3)[:m].astype(np.float32), np.zeros(m - 1, np.float32)


def _graded(m):
    """Entries graded over 12 decades: eigenvalues from ~1 down to ~1e-12 of it."""
    i = np.arange(m)
    d = 10.0 ** (-12.0 * i / (m - 1))
    e = 0.1 * 10.0 ** (-12.0 * (i[:-1] + 0.5) / (m - 1))
    return d.astype(np.float32), e.astype(np.float32)


def _split(m):
    rs = np.random.RandomState(m)
    d = rs.normal(size=m).astype(np.float32)
    e = rs.uniform(0.1, 1.0, size=m - 1).astype(np.float32)
    e[::17] = 0.0
    e[5::13] = 1e-30
    return d, e


def _lanczos_laplacian(m):
    """T of m unreorthogonalised Lanczos steps on a 1-D Laplacian (n = 100 < m): converged Ritz
    values reappear as ghost copies (45 gaps below 1e-8 for m = 250), tight clusters the
    inverse iteration has to separate."""
    n = 100
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    q = np.random.RandomState(3).normal(size=n)
    q /= np.linalg.norm(q)
    qp, beta = np.zeros(n), 0.0
    al, be = [], []
    for _ in range(m):
        w = A @ q - beta * qp
        a = q @ w
        w -= a * q
        beta = np.linalg.norm(w)
        al.append(a)
        be.append(beta)
        qp, q = q, w / beta
    return np.array(al, np.float32), np.array(be[:-1], np.float32)


@pytest.mark.parametrize("kind", ["wilkinson", "repeated", "graded", "split", "lanczos_ghosts"])
def test_tridiag_bisect_hard_spectra(dev, kind):
    """Clustered, repeated, graded and split spectra: the cluster rule and the perturbed shifts
    keep the vectors orthogonal and their residuals at FP64 level."""
    gen = {"wilkinson": lambda: _wilkinson(41), "repeated": lambda: _repeated(60), "graded": lambda: _graded(50),
           "split": lambda: _split(127), "lanczos_ghosts": lambda: _lanczos_laplacian(250)}[kind]
    d, e = gen()
    w, V, fr, st = _bisect(dev, d[None], e[None])
    assert st[0] == 0
    _check_decomposition(d, e, w[0], V[0], fr[0], kind)


def test_tridiag_bisect_matches_ql_on_lanczos_run(dev):
    """The eigenvector solver the Lanczos wrappers use by default against the QL kernel on one
    FP32 Lanczos tridiagonal (a 64^2-grid Laplacian, 120 steps): eigenvalues to 1e-12 of ||T||,
    Ritz vectors equal up to sign, to 1e-9."""
    import cola_amd as ca
    from cola_amd.linalg.lanczos import run_lanczos, tridiag_eigh
    from oracle import operators as oop
    ip, ix, v = oop.laplacian_2d_csr(64, 0.0)
    A = ca.Sparse.from_csr(torch.from_numpy(ip).to(dev), torch.from_numpy(ix).to(dev), torch.from_numpy(v).to(dev),
                           (4096, 4096))
    run = run_lanczos(ca.SelfAdjoint(A), torch.randn(4096, generator=torch.Generator().manual_seed(1)).to(dev), 120,
                      0.0)
    wq, Vq, _, sq = tridiag_eigh(run, vectors=True, method="ql")
    wb, Vb, _, sb = tridiag_eigh(run, vectors=True)
    assert int(sq.max()) == 0 and int(sb.max()) == 0
    wq, Vq, wb, Vb = wq[0].cpu().numpy(), Vq[0].cpu().numpy(), wb[0].cpu().numpy(), Vb[0].cpu().numpy()
    tn = np.abs(wq).max()
    assert np.abs(wq - wb).max() <= 1e-12 * tn
    # distinct eigenvalues: vectors agree up to sign
    gap = np.minimum(np.diff(wq, prepend=-np.inf), np.diff(wq, append=np.inf))
    ok = gap > 1e-6 * tn
    sgn = np.sign(np.sum(Vq * Vb, axis=0))
    print("vector diff", np.abs(Vq * sgn - Vb)[:, ok].max(), "distinct", int(ok.sum()), flush=True)
    assert np.abs(Vq * sgn - Vb)[:, ok].max() < 1e-9


def test_tridiag_bisect_nonfinite_input_reports_nan(dev):
    """A NaN or infinite entry (in d or e) gives CK_ERR_NAN for that problem and leaves the
    others as they are."""
    rs = np.random.RandomState(5)
    d = rs.normal(size=(3, 20)).astype(np.float32)
    e = rs.uniform(0.1, 1.0, size=(3, 19)).astype(np.float32)
    d[1, 7] = np.nan
    e[2, 3] = np.inf
    w, V, fr, st = _bisect(dev, d, e)
    assert st[0] == 0 and st[1] == N.CK_ERR_NAN and st[2] == N.CK_ERR_NAN
    _check_decomposition(d[0], e[0], w[0], V[0], fr[0], "finite neighbour")


def test_tridiag_bisect_large_batch_without_workspace(dev):
    """Without a workspace a large batch stays on the LDS kernel: same bars."""
    rs = np.random.RandomState(7)
    m, nb = 30, 300
    d = rs.normal(size=(nb, m)).astype(np.float32)
    e = rs.uniform(0.1, 1.0, size=(nb, m - 1)).astype(np.float32)
    w, V, fr, st = _bisect(dev, d, e, use_work=False)
    assert (st == 0).all()
    for b in range(0, nb, 7):
        _check_decomposition(d[b], e[b], w[b], V[b], fr[b], "random")


def test_tridiag_auto_method_on_a_clustered_spectrum(dev):
    """The Lanczos tridiagonal of an RBF kernel matrix: most Ritz values sit in one cluster, so
    the automatic choice takes QL; eigenvalues and vectors pass the same bars."""
    from cola_amd.linalg.lanczos import CLUSTER_LIMIT, largest_cluster, tridiag_eigh_batch
    rs = np.random.RandomState(3)
    pts = rs.uniform(size=(400, 3))
    K = np.exp(-((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1) / (2 * 0.5 ** 2))
    m = 100
    # Lanczos in FP64 with full re-orthogonalisation on the host: T's diagonals
    Q = np.zeros((400, m + 1))
    q = rs.normal(size=400)
    Q[:, 0] = q / np.linalg.norm(q)
    d, e = np.zeros(m), np.zeros(m)
    for j in range(m):
        w = K @ Q[:, j]
        d[j] = Q[:, j] @ w
        w -= Q[:, :j + 1] @ (Q[:, :j + 1].T @ w)
        w -= Q[:, :j + 1] @ (Q[:, :j + 1].T @ w)
        e[j] = np.linalg.norm(w)
        Q[:, j + 1] = w / e[j]
    d32, e32 = d.astype(np.float32)[None, :], e[:m - 1].astype(np.float32)[None, :]
    assert largest_cluster(np.linalg.eigvalsh(np.diag(d32[0].astype(np.float64)) + np.diag(e32[0].astype(np.float64), 1)
                                              + np.diag(e32[0].astype(np.float64), -1))[None, :]) > CLUSTER_LIMIT
    ev, V, fr, st = tridiag_eigh_batch(torch.from_numpy(d32).to(dev), torch.from_numpy(e32).to(dev), vectors=True)
    assert int(st.max()) == 0
    # the QL bars (its vectors follow no sign rule): eigenvalues, residual, orthogonality
    w, Vn, frn = ev[0].cpu().numpy(), V[0].cpu().numpy(), fr[0].cpu().numpy()
    T = np.diag(d32[0].astype(np.float64)) + np.diag(e32[0].astype(np.float64), 1) + np.diag(e32[0].astype(np.float64), -1)
    wl = np.linalg.eigvalsh(T)
    tn = np.abs(wl).max()
    assert np.abs(w - wl).max() <= 1e-12 * tn
    assert np.abs(T @ Vn - Vn * w[None, :]).max() / tn < 1e-11
    assert np.abs(Vn.T @ Vn - np.eye(m)).max() < 1e-11
    assert np.allclose(frn, Vn[0], rtol=0, atol=1e-14)
