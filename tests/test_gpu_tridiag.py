"""Device FP64 implicit-QL tridiagonal eigensolver against LAPACK (scipy)."""
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
