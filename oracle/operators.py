"""Matvecs of the operators on the hot path, (n, B) row-major like the reference.

* Dense      -- cola/ops/operators.py:26-28   (cast to the promoted dtype, matmul)
* Sparse     -- cola/ops/operators.py:64-78   (COO sorted by row -> CSR int32 -> A @ V)
* ScalarMul  -- cola/ops/operators.py:97-98   (c * v)
* Sum        -- cola/ops/operators.py:182-183 (sum of member matvecs)
* Kronecker  -- cola/ops/operators.py:216-223 (reshape to factor axes, apply each
                factor along its axis with moveaxis/reshape/matmul)
"""
import numpy as np
import torch


def dense_matmat(A, X):
    """operators.py:26-28."""
    dtype = torch.promote_types(A.dtype, X.dtype)
    return A.to(dtype) @ X.to(dtype)


def coo_to_csr(data, rows, cols, shape):
    """operators.py:64-75: argsort by row, then scipy COO->CSR with int32 indices.

    scipy's tocsr sums duplicates and sorts columns within a row; restated with
    numpy here (stable sort by (row, col), duplicates summed)."""
    data = np.asarray(data)
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    order = np.lexsort((cols, rows))
    rows, cols, data = rows[order], cols[order], data[order]
    key = rows * shape[1] + cols
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(data, start) if len(data) else data
    r = (uniq // shape[1]).astype(np.int64)
    c = (uniq % shape[1]).astype(np.int32)
    indptr = np.zeros(shape[0] + 1, dtype=np.int32)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr).astype(np.int32)
    return indptr, c, summed.astype(data.dtype)


def csr_matmat(indptr, indices, values, X):
    """operators.py:77-78 (``torch.sparse_csr_tensor @ X``), restated as a row loop."""
    indptr = torch.as_tensor(np.asarray(indptr, dtype=np.int64))
    indices = torch.as_tensor(np.asarray(indices, dtype=np.int64))
    values = torch.as_tensor(values)
    n = indptr.numel() - 1
    row_of = torch.repeat_interleave(torch.arange(n), indptr[1:] - indptr[:-1])
    out = torch.zeros((n, X.shape[1]), dtype=torch.promote_types(values.dtype, X.dtype))
    out.index_add_(0, row_of, values[:, None].to(out.dtype) * X[indices].to(out.dtype))
    return out


def kron_matmat(factors, v):
    """operators.py:216-223: v (N, B) -> (K1 x K2 x ... ) v."""
    ev = v.reshape(*[M.shape[-1] for M in factors], -1)
    for i, M in enumerate(factors):
        ev_front = torch.movedim(ev, i, 0)
        shape = (M.shape[0], *ev_front.shape[1:])
        Mev_front = (M @ ev_front.reshape(M.shape[-1], -1)).reshape(shape)
        ev = torch.movedim(Mev_front, 0, i)
    rows = int(np.prod([M.shape[0] for M in factors]))
    return ev.reshape(rows, ev.shape[-1])


def shifted(matmat, shift):
    """Sum(A, ScalarMul(shift)) (operators.py:182-183 with :97-98): A v + shift v."""
    def f(X):
        return matmat(X) + shift * X
    return f


# ------------------------------------------------------------- synthetic operators
def laplacian_2d_csr(m, shift=0.0, dtype=np.float32):
    """5-point Dirichlet Laplacian on an m x m grid: diagonal 4+shift, neighbours -1.

    Row/column index = y*m + x; columns sorted within each row."""
    n = m * m
    y, x = np.divmod(np.arange(n), m)
    rows, cols, vals = [], [], []
    for dy, dx, v in ((-1, 0, -1.0), (0, -1, -1.0), (0, 0, 4.0 + shift), (0, 1, -1.0), (1, 0, -1.0)):
        ok = (y + dy >= 0) & (y + dy < m) & (x + dx >= 0) & (x + dx < m)
        r = np.nonzero(ok)[0]
        rows.append(r)
        cols.append(r + dy * m + dx)
        vals.append(np.full(r.size, v))
    return coo_to_csr(np.concatenate(vals).astype(dtype), np.concatenate(rows), np.concatenate(cols), (n, n))
