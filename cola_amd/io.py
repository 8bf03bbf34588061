"""Matrix Market (.mtx) files to the engine's Sparse operator.

The reference's sparse tests read SuiteSparse matrices with scipy.io.mmread and build
``Sparse(data, row, col, shape)`` from the COO triplets (cola/ops/operators.py:64-75);
``load_mtx`` does the same in one call, with the COO -> CSR step on the device.

``read_mtx`` parses the text format itself: the ``%%MatrixMarket matrix`` banner,
``coordinate`` or ``array`` layout, ``real``/``double``/``integer``/``pattern`` fields and
``general``/``symmetric``/``skew-symmetric``/``hermitian`` (real) symmetry.  Symmetric files
store one triangle; the other is mirrored (negated for skew-symmetric), as mmread does.
Indices in the file are 1-based; the triplets returned are 0-based.
"""
import numpy as np
import torch

from .ops import Sparse

_FIELDS = {"real", "double", "integer", "pattern"}
_SYMMETRIES = {"general", "symmetric", "skew-symmetric", "hermitian"}


def _banner(line):
    parts = line.strip().lower().split()
    if len(parts) != 5 or parts[0] != "%%matrixmarket" or parts[1] != "matrix":
        raise ValueError(f"not a Matrix Market matrix banner: {line.strip()!r}")
    layout, field, symmetry = parts[2], parts[3], parts[4]
    if layout not in ("coordinate", "array"):
        raise ValueError(f"unknown Matrix Market layout {layout!r}")
    if field not in _FIELDS:
        raise ValueError(f"Matrix Market field {field!r} is not supported (real, integer or pattern)")
    if symmetry not in _SYMMETRIES:
        raise ValueError(f"unknown Matrix Market symmetry {symmetry!r}")
    if layout == "array" and field == "pattern":
        raise ValueError("a pattern matrix must use the coordinate layout")
    return layout, field, symmetry


def read_mtx(path):
    """(data float64, rows int64, cols int64, shape) of a Matrix Market file, 0-based."""
    with open(path, "r") as f:
        layout, field, symmetry = _banner(f.readline())
        line = f.readline()
        while line and (line.startswith("%") or not line.strip()):
            line = f.readline()
        if not line:
            raise ValueError("Matrix Market file has no size line")
        size = [int(t) for t in line.split()]
        body = f.read().split()
    if layout == "coordinate":
        if len(size) != 3:
            raise ValueError(f"coordinate size line needs 'rows cols entries', got {line.strip()!r}")
        n, m, nnz = size
        width = 2 if field == "pattern" else 3
        if len(body) != width * nnz:
            raise ValueError(f"expected {nnz} entries of {width} fields, found {len(body)} fields")
        vals = np.array(body, dtype=np.float64).reshape(nnz, width) if nnz else np.zeros((0, width))
        rows = vals[:, 0].astype(np.int64) - 1
        cols = vals[:, 1].astype(np.int64) - 1
        data = np.ones(nnz) if field == "pattern" else vals[:, 2].copy()
    else:
        if len(size) != 2:
            raise ValueError(f"array size line needs 'rows cols', got {line.strip()!r}")
        n, m = size
        vals = np.array(body, dtype=np.float64)
        if symmetry == "general":
            if vals.size != n * m:
                raise ValueError(f"expected {n * m} array values, found {vals.size}")
            cols, rows = np.divmod(np.arange(n * m, dtype=np.int64), n)  # column-major
            data = vals
        else:  # lower triangle by columns (strictly lower for skew-symmetric)
            lo = 1 if symmetry == "skew-symmetric" else 0
            rr, cc = [], []
            for j in range(m):
                r = np.arange(j + lo, n, dtype=np.int64)
                rr.append(r)
                cc.append(np.full(r.size, j, dtype=np.int64))
            rows, cols = np.concatenate(rr), np.concatenate(cc)
            if vals.size != rows.size:
                raise ValueError(f"expected {rows.size} array values, found {vals.size}")
            data = vals
    if rows.size and (rows.min() < 0 or rows.max() >= n or cols.min() < 0 or cols.max() >= m):
        raise ValueError("Matrix Market entry index outside the stated size")
    if symmetry != "general":
        off = rows != cols
        sign = -1.0 if symmetry == "skew-symmetric" else 1.0
        rows, cols, data = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([data, sign * data[off]]))
    return data, rows, cols, (n, m)


def load_mtx(path, device="cuda", dtype=torch.float32):
    """A Matrix Market file as ``Sparse`` on ``device`` (COO -> CSR by the engine's kernels
    when the device is the GPU)."""
    data, rows, cols, shape = read_mtx(path)
    dev = torch.device(device)
    return Sparse(torch.from_numpy(data).to(dtype=dtype, device=dev), torch.from_numpy(rows).to(dev),
                  torch.from_numpy(cols).to(dev), shape)
