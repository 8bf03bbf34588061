"""The Matrix Market reader against scipy.io.mmread on the reference's fixture and on small
files of every supported layout, field and symmetry (CPU)."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse
import torch

import cola_amd as ca
from cola_amd.io import load_mtx, read_mtx

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dense_of(data, rows, cols, shape):
    return scipy.sparse.coo_matrix((data, (rows, cols)), shape=shape).toarray()


def test_1138_bus_matches_mmread():
    path = os.path.join(GOLD, "1138_bus.mtx")
    data, rows, cols, shape = read_mtx(path)
    M = scipy.io.mmread(path)
    assert shape == M.shape and data.size == M.nnz
    assert np.array_equal(dense_of(data, rows, cols, shape), M.toarray())
    A = load_mtx(path, device="cpu", dtype=torch.float64)
    C = M.tocsr()
    assert np.array_equal(A.indptr.numpy(), C.indptr) and np.array_equal(A.indices.numpy(), C.indices)
    assert np.array_equal(A.data.numpy(), C.data)


CASES = {
    "general_real": "%%MatrixMarket matrix coordinate real general\n% c\n3 4 4\n1 2 1.5\n3 1 -2\n2 4 3e-1\n1 2 0.25\n",
    "symmetric_int": "%%MatrixMarket matrix coordinate integer symmetric\n3 3 4\n1 1 2\n2 1 -1\n3 2 -1\n3 3 2\n",
    "skew": "%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 4.0\n3 1 -1.5\n",
    "pattern": "%%MatrixMarket matrix coordinate pattern general\n2 3 3\n1 1\n2 3\n1 3\n",
    "array_general": "%%MatrixMarket matrix array real general\n2 3\n1\n2\n3\n4\n5\n6\n",
    "array_symmetric": "%%MatrixMarket matrix array real symmetric\n3 3\n1\n2\n3\n4\n5\n6\n",
    "empty": "%%MatrixMarket matrix coordinate real general\n4 4 0\n",
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_small_files_match_mmread(tmp_path, name):
    path = tmp_path / f"{name}.mtx"
    path.write_text(CASES[name])
    data, rows, cols, shape = read_mtx(str(path))
    ref = scipy.io.mmread(str(path))
    ref = ref.toarray() if scipy.sparse.issparse(ref) else np.asarray(ref)
    assert shape == ref.shape
    assert np.array_equal(dense_of(data, rows, cols, shape), ref)


@pytest.mark.parametrize("text", ["%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
                                  "%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n",
                                  "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",
                                  "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n"])
def test_bad_files_raise(tmp_path, text):
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    with pytest.raises(ValueError):
        read_mtx(str(path))


def test_sparse_cpu_duplicates_and_order():
    """The CPU Sparse keeps the reference's CSR: columns ascending, duplicates summed."""
    data = torch.tensor([1., 2., 3., 4., 5.], dtype=torch.float64)
    rows = torch.tensor([2, 0, 2, 0, 2])
    cols = torch.tensor([1, 3, 0, 3, 1])
    A = ca.Sparse(data, rows, cols, (3, 4))
    assert A.indptr.tolist() == [0, 1, 1, 3]
    assert A.indices.tolist() == [3, 0, 1]
    assert A.data.tolist() == [6., 3., 6.]
