"""FP64 symmetric tridiagonal eigensolve (LAPACK via scipy), the check for the
engine's device QL solver.  The reference calls ``xnp.eigh(T.to_dense())`` in A's
dtype (lanczos.py:57, unary.py:52, slq.py:44); the engine keeps FP64 here."""
import numpy as np
from scipy.linalg import eigh_tridiagonal


def eigh_tridiag(diag, offdiag, first_row_only=False):
    """Ascending eigenvalues and eigenvectors of the symmetric tridiagonal (diag, offdiag).

    Eigenvector signs are normalised so that the first nonzero component is
    positive is NOT imposed: callers compare sign-invariant quantities."""
    d = np.asarray(diag, dtype=np.float64)
    e = np.asarray(offdiag, dtype=np.float64)
    if d.size == 1:
        return d.copy(), np.ones((1, 1))
    w, v = eigh_tridiagonal(d, e)
    return (w, v[:1, :]) if first_row_only else (w, v)
