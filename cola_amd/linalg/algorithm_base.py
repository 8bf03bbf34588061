"""Algorithm objects (cola/linalg/algorithm_base.py:11-34)."""
from types import SimpleNamespace

from ..ops import LinearOperator


class Algorithm:
    """Marker base class of every algorithm choice."""


class IterativeOperatorWInfo(LinearOperator):
    """Lazy operator whose matmat runs ``alg(A, X) -> (Y, info)`` and keeps info
    (algorithm_base.py:16-29)."""

    def __init__(self, A, alg):
        self.device = A.device
        super().__init__(A.dtype, A.shape)
        self.A = A
        self.alg = alg
        self.info = {}

    def _matmat(self, X):
        Y, self.info = self.alg(self.A, X)
        return Y

    def __str__(self):
        return f"{self.alg}({self.A})"


class Auto(SimpleNamespace, Algorithm):
    """Chooses an algorithm from the operator's annotations and size; keyword
    arguments are forwarded to the chosen algorithm (cola/linalg/inverse/inv.py:85)."""
