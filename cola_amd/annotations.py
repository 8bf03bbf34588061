"""Operator annotations (cola/annotations.py:27-193).

``PSD(A)`` returns a copy of ``A`` carrying the annotation, like cola's
``WrapMeta.__call__`` (annotations.py:34-42).  Inference rules for composites
follow annotations.py:80-193: Kronecker and Sum keep the intersection of their
members' annotations (Sum drops Unitary/Stiefel), a Product with one
non-scalar member inherits that member's, Identity is Unitary and PSD.
"""
import copy
from functools import reduce


class WrapMeta(type):
    def __str__(cls):
        return cls.__name__

    def __repr__(cls):
        return cls.__name__

    def __call__(cls, obj):
        new = copy.copy(obj)
        new.annotations = set(obj.annotations) | {cls}
        return new


class Annotation(metaclass=WrapMeta):
    pass


class SelfAdjoint(Annotation):
    """A^H = A (symmetric for real operators)."""


Hermitian = SelfAdjoint


class PSD(SelfAdjoint):
    """Positive semi-definite and self-adjoint."""


class Stiefel(Annotation):
    """A^H A = I."""


class Unitary(Stiefel):
    """A^H A = A A^H = I."""


def intersect_annotations(ops):
    return reduce(lambda x, y: x & y, (set(op.annotations) for op in ops))


def get_annotations(A):
    """Annotations inferred from structure when an operator is built."""
    from . import ops as O
    if isinstance(A, O.Identity):
        return {Unitary, PSD}
    if isinstance(A, O.Kronecker):
        return intersect_annotations(A.Ms)
    if isinstance(A, O.Sum):
        return intersect_annotations(A.Ms) - {Unitary, Stiefel}
    if isinstance(A, O.Product):
        not_commuting = [M for M in A.Ms if not isinstance(M, O.ScalarMul)]
        if len(not_commuting) == 1:
            return set(not_commuting[0].annotations)
        return intersect_annotations(A.Ms) & {Unitary, Stiefel}
    return set()
