"""A small multiple-dispatch registry with the semantics cola relies on.

cola routes ``inv``/``eig``/``slogdet``/``diag``/``trace``/``apply_unary`` through
plum's ``@dispatch`` on (operator type, algorithm type) (cola/linalg/inverse/inv.py:42-165,
cola/linalg/eig/eigs.py:19-182, cola/linalg/logdet/logdet.py:52-184,
cola/linalg/trace/diag_trace.py:22-147, cola/linalg/unary/unary.py:94-210).
The rules used there are:

* a method applies when every argument is an instance of the annotated type
  (``A | B`` unions allowed) and its optional ``cond`` predicate holds;
* among applicable methods the most specific signature wins (each parameter
  type a subclass of the other's); structural rules therefore beat the
  algorithm base cases, which cola also marks ``precedence=-1``;
* remaining ties are broken by ``precedence`` (higher wins); a real tie is an
  error, as in plum.
"""
import inspect
import types
import typing
from functools import wraps

_REGISTRY = {}


def _types_of(ann):
    if ann is inspect.Parameter.empty or ann is typing.Any:
        return (object, )
    origin = typing.get_origin(ann)
    if origin is typing.Union or isinstance(ann, types.UnionType):
        out = []
        for a in typing.get_args(ann):
            out.extend(_types_of(a))
        return tuple(out)
    if isinstance(ann, type):
        return (ann, )
    return (object, )


class _Method:
    def __init__(self, fn, precedence, cond):
        self.fn = fn
        self.precedence = precedence
        self.cond = cond
        sig = inspect.signature(fn)
        self.params = [p for p in sig.parameters.values() if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
        self.types = [_types_of(p.annotation) for p in self.params]
        self.defaults = [p.default for p in self.params]

    def bind(self, args):
        if len(args) > len(self.params):
            return None
        full = list(args)
        for d in self.defaults[len(args):]:
            if d is inspect.Parameter.empty:
                return None
            full.append(d)
        for a, ts in zip(full, self.types):
            if not isinstance(a, ts):
                return None
        if self.cond is not None and not self.cond(*full):
            return None
        return full

    def more_specific(self, other):
        """Every parameter type of self is covered by (a subclass of) other's."""
        ge = True
        strict = False
        for ts, us in zip(self.types, other.types):
            if not all(any(issubclass(t, u) for u in us) for t in ts):
                ge = False
                break
            if not all(any(issubclass(u, t) for t in ts) for u in us):
                strict = True
        return ge and strict


class Function:
    def __init__(self, name):
        self.name = name
        self.methods = []
        self.abstract = None

    def register(self, fn, precedence=0, cond=None):
        self.methods.append(_Method(fn, precedence, cond))

    def resolve(self, args):
        cands = [(m, m.bind(args)) for m in self.methods]
        cands = [(m, b) for m, b in cands if b is not None]
        if not cands:
            tys = ", ".join(type(a).__name__ for a in args)
            raise NotImplementedError(f"no method of {self.name} for ({tys})")
        best = [c for c in cands if not any(o[0].more_specific(c[0]) for o in cands if o is not c)]
        if len(best) > 1:
            top = max(m.precedence for m, _ in best)
            best = [c for c in best if c[0].precedence == top]
        if len(best) > 1:
            tys = ", ".join(type(a).__name__ for a in args)
            raise TypeError(f"ambiguous dispatch of {self.name} for ({tys})")
        return best[0]

    def __call__(self, *args, **kwargs):
        if kwargs or self.abstract is not None:
            args = self._merge_kwargs(args, kwargs)
        m, full = self.resolve(args)
        return m.fn(*full)

    def _merge_kwargs(self, args, kwargs):
        ref = self.abstract or self.methods[0].fn
        sig = inspect.signature(ref)
        bound = sig.bind_partial(*args, **kwargs)
        bound.apply_defaults()
        return tuple(bound.arguments.values())


def dispatch(fn=None, *, precedence=0, cond=None):
    """Register ``fn`` as a method of the generic function with the same name."""

    def deco(f):
        key = f.__name__
        gen = _REGISTRY.setdefault(key, Function(f.__name__))
        gen.register(f, precedence, cond)

        @wraps(f)
        def call(*args, **kwargs):
            return gen(*args, **kwargs)

        call._ck_function = gen
        return call

    if fn is not None:
        return deco(fn)
    return deco


def abstract(fn):
    """Declare the generic signature (defaults and docs) without a method."""
    key = fn.__name__
    gen = _REGISTRY.setdefault(key, Function(fn.__name__))
    gen.abstract = fn

    @wraps(fn)
    def call(*args, **kwargs):
        return gen(*args, **kwargs)

    call._ck_function = gen
    return call


dispatch.abstract = abstract
