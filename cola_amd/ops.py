"""Linear operators (cola/ops/operator_base.py, cola/ops/operators.py) for the engine.

Same classes and contract as cola: ``LinearOperator(dtype, shape, matmat)``,
``_matmat(X)`` on an (n, k) block, ``A @ x`` with 1-D promotion
(operator_base.py:97-106), ``A.isa(PSD)`` (operator_base.py:52-54).

On the HIP device the structures the engine knows -- Dense (fp32 or bf16
storage), Sparse CSR, Kronecker of dense factors, Diagonal, Identity,
ScalarMul, an on-the-fly RBF kernel, a Woodbury inverse, and any of these
scaled by a scalar and/or shifted by c*I -- lower to one ``ck_op`` descriptor
(``hip_op()``) and run as the engine's own kernels.  Operators on CPU tensors
keep plain torch arithmetic; the Krylov algorithms themselves refuse CPU input.
"""
from functools import reduce

import numpy as np
import torch

from . import _native as N
from . import annotations as ann
from ._blocks import from_blocks, ld_of, to_blocks


_ENGINE = {"enabled": True}


class engine_disabled:
    """Context in which operators on HIP tensors use their torch arithmetic instead of the
    engine's kernels, so that torch autograd can differentiate A(theta) @ X with respect to the
    operator's tensors (the vector-Jacobian products of the iterative solvers' backward rules)."""

    def __enter__(self):
        self._prev = _ENGINE["enabled"]
        _ENGINE["enabled"] = False
        return self

    def __exit__(self, *exc):
        _ENGINE["enabled"] = self._prev
        return False


class _Leaf:
    __slots__ = ("i", )

    def __init__(self, i):
        self.i = i


class _OpTemplate:
    __slots__ = ("cls", "fields")

    def __init__(self, cls, fields):
        self.cls, self.fields = cls, fields


def _flatten(obj, leaves):
    if isinstance(obj, torch.Tensor):
        leaves.append(obj)
        return _Leaf(len(leaves) - 1)
    if isinstance(obj, LinearOperator):
        return _OpTemplate(type(obj), {k: _flatten(v, leaves) for k, v in sorted(vars(obj).items())})
    if type(obj) in (list, tuple):
        return type(obj)(_flatten(v, leaves) for v in obj)
    if type(obj) is dict:
        return {k: _flatten(v, leaves) for k, v in obj.items()}
    return obj


def _rebuild(t, params):
    if isinstance(t, _Leaf):
        return params[t.i]
    if isinstance(t, _OpTemplate):
        obj = object.__new__(t.cls)
        for k, v in t.fields.items():
            object.__setattr__(obj, k, _rebuild(v, params))
        obj._after_unflatten()
        return obj
    if type(t) in (list, tuple):
        return type(t)(_rebuild(v, params) for v in t)
    if type(t) is dict:
        return {k: _rebuild(v, params) for k, v in t.items()}
    return t


def _find_device(objs):
    for o in objs:
        if isinstance(o, torch.Tensor):
            return o.device
        if isinstance(o, LinearOperator):
            return o.device
        if isinstance(o, (list, tuple)):
            d = _find_device(o)
            if d is not None:
                return d
    return None


class LinearOperator:
    """Base class (operator_base.py:15-219)."""
    __array_ufunc__ = None

    def __init__(self, dtype, shape, matmat=None, annotations=frozenset(), device=None):
        self.dtype = dtype
        self.shape = tuple(int(s) for s in shape)
        if matmat is not None:
            self._matmat = matmat
        if not hasattr(self, "device") or self.device is None:
            self.device = device if device is not None else torch.device("cpu")
        self.annotations = ann.get_annotations(self)
        self.annotations.update(annotations)

    def isa(self, annotation):
        return any(issubclass(a, annotation) for a in self.annotations)

    def flatten(self):
        """(tensors, unflatten): every tensor reachable from the operator's attributes (nested
        operators, lists, tuples, dicts included) in a fixed order, and the function that builds
        the same operator around replacement tensors (operator_base.py:89-95, 191-219)."""
        leaves = []
        tmpl = _flatten(self, leaves)

        def unflatten(params):
            return _rebuild(tmpl, list(params))

        return leaves, unflatten

    def _after_unflatten(self):
        """Hook for operators that cache values derived from their tensors."""

    def _matmat(self, X):
        raise NotImplementedError

    # ---- structure recognition --------------------------------------------
    def hip_op(self):
        """(CkOp, keepalive list) when the engine has a kernel for this operator, else None."""
        return None

    def _hip_apply(self, X):
        op, keep = self.hip_op()
        n = self.shape[0]
        Xb = to_blocks(X)
        ld = ld_of(n)
        B = Xb.shape[0]
        Yb = torch.empty_like(Xb)
        wb = N.q("ck_op_workspace_bytes", op, B)
        work = torch.empty(max(1, (wb + 3) // 4), dtype=torch.float32, device=Xb.device) if wb else None
        N.call("ck_op_apply", op, B, N.ptr(Xb), ld, N.ptr(Yb), ld, N.ptr(work), wb, N.stream_ptr(Xb.device))
        del keep
        return from_blocks(Yb, n).to(X.dtype)

    def _use_hip(self, X):
        return (_ENGINE["enabled"] and isinstance(X, torch.Tensor) and X.is_cuda and X.dtype == torch.float32
                and self.hip_op() is not None)

    # ---- algebra ------------------------------------------------------------
    def __matmul__(self, X):
        if isinstance(X, LinearOperator):
            from .fns import dot
            return dot(self, X)
        assert X.shape[0] == self.shape[-1], f"dimension mismatch {self.shape} vs {tuple(X.shape)}"
        if X.dim() == 1:
            return self._matmat(X.reshape(-1, 1)).reshape(-1)
        return self._matmat(X)

    def __rmatmul__(self, X):
        if isinstance(X, LinearOperator):
            from .fns import dot
            return dot(X, self)
        if self.isa(ann.SelfAdjoint):
            return (self @ X.T).T
        return X @ self.to_dense()

    def to_dense(self):
        n = self.shape[-1]
        return self @ torch.eye(n, dtype=self.dtype, device=self.device)

    @property
    def T(self):
        if self.isa(ann.SelfAdjoint):
            return self
        return Dense(self.to_dense().T.contiguous())

    @property
    def H(self):
        return self.T

    def __add__(self, other):
        if isinstance(other, (int, float)) and other == 0:
            return self
        from .fns import add
        return add(self, other)

    def __radd__(self, other):
        return self.__add__(other)

    def __mul__(self, c):
        from .fns import mul
        return mul(self, c)

    def __rmul__(self, c):
        return self * c

    def __neg__(self):
        return -1 * self

    def __sub__(self, x):
        return self.__add__(-x)

    def to(self, device):
        raise NotImplementedError(f"{type(self).__name__}.to")

    def __repr__(self):
        return f"<{type(self).__name__} {self.shape[0]}x{self.shape[1]} {self.dtype}>"


def _op(kind, n, scale=1.0, shift=0.0):
    op = N.CkOp()
    op.kind, op.n, op.scale, op.shift = kind, n, scale, shift
    return op


# Order from which a self-adjoint dense operator is lowered with CK_OP_SYMMETRIC.  Below it
# the matrix (<= 512 MiB) is served largely from the 256 MB Infinity Cache and the full product
# is as fast or faster (MI355X, 16384 x 16384: 1 rhs 109 vs 177 us, 8 rhs 126 vs 170 us;
# 4096 x 4096: 28 vs 11 us).
SYMMETRIC_MIN_ORDER = 11776


def _mark_symmetric(A, got):
    """Sets CK_OP_SYMMETRIC on a dense lowering of an operator annotated self-adjoint and of
    order >= SYMMETRIC_MIN_ORDER: the engine then reads only the blocks on and above the
    diagonal (half the matrix bytes)."""
    if (got is not None and A.isa(ann.SelfAdjoint) and got[0].kind == N.CK_OP_DENSE_F32
            and got[0].n >= SYMMETRIC_MIN_ORDER):
        got[0].flags |= N.CK_OP_SYMMETRIC
    return got


class Dense(LinearOperator):
    """operators.py:12-38."""

    def __init__(self, A):
        self.A = A
        self.device = A.device
        super().__init__(A.dtype, A.shape)

    def _matmat(self, X):
        if self._use_hip(X):
            return self._hip_apply(X)
        dtype = torch.promote_types(self.dtype, X.dtype)
        return self.A.to(dtype) @ X.to(dtype)

    def hip_op(self):
        A = self.A
        if not (A.is_cuda and A.dtype == torch.float32 and A.dim() == 2 and A.shape[0] == A.shape[1]
                and A.stride(1) == 1):
            return None
        op = _op(N.CK_OP_DENSE_F32, A.shape[0])
        op.a, op.lda = A.data_ptr(), A.stride(0)
        return _mark_symmetric(self, (op, [A]))

    def to_dense(self):
        return self.A


class DenseBF16(LinearOperator):
    """A dense matrix stored in bfloat16; products accumulate in fp32 (engine-only)."""

    def __init__(self, A):
        self.A = A.to(torch.bfloat16).contiguous()
        self.device = A.device
        super().__init__(torch.float32, A.shape)

    def hip_op(self):
        A = self.A
        if not (A.is_cuda and A.shape[0] % 8 == 0):
            return None
        op = _op(N.CK_OP_DENSE_BF16, A.shape[0])
        op.a, op.lda = A.data_ptr(), A.stride(0)
        return op, [A]

    def _matmat(self, X):
        if self._use_hip(X):
            return self._hip_apply(X)
        return self.A.to(torch.float32) @ X.to(torch.float32)

    def to_dense(self):
        return self.A.to(torch.float32)


def coo_to_csr(data, rows, cols, shape):
    """COO -> CSR: rows in order, columns ascending inside each row, duplicates summed,
    int32 indices -- the arrays the reference keeps after argsort + scipy's
    ``coo_array(...).tocsr()`` (operators.py:64-75).

    Device tensors go through the engine's kernels (``ck_coo_to_csr_f32``/``_f64``: row
    counts, scan, scatter, per-row sort, duplicate merge); CPU tensors through torch."""
    if data.is_cuda:
        return _coo_to_csr_device(data, rows, cols, shape)
    dev = data.device
    rows = rows.to(device=dev, dtype=torch.int64)
    cols = cols.to(device=dev, dtype=torch.int64)
    key = rows * int(shape[1]) + cols
    key, order = torch.sort(key, stable=True)
    vals = data[order]
    uniq, inverse = torch.unique_consecutive(key, return_inverse=True)
    summed = torch.zeros(uniq.numel(), dtype=data.dtype, device=dev).index_add_(0, inverse, vals)
    r = uniq // int(shape[1])
    c = (uniq % int(shape[1])).to(torch.int32)
    counts = torch.bincount(r, minlength=int(shape[0]))
    indptr = torch.zeros(int(shape[0]) + 1, dtype=torch.int64, device=dev)
    indptr[1:] = torch.cumsum(counts, 0)
    return indptr.to(torch.int32), c, summed


def _coo_to_csr_device(data, rows, cols, shape):
    fn = {torch.float32: "ck_coo_to_csr_f32", torch.float64: "ck_coo_to_csr_f64"}.get(data.dtype)
    if fn is None:
        raise TypeError(f"sparse values on the device must be float32 or float64, got {data.dtype}")
    dev = data.device
    n, m = int(shape[0]), int(shape[1])
    data = data.contiguous()
    rows = rows.to(device=dev, dtype=torch.int64).contiguous()
    cols = cols.to(device=dev, dtype=torch.int64).contiguous()
    if rows.numel() != data.numel() or cols.numel() != data.numel():
        raise ValueError("data, row and column index arrays differ in length")
    if n >= 2**31 or m >= 2**31 or data.numel() >= 2**31:
        raise ValueError("CSR with int32 indices needs n, nnz < 2**31")
    nnz = data.numel()
    indptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
    indices = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    values = torch.empty(max(nnz, 1), dtype=data.dtype, device=dev)
    info = torch.zeros(2, dtype=torch.int32, device=dev)  # [unique count, index status]
    wb = N.q("ck_coo_to_csr_workspace_bytes", n, nnz)
    work = torch.empty(max(wb, 1), dtype=torch.uint8, device=dev)
    N.call(fn, n, m, nnz, N.ptr(data), N.ptr(rows), N.ptr(cols), N.ptr(indptr), N.ptr(indices), N.ptr(values),
           N.ptr(info), N.ptr(info[1:]), N.ptr(work), wb, N.stream_ptr(dev))
    k, status = info.tolist()
    if status:
        raise ValueError("sparse indices out of range for shape")
    return indptr, indices[:k], values[:k]


def _check_coo(rows, cols, shape):
    # kernels index x[col] and y[row] without bounds checks: validate once here
    if rows.numel() != cols.numel():
        raise ValueError("row and column index arrays differ in length")
    if rows.numel() and (int(rows.min()) < 0 or int(rows.max()) >= shape[0] or int(cols.min()) < 0
                         or int(cols.max()) >= shape[1]):
        raise ValueError("sparse indices out of range for shape")
    if shape[0] >= 2**31 or shape[1] >= 2**31 or rows.numel() >= 2**31:
        raise ValueError("CSR with int32 indices needs n, nnz < 2**31")


def _check_csr(indptr, indices, data, shape):
    if indptr.numel() != shape[0] + 1 or int(indptr[0]) != 0 or int(indptr[-1]) != data.numel():
        raise ValueError("indptr must have n+1 entries from 0 to nnz")
    if indices.numel() != data.numel():
        raise ValueError("indices and data differ in length")
    if shape[0] > 0 and bool((indptr[1:] < indptr[:-1]).any()):
        raise ValueError("indptr must be non-decreasing")
    if indices.numel() and (int(indices.min()) < 0 or int(indices.max()) >= shape[1]):
        raise ValueError("column indices out of range for shape")


class Sparse(LinearOperator):
    """operators.py:48-81 (COO input, CSR storage with int32 indices)."""

    def __init__(self, data, row_indices, col_indices, shape):
        self.device = data.device
        super().__init__(data.dtype, shape)
        if not data.is_cuda:  # the device kernels check the indices themselves
            _check_coo(row_indices, col_indices, shape)
        self.indptr, self.indices, self.data = coo_to_csr(data, row_indices, col_indices, shape)

    @classmethod
    def from_csr(cls, indptr, indices, data, shape):
        obj = cls.__new__(cls)
        obj.device = data.device
        LinearOperator.__init__(obj, data.dtype, shape)
        obj.indptr = indptr.to(device=data.device, dtype=torch.int32)
        obj.indices = indices.to(device=data.device, dtype=torch.int32)
        obj.data = data
        _check_csr(obj.indptr, obj.indices, data, shape)
        return obj

    def hip_op(self):
        if not (self.data.is_cuda and self.data.dtype == torch.float32 and self.shape[0] == self.shape[1]):
            return None
        op = _op(N.CK_OP_CSR_F32, self.shape[0])
        op.rowptr, op.colind, op.values = self.indptr.data_ptr(), self.indices.data_ptr(), self.data.data_ptr()
        op.nnz = self.data.numel()
        return op, [self.indptr, self.indices, self.data]

    def _matmat(self, V):
        if self._use_hip(V):
            return self._hip_apply(V)
        n = self.shape[0]
        counts = (self.indptr[1:] - self.indptr[:-1]).to(torch.int64)
        row_of = torch.repeat_interleave(torch.arange(n, device=V.device), counts)
        out = torch.zeros((n, V.shape[1]), dtype=torch.promote_types(self.dtype, V.dtype), device=V.device)
        out.index_add_(0, row_of, self.data[:, None].to(out.dtype) * V[self.indices.long()].to(out.dtype))
        return out

    def to_dense(self):
        n, m = self.shape
        D = torch.zeros((n, m), dtype=self.dtype, device=self.device)
        counts = (self.indptr[1:] - self.indptr[:-1]).to(torch.int64)
        row_of = torch.repeat_interleave(torch.arange(n, device=self.device), counts)
        D.index_put_((row_of, self.indices.long()), self.data, accumulate=True)
        return D


class ScalarMul(LinearOperator):
    """c * I (operators.py:84-101)."""

    def __init__(self, c, shape, dtype=None, device=None):
        self.device = device if device is not None else torch.device("cpu")
        # host copy of c for kernel arguments: reading a device scalar each call would be a
        # device-to-host copy that synchronises the stream.  A tensor c is re-read whenever its
        # version counter moves (an in-place update such as an optimizer step on a noise
        # hyperparameter), so the kernels never see a stale coefficient.
        self._c_host = None if isinstance(c, torch.Tensor) else float(c)
        self._c_version = None
        self.c = c if isinstance(c, torch.Tensor) else torch.tensor(c, dtype=dtype or torch.float32,
                                                                    device=self.device)
        self._c_is_tensor = isinstance(c, torch.Tensor)
        super().__init__(dtype or self.c.dtype, shape)

    def c_float(self):
        """c as a Python float: a Python-number coefficient is fixed; a tensor coefficient is
        read from the device the first time and again after every in-place change."""
        if self._c_is_tensor:
            ver = self.c._version
            if self._c_host is None or ver != self._c_version:
                self._c_host = float(self.c)
                self._c_version = ver
        return self._c_host

    def _matmat(self, v):
        return self.c.to(v.device) * v

    def _after_unflatten(self):
        # the coefficient is now whatever tensor was passed in: read it again when needed
        self._c_is_tensor = True
        self._c_host = None
        self._c_version = None

    def hip_op(self):
        if self.device.type != "cuda":
            return None
        return _op(N.CK_OP_IDENTITY, self.shape[0], scale=self.c_float()), []


class Identity(LinearOperator):
    """operators.py:104-128."""

    def __init__(self, shape, dtype, device=None):
        self.device = device if device is not None else torch.device("cpu")
        super().__init__(dtype, shape)

    def _matmat(self, X):
        return X

    def hip_op(self):
        if self.device.type != "cuda":
            return None
        return _op(N.CK_OP_IDENTITY, self.shape[0]), []


def I_like(A):
    """operators.py:130-135."""
    return Identity(A.shape, A.dtype, device=A.device)


class Product(LinearOperator):
    """operators.py:138-164."""

    def __init__(self, *Ms):
        from .fns import lazify
        self.Ms = tuple(lazify(M) for M in Ms)
        for M1, M2 in zip(self.Ms[:-1], self.Ms[1:]):
            if M1.shape[-1] != M2.shape[-2]:
                raise ValueError(f"dimension mismatch {M1.shape} vs {M2.shape}")
        self.device = self.Ms[0].device
        shape = (self.Ms[0].shape[-2], self.Ms[-1].shape[-1])
        dtype = reduce(torch.promote_types, (M.dtype for M in self.Ms))
        super().__init__(dtype, shape)

    def _matmat(self, v):
        for M in self.Ms[::-1]:
            v = M @ v
        return v

    def hip_op(self):
        # scalar * structured: fold the scalars into the structured op's scale
        scal = 1.0
        rest = []
        for M in self.Ms:
            if isinstance(M, ScalarMul):
                scal *= M.c_float()
            elif isinstance(M, Identity):
                continue
            else:
                rest.append(M)
        if len(rest) > 1:
            return None
        if not rest:
            if self.device.type != "cuda":
                return None
            return _op(N.CK_OP_IDENTITY, self.shape[0], scale=scal), []
        got = rest[0].hip_op()
        if got is None:
            return None
        op, keep = got
        op.scale *= scal
        op.shift *= scal
        return _mark_symmetric(self, (op, keep))


class Sum(LinearOperator):
    """operators.py:167-191."""

    def __init__(self, *Ms):
        from .fns import lazify
        self.Ms = tuple(lazify(M) for M in Ms)
        shape = self.Ms[0].shape
        for M in self.Ms:
            if M.shape != shape:
                raise ValueError(f"dimension mismatch {M.shape} vs {shape}")
        self.device = self.Ms[0].device
        super().__init__(self.Ms[0].dtype, shape)

    def _matmat(self, v):
        if self._use_hip(v):
            return self._hip_apply(v)
        return sum(M @ v for M in self.Ms)

    def hip_op(self):
        # one structured member plus any number of c*I members -> shift epilogue
        shift = 0.0
        structured = []
        for M in self.Ms:
            got = M.hip_op()
            if got is None:
                return None
            op, keep = got
            if op.kind == N.CK_OP_IDENTITY:
                shift += op.scale + op.shift
            else:
                structured.append((op, keep))
        if len(structured) > 1:
            return None
        if not structured:
            return _op(N.CK_OP_IDENTITY, self.shape[0], scale=shift), []
        op, keep = structured[0]
        op.shift += shift
        return _mark_symmetric(self, (op, keep))


class Triangular(Dense):
    """operators.py:41-45: a dense triangular matrix (lower or upper)."""

    def __init__(self, A, lower=True):
        super().__init__(A)
        self.lower = lower


class TriangularInv(LinearOperator):
    """inv.py:154-165: A^-1 for a Triangular A by triangular solves."""

    def __init__(self, A):
        self.A = A.to_dense()
        self.lower = A.lower
        self.device = self.A.device
        super().__init__(A.dtype, A.shape)

    def _matmat(self, X):
        return torch.linalg.solve_triangular(self.A.to(X.dtype), X, upper=not self.lower)

    def to_dense(self):
        return torch.linalg.solve_triangular(self.A, torch.eye(self.shape[0], dtype=self.A.dtype, device=self.device),
                                             upper=not self.lower)


class Permutation(LinearOperator):
    """operators.py:530-546: (P v)_i = v_{perm[i]}."""

    def __init__(self, perm, dtype=None):
        self.perm = perm
        self.device = perm.device
        super().__init__(torch.float32 if dtype is None else dtype, (len(perm), len(perm)))

    def _matmat(self, v):
        return v[self.perm]

    def to_dense(self):
        n = self.shape[0]
        return torch.eye(n, dtype=self.dtype, device=self.device)[self.perm]


class KronSum(LinearOperator):
    """operators.py:241-273: A (+) B = A (x) I + I (x) B, applied factor by factor."""

    def __init__(self, *Ms):
        from .fns import lazify
        self.Ms = tuple(lazify(M) for M in Ms)
        self.device = self.Ms[0].device
        shape = (int(np.prod([M.shape[-2] for M in self.Ms])), int(np.prod([M.shape[-1] for M in self.Ms])))
        dtype = reduce(torch.promote_types, (M.dtype for M in self.Ms))
        super().__init__(dtype, shape)

    def _matmat(self, v):
        ev = v.reshape(*[M.shape[-1] for M in self.Ms], -1)
        out = torch.zeros_like(ev)
        for i, M in enumerate(self.Ms):
            front = torch.movedim(ev, i, 0)
            Mf = (M @ front.reshape(M.shape[-1], -1)).reshape(M.shape[0], *front.shape[1:])
            out = out + torch.movedim(Mf, 0, i)
        return out.reshape(self.shape[-2], ev.shape[-1])

    def to_dense(self):
        def ks(A, B):
            IA = torch.eye(A.shape[0], dtype=A.dtype, device=A.device)
            IB = torch.eye(B.shape[0], dtype=B.dtype, device=B.device)
            return torch.kron(A, IB) + torch.kron(IA, B)
        return reduce(ks, [M.to_dense() for M in self.Ms])


class BlockDiag(LinearOperator):
    """operators.py:278-315: [M_0 0 ...; 0 M_1 ...] with optional multiplicities (block i
    repeated multiplicities[i] times along the diagonal)."""

    def __init__(self, *Ms, multiplicities=None):
        from .fns import lazify
        self.Ms = tuple(lazify(M) for M in Ms)
        self.multiplicities = [1 for _ in Ms] if multiplicities is None else list(multiplicities)
        shape = (sum(M.shape[-2] * c for M, c in zip(self.Ms, self.multiplicities)),
                 sum(M.shape[-1] * c for M, c in zip(self.Ms, self.multiplicities)))
        self.device = self.Ms[0].device
        dtype = reduce(torch.promote_types, (M.dtype for M in self.Ms))
        super().__init__(dtype, shape)

    def blocks(self):
        """The diagonal blocks in order, multiplicities expanded."""
        return [M for M, c in zip(self.Ms, self.multiplicities) for _ in range(c)]

    def _matmat(self, v):
        k = v.shape[1]
        i = 0
        y = []
        for M, mult in zip(self.Ms, self.multiplicities):
            i_end = i + mult * M.shape[-1]
            elems = M @ v[i:i_end].T.reshape(k * mult, M.shape[-1]).T
            y.append(elems.T.reshape(k, mult * M.shape[0]).T)
            i = i_end
        return torch.cat(y, dim=0)

    def to_dense(self):
        return torch.block_diag(*[M.to_dense() for M in self.blocks()])


class Kronecker(LinearOperator):
    """operators.py:198-230."""

    def __init__(self, *Ms):
        from .fns import lazify
        self.Ms = tuple(lazify(M) for M in Ms)
        self.device = self.Ms[0].device
        shape = (int(np.prod([M.shape[-2] for M in self.Ms])), int(np.prod([M.shape[-1] for M in self.Ms])))
        dtype = reduce(torch.promote_types, (M.dtype for M in self.Ms))
        super().__init__(dtype, shape)

    def _matmat(self, v):
        if self._use_hip(v):
            return self._hip_apply(v)
        ev = v.reshape(*[M.shape[-1] for M in self.Ms], -1)
        for i, M in enumerate(self.Ms):
            ev_front = torch.movedim(ev, i, 0)
            shape = (M.shape[0], *ev_front.shape[1:])
            Mev_front = (M @ ev_front.reshape(M.shape[-1], -1)).reshape(shape)
            ev = torch.movedim(Mev_front, 0, i)
        return ev.reshape(self.shape[-2], ev.shape[-1])

    def hip_op(self):
        if len(self.Ms) not in (2, 3):
            return None
        facs = []
        for M in self.Ms:
            if not isinstance(M, Dense):
                return None
            A = M.A
            if not (A.is_cuda and A.dtype == torch.float32 and A.shape[0] == A.shape[1] and A.shape[0] <= 64):
                return None
            facs.append(A.contiguous())
        op = _op(N.CK_OP_KRON_F32, self.shape[0])
        op.nfactors = len(facs)
        for i, A in enumerate(facs):
            op.fdim[i] = A.shape[0]
            op.factor[i] = A.data_ptr()
        return op, facs

    def to_dense(self):
        return reduce(torch.kron, [M.to_dense() for M in self.Ms])


class Diagonal(LinearOperator):
    """operators.py:323-348."""

    def __init__(self, diag):
        assert diag.dim() == 1
        self.diag = diag
        self.device = diag.device
        super().__init__(diag.dtype, (diag.numel(), diag.numel()))

    def _matmat(self, X):
        if self._use_hip(X):
            return self._hip_apply(X)
        return self.diag[:, None] * X

    def hip_op(self):
        d = self.diag
        if not (d.is_cuda and d.dtype == torch.float32 and d.is_contiguous()):
            return None
        op = _op(N.CK_OP_DIAG_F32, d.numel())
        op.a = d.data_ptr()
        return op, [d]

    def to_dense(self):
        return torch.diag(self.diag)


class Tridiagonal(LinearOperator):
    """operators.py:351-380 (alpha lower band, beta diagonal, gamma upper band)."""

    def __init__(self, alpha, beta, gamma):
        self.alpha, self.beta, self.gamma = alpha, beta, gamma
        self.device = beta.device
        super().__init__(beta.dtype, (beta.shape[-1], beta.shape[-1]))

    def to_dense(self):
        T = torch.diag(self.beta)
        if self.alpha.numel():
            T = T + torch.diag(self.alpha, -1) + torch.diag(self.gamma, 1)
        return T

    def _matmat(self, X):
        return self.to_dense().to(X.dtype) @ X


class StationaryKernelFn:
    """A stationary kernel k(x1_i, x2_j) = g(|x1_i - x2_j|) usable as the ``fn`` of ``Kernel``:
    called on two point blocks it returns the dense block (torch, as cola's Kernel expects),
    and its ``ck_family`` lets ``Kernel`` lower to the engine's on-the-fly kernel op
    (CK_OP_RBF_F32).  Families: 0 RBF exp(-r^2/(2 l^2)), 1 Matern 1/2, 2 Matern 3/2,
    3 Matern 5/2."""

    def __init__(self, family, lengthscale):
        if family not in (0, 1, 2, 3):
            raise ValueError("family must be 0 (RBF), 1, 2 or 3 (Matern 1/2, 3/2, 5/2)")
        self.ck_family, self.lengthscale = int(family), float(lengthscale)

    def __call__(self, x1, x2):
        # direct differences: the matmul form of cdist loses the small distances to cancellation
        r = torch.cdist(x1, x2, compute_mode="donot_use_mm_for_euclid_dist")
        ell = self.lengthscale
        if self.ck_family == 0:
            return torch.exp(-r**2 / (2 * ell * ell))
        if self.ck_family == 1:
            return torch.exp(-r / ell)
        a = (3.0**0.5 if self.ck_family == 2 else 5.0**0.5) * r / ell
        if self.ck_family == 2:
            return (1 + a) * torch.exp(-a)
        return (1 + a + a * a / 3) * torch.exp(-a)


def rbf(lengthscale):
    return StationaryKernelFn(0, lengthscale)


def matern12(lengthscale):
    return StationaryKernelFn(1, lengthscale)


def matern32(lengthscale):
    return StationaryKernelFn(2, lengthscale)


def matern52(lengthscale):
    return StationaryKernelFn(3, lengthscale)


class Kernel(LinearOperator):
    """Blocked on-the-fly kernel matvec [Kv]_i = sum_j f(x1_i, x2_j) v_j (operators.py:609-646).

    With x2 the same points as x1 (at most 8 coordinates, fp32 on the device) and ``fn`` a
    ``StationaryKernelFn`` (``rbf``, ``matern12``, ``matern32``, ``matern52``) the product runs
    in the engine's kernel, which forms each entry inside the matvec; any other ``fn`` is
    evaluated block by block as in the reference."""

    def __init__(self, x1, x2, fn, block_size1, block_size2):
        self.x1, self.x2, self.fn = x1, x2, fn
        self.block_size1, self.block_size2 = block_size1, block_size2
        self.device = x1.device
        super().__init__(x1.dtype, (x1.shape[0], x2.shape[0]))

    def hip_op(self):
        fam = getattr(self.fn, "ck_family", None)
        x1, x2 = self.x1, self.x2
        same = x2 is x1 or (x1.shape == x2.shape and x1.data_ptr() == x2.data_ptr() and x1.stride() == x2.stride())
        if fam is None or not same or not (x1.is_cuda and x1.dtype == torch.float32 and x1.dim() == 2
                                           and x1.shape[1] <= 8 and x1.stride(1) == 1):
            return None
        op = _op(N.CK_OP_RBF_F32, x1.shape[0])
        op.a, op.lda, op.dim, op.aux_f, op.rank = x1.data_ptr(), x1.stride(0), x1.shape[1], self.fn.lengthscale, fam
        return op, [x1]

    def _matmat(self, V):
        if self._use_hip(V):
            return self._hip_apply(V)
        out = torch.zeros((self.shape[0], V.shape[1]), dtype=V.dtype, device=V.device)
        for i0 in range(0, self.shape[0], self.block_size1):
            blk = slice(i0, i0 + self.block_size1)
            acc = torch.zeros_like(out[blk])
            for j0 in range(0, self.shape[1], self.block_size2):
                blk2 = slice(j0, j0 + self.block_size2)
                acc += self.fn(self.x1[blk], self.x2[blk2]) @ V[blk2]
            out[blk] = acc
        return out


class RBFKernel(LinearOperator):
    """exp(-|x_i - x_j|^2 / (2 l^2)) on points x (n x d, d <= 8), never stored.

    The engine builds each entry inside the matvec kernel (CK_OP_RBF_F32)."""

    def __init__(self, points, lengthscale):
        self.points = points.contiguous()
        self.lengthscale = float(lengthscale)
        self.device = points.device
        super().__init__(points.dtype, (points.shape[0], points.shape[0]), annotations={ann.PSD})

    def hip_op(self):
        P = self.points
        if not (P.is_cuda and P.dtype == torch.float32 and P.shape[1] <= 8):
            return None
        op = _op(N.CK_OP_RBF_F32, P.shape[0])
        op.a, op.lda, op.dim, op.aux_f = P.data_ptr(), P.stride(0), P.shape[1], self.lengthscale
        return op, [P]

    def _matmat(self, V):
        if self._use_hip(V):
            return self._hip_apply(V)
        return self.to_dense().to(V.dtype) @ V

    def to_dense(self):
        d2 = torch.cdist(self.points, self.points) ** 2
        return torch.exp(-d2 / (2 * self.lengthscale**2))


class WoodburyInverse(LinearOperator):
    """(s I + F F^T)^{-1} applied by the Woodbury identity; F is (n, k), k <= 128.

    The core (s I + F^T F)^{-1} is kept in FP64 and the engine applies it with FP64
    arithmetic (see CK_OP_WOODBURY_INV)."""

    def __init__(self, F, s, core_inv=None):
        self.F = F.contiguous()
        self.s = float(s)
        k = F.shape[1]
        if core_inv is not None:
            self.core_inv = core_inv.to(torch.float64).contiguous()
        elif self.F.is_cuda and self.F.dtype == torch.float32 and k <= 64:
            from .linalg.learned import woodbury_core
            self.core_inv = woodbury_core(self.F, self.s)[0]
        else:
            core = self.s * torch.eye(k, dtype=torch.float64, device=F.device) + (F.T.double() @ F.double())
            self.core_inv = torch.linalg.inv(core).contiguous()
        self.device = F.device
        super().__init__(F.dtype, (F.shape[0], F.shape[0]), annotations={ann.PSD})

    def hip_op(self):
        if not (self.F.is_cuda and self.F.dtype == torch.float32 and self.F.shape[1] <= 128):
            return None
        op = _op(N.CK_OP_WOODBURY_INV, self.F.shape[0], scale=1.0 / self.s)
        op.a, op.lda, op.aux, op.rank = self.F.data_ptr(), self.F.stride(0), self.core_inv.data_ptr(), self.F.shape[1]
        return op, [self.F, self.core_inv]

    def _matmat(self, V):
        if self._use_hip(V):
            return self._hip_apply(V)
        Fd = self.F.double()
        t = Fd.T @ V.double()
        return ((V.double() - Fd @ (self.core_inv @ t)) / self.s).to(V.dtype)
