"""The closed set of built-in operators.

The reference compiles user-written CubeCL traits into each kernel:
``ReductionOp`` (massively/src/core/reduce.rs:35-66), ``BinaryPredicateOp``
(core/ordering/mod.rs:29-65) and ``PredicateOp`` (core/predicate.rs:40-52).
This engine serves the built-in members of those families:

* reductions ``Add`` (integer add wraps modulo 2^bits), ``Min``, ``Max``,
  ``Xor`` (integers only); ``Min(a, b) = b if b < a else a``;
* key comparisons ``Less`` and ``Greater``;
* predicates ``NonZero``, ``Zero``, thresholds ``Lt/Le/Gt/Ge(t)`` and bitmask
  tests ``BitsAny(mask)`` / ``BitsEq(mask, value)`` on the raw element bits.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class ReductionOp:
    code: int = -1
    name: str = ""

    def __repr__(self):
        return self.name


class _Add(ReductionOp):
    code, name = 0, "Add"


class _Min(ReductionOp):
    code, name = 1, "Min"


class _Max(ReductionOp):
    code, name = 2, "Max"


class _Xor(ReductionOp):
    code, name = 3, "Xor"


Add, Min, Max, Xor = _Add(), _Min(), _Max(), _Xor()


class BinaryPredicateOp:
    code: int = -1
    name: str = ""

    def __repr__(self):
        return self.name


class _Less(BinaryPredicateOp):
    code, name = 0, "Less"


class _Greater(BinaryPredicateOp):
    code, name = 1, "Greater"


class _Equal(BinaryPredicateOp):
    """The built-in Equal of the by-key family and seg::Unique: numeric ==
    in the key dtype (-0.0 == +0.0; a NaN equals nothing)."""

    code, name = -1, "Equal"


Less, Greater, Equal = _Less(), _Greater(), _Equal()


@dataclass(frozen=True)
class PredicateOp:
    kind: int
    a: object = 0
    b: object = 0

    def c_struct(self, dtype: np.dtype) -> L.MsvPredicate:
        return L.MsvPredicate(self.kind, 0, _bits(self.a, dtype), _bits(self.b, dtype))


def _bits(v, dtype: np.dtype) -> int:
    """Bit pattern of `v` as an element of `dtype`, zero-extended to 64 bits."""
    dtype = np.dtype(dtype)
    arr = np.array([v], dtype=dtype)
    return int(arr.view(np.uint32 if dtype.itemsize == 4 else np.uint64)[0])


NonZero = PredicateOp(0)
Zero = PredicateOp(1)


def Lt(t):
    return PredicateOp(2, t)


def Le(t):
    return PredicateOp(3, t)


def Gt(t):
    return PredicateOp(4, t)


def Ge(t):
    return PredicateOp(5, t)


def BitsAny(mask):
    return PredicateOp(6, mask)


def BitsEq(mask, value):
    return PredicateOp(7, mask, value)


def host_scalar(value, dtype) -> C.Array:
    """A one-element host buffer holding `value` as `dtype` (init arguments)."""
    arr = np.array([value], dtype=np.dtype(dtype))
    return arr


class _UserOp:
    """A user-defined operator compiled at run time (msv_user_op_create): one
    compiled form per (executor, dtype), made on first use."""

    kind = 0

    def __init__(self, expr: str, identity: str | None = None):
        self.expr = expr
        self.identity = identity
        self._compiled = {}

    def handle(self, exec, dtype):
        import ctypes as C

        from . import _lib as L
        from .error import check

        key = (id(exec), np.dtype(dtype).str)
        h = self._compiled.get(key)
        if h is None:
            h = C.c_void_p()
            check(L.lib().msv_user_op_create(
                exec.handle, self.kind, L.DTYPE_OF_NP[np.dtype(dtype)], self.expr.encode(),
                None if self.identity is None else self.identity.encode(), C.byref(h)))
            self._compiled[key] = (exec, h)
            return h
        return h[1]

    def __del__(self):
        try:
            from . import _lib as L

            for _, h in self._compiled.values():
                L.lib().msv_user_op_destroy(h)
            self._compiled.clear()
        except Exception:
            pass


class UserReductionOp(_UserOp, ReductionOp):
    """A reduction given as a C++ expression of `a` and `b` over the element
    type, with an identity expression, e.g. UserReductionOp("a > b ? a : b",
    "0"). It must be associative; the scans and the reduce combine in index
    order, so it need not be commutative."""

    kind = 1
    name = "UserReductionOp"


class UserPredicate(_UserOp):
    """A predicate given as a C++ expression of `x` (nonzero keeps it),
    applied by ``vector.copy_where_if(exec, x, x, pred)``."""

    kind = 2

    def __init__(self, expr: str):
        _UserOp.__init__(self, expr, None)


class UserLess(_UserOp, BinaryPredicateOp):
    """A BinaryPredicateOp used as Less (massively/src/core/ordering/mod.rs:53-56):
    a C++ expression of `a` and `b` over the key type, nonzero when a comes
    strictly before b, e.g. UserLess("fabs(a) > fabs(b)") (descending by
    magnitude). It must be a strict weak ordering. Accepted by sort,
    sort_by_key, merge, merge_by_key, the set operations and seg.Sort; every
    one of them is stable under it."""

    kind = 4
    name = "UserLess"

    def __init__(self, expr: str):
        _UserOp.__init__(self, expr, None)


class UserEqual(_UserOp, BinaryPredicateOp):
    """A BinaryPredicateOp used as Equal: a C++ expression of `a` and `b` over
    the key type, nonzero when they are equal keys, e.g. UserEqual("a / 10 ==
    b / 10"). Accepted by the by-key family and seg.Unique; a run is a maximal
    range of adjacent keys with equal(k[i-1], k[i])."""

    kind = 5
    name = "UserEqual"

    def __init__(self, expr: str):
        _UserOp.__init__(self, expr, None)


class UnaryOp:
    """``op::UnaryOp`` (massively/src/core/op.rs:33-60): a C++ expression of
    `x` (the input element; `i` is its index) giving the output element, of
    dtype `output` (the input's dtype when None). Used by ``lazy.map`` and
    ``vector.map``; compiled into the consuming kernel at run time."""

    def __init__(self, expr: str, output=None):
        self.expr = expr
        self.output = None if output is None else np.dtype(output)

    def __repr__(self):
        return f"UnaryOp({self.expr!r})"


# op::Identity (massively/src/core/op.rs:86-96).
Identity = UnaryOp("x")
