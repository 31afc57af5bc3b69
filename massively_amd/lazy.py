"""``massively::lazy`` (massively/src/api/lazy.rs): inputs that are computed
inside the kernel that consumes them.

A lazy input is a tree of the reference's constructors: ``constant``,
``counting``, ``stride`` (each with ``.take(len)``), ``map``, ``permute``,
``reverse``, ``repeat_each``, ``tile``, ``adjacent_difference``, ``identity``,
``with_table`` (through ``map``) and ``.slice``, whose leaves are device vectors or slices. The algorithms that
fuse it (``vector.reduce``, ``inclusive_scan``, ``exclusive_scan``,
``copy_where_if``, ``map``, ``adjacent_difference``) lower the tree to C++
source giving element ``i`` (``msv_read``) over at most 13 device slots and 16
run-time parameters, the layout of the reference's read-expression ABI
(core/eval.rs:21-107), and the library compiles that source into its own
kernels once per shape (``msv_lazy_op_create``; the compiled form is kept by
the executor). Element values are computed, never stored. The other
algorithms take a lazy input by writing it out first (``materialize``).

Lengths follow the reference: map / identity / reverse / adjacent_difference
keep the input's length, permute has the indices' length, repeat_each and
tile multiply it (``LengthTooLarge`` past 2^32 - 1). Out-of-range permute
indices read 0 (as ``gather`` does here); the reference leaves them undefined.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .error import LengthMismatch, LengthTooLarge, check
from .op import ReductionOp, UnaryOp, UserReductionOp
from .runtime import DeviceSlice, DeviceVec, Executor, as_slice

_CTYPE = {
    np.dtype(np.uint32): "uint32_t",
    np.dtype(np.int32): "int32_t",
    np.dtype(np.float32): "float",
    np.dtype(np.uint64): "uint64_t",
    np.dtype(np.int64): "int64_t",
    np.dtype(np.float64): "double",
}
_U32 = np.dtype(np.uint32)
_MAX_LEN = 0xFFFFFFFF


def _bits_u64(value, dtype: np.dtype) -> int:
    a = np.array([value], dtype=dtype)
    return int(a.view(np.uint32 if a.itemsize == 4 else np.uint64)[0])


def _checked_len(n: int) -> int:
    if n > _MAX_LEN:
        raise LengthTooLarge(n)
    return n


class _Gen:
    """Collects the slots, parameters and device functions of one tree."""

    def __init__(self):
        self.slots: list[DeviceSlice] = []
        self.params: list[int] = []
        self.lines: list[str] = []
        self.memo: dict[int, str] = {}
        self.memo_v: dict[int, str] = {}
        self.info: dict[int, tuple] = {}  # slot / parameter indices of a node's scalar form

    def slot(self, s: DeviceSlice) -> int:
        if len(self.slots) == L.LAZY_SLOTS:
            raise ValueError(f"a lazy input reads at most {L.LAZY_SLOTS} device slices")
        self.slots.append(s)
        return len(self.slots) - 1

    def param(self, v: int) -> int:
        if len(self.params) == L.LAZY_PARAMS:
            raise ValueError(f"a lazy input takes at most {L.LAZY_PARAMS} parameters")
        self.params.append(int(v) & 0xFFFFFFFFFFFFFFFF)
        return len(self.params) - 1

    def fn(self, node: "Lazy", body: str) -> str:
        name = f"lz{len(self.memo)}"
        self.memo[id(node)] = name
        self.lines.append(
            f"__device__ inline {_CTYPE[node.dtype]} {name}(const LazyArgs& la, size_t i) {{ {body} }}")
        return name

    def emit(self, node: "Lazy") -> str:
        got = self.memo.get(id(node))
        return got if got is not None else node._emit(self)

    def fn_v(self, node: "Lazy", body: str, ok: str = "false") -> str:
        """The K-element form (o[k] = element i0 + k; F: unchecked slot reads,
        valid where the node's tile_ok holds) and the node's tile_ok(lo, hi):
        whether its slots can be read unchecked over [lo, hi)."""
        name = self.memo[id(node)] + "_v"
        self.memo_v[id(node)] = name
        self.lines.append(
            f"template <int K, bool F>\n__device__ inline void {name}(const LazyArgs& la, size_t i0, "
            f"{_CTYPE[node.dtype]} (&o)[K]) {{ {body} }}")
        self.lines.append(
            f"__device__ inline bool {name}_ok(const LazyArgs& la, size_t lo, size_t hi) "
            f"{{ (void)la; (void)lo; (void)hi; return {ok}; }}")
        return name

    def emit_v(self, node: "Lazy") -> str:
        got = self.memo_v.get(id(node))
        if got is not None:
            return got
        self.emit(node)  # the scalar form first: it assigns the slots and parameters
        return node._emit_v(self)


class Lazy:
    """A lazy input of ``dtype`` and length ``len(self)``."""

    dtype: np.dtype

    def length(self) -> int:
        raise NotImplementedError

    def __len__(self) -> int:
        return self.length()

    def _emit(self, g: _Gen) -> str:
        raise NotImplementedError

    def _emit_v(self, g: _Gen) -> str:
        """Default K-element form: the scalar form per element (nodes whose
        elements come from scattered indices)."""
        sc = g.memo[id(self)]
        return g.fn_v(self, f"for (int k = 0; k < K; ++k) o[k] = {sc}(la, i0 + k);")

    def slice(self, start: int = 0, end: int | None = None) -> "Lazy":
        """The lazy sub-range [start, end) (``MIter::slice``)."""
        n = self.length()
        end = n if end is None else end
        if not (0 <= start <= end <= n):
            raise ValueError(f"bad sub-slice {start}..{end} of length {n}")
        return _Slice(self, start, end - start)

    def lower(self) -> tuple[str, L.MsvLazyArgs]:
        """The read source (``msv_read``) and the slots / parameters of a call."""
        g = _Gen()
        root = g.emit(self)
        root_v = g.emit_v(self)
        src = "\n".join(g.lines)
        t = _CTYPE[self.dtype]
        src += (f"\n__device__ inline UT msv_read(const LazyArgs& la, size_t i) {{ return (UT){root}(la, i); }}\n"
                f"template <int K>\n__device__ inline void msv_read_v(const LazyArgs& la, size_t i, UT (&o)[K]) "
                f"{{ {t} t[K]; {root_v}<K, false>(la, i, t); for (int k = 0; k < K; ++k) o[k] = (UT)t[k]; }}\n"
                f"template <int K>\n__device__ inline void msv_read_vf(const LazyArgs& la, size_t i, UT (&o)[K]) "
                f"{{ {t} t[K]; {root_v}<K, true>(la, i, t); for (int k = 0; k < K; ++k) o[k] = (UT)t[k]; }}\n"
                f"__device__ inline bool msv_tile_ok(const LazyArgs& la, size_t lo, size_t hi) "
                f"{{ return {root_v}_ok(la, lo, hi); }}\n")
        args = L.MsvLazyArgs()
        for k, s in enumerate(g.slots):
            args.slots[k] = s._c()
        args.nslots = len(g.slots)
        args.len = self.length()
        for k, v in enumerate(g.params):
            args.params[k] = v
        args._keep = list(g.slots)  # the slices' vectors stay alive during the call
        return src, args

    def materialize(self, exec: Executor) -> DeviceVec:
        """The lazy input written out to a new vector (``vector::map`` with
        ``op::Identity``)."""
        return map_into_vec(exec, self)


class _Slot(Lazy):
    def __init__(self, s: DeviceSlice):
        self.s = s
        self.dtype = s.dtype
        self.n = s.vec.len() - s.offset if s.length == L.WHOLE else s.length

    def length(self) -> int:
        return self.n

    def _emit(self, g: _Gen) -> str:
        k = g.slot(self.s)
        g.info[id(self)] = (k,)
        return g.fn(self, f"return MSV_SLOT({k}, {_CTYPE[self.dtype]}, i);")

    def _emit_v(self, g: _Gen) -> str:
        (k,) = g.info[id(self)]
        t = _CTYPE[self.dtype]
        return g.fn_v(self, f"if constexpr (F) slot_read_fast<K, {t}>(la.p[{k}], i0, o); "
                            f"else slot_read<K, {t}>(la.p[{k}], la.n[{k}], i0, o);",
                      ok=f"hi <= la.n[{k}] && ((uintptr_t)((const {t}*)la.p[{k}] + lo) & 15u) == 0")


def _node(x) -> Lazy:
    if isinstance(x, Lazy):
        return x
    if isinstance(x, (DeviceVec, DeviceSlice)):
        return _Slot(as_slice(x))
    if isinstance(x, (Constant, Counting, Stride)):
        raise TypeError("an unbounded lazy stream needs .take(len) before use")
    raise TypeError(f"expected a lazy input, DeviceVec or DeviceSlice, got {type(x).__name__}")


class _Slice(Lazy):
    def __init__(self, inner: Lazy, start: int, count: int):
        self.inner, self.start, self.count = inner, start, count
        self.dtype = inner.dtype

    def length(self) -> int:
        return self.count

    def _emit(self, g: _Gen) -> str:
        c = g.emit(self.inner)
        j = g.param(self.start)
        g.info[id(self)] = (j,)
        return g.fn(self, f"return {c}(la, i + la.q[{j}]);")

    def _emit_v(self, g: _Gen) -> str:
        (j,) = g.info[id(self)]
        c = g.emit_v(self.inner)
        return g.fn_v(self, f"{c}<K, F>(la, i0 + la.q[{j}], o);",
                      ok=f"{c}_ok(la, lo + la.q[{j}], hi + la.q[{j}])")


# ---- unbounded sources (lazy.rs:566-720) ------------------------------------

class Constant:
    """``lazy::constant(value)``; ``dtype`` defaults to u32 for Python ints and
    f32 for Python floats (numpy scalars keep theirs)."""

    def __init__(self, value, dtype=None):
        if dtype is None:
            if isinstance(value, np.generic):
                dtype = value.dtype
            elif isinstance(value, float):
                dtype = np.float32
            else:
                dtype = np.uint32
        self.dtype = np.dtype(dtype)
        self.value = value

    def take(self, n: int) -> Lazy:
        return _Constant(self, _checked_len(int(n)))


class _Constant(Lazy):
    def __init__(self, c: Constant, n: int):
        self.c, self.n, self.dtype = c, n, c.dtype

    def length(self) -> int:
        return self.n

    def _emit(self, g: _Gen) -> str:
        j = g.param(_bits_u64(self.c.value, self.dtype))
        g.info[id(self)] = (j,)
        u = "uint32_t" if self.dtype.itemsize == 4 else "uint64_t"
        return g.fn(self, f"return bit_cast<{_CTYPE[self.dtype]}>(({u})la.q[{j}]);")

    def _emit_v(self, g: _Gen) -> str:
        sc = g.memo[id(self)]
        return g.fn_v(self, f"const {_CTYPE[self.dtype]} v = {sc}(la, 0); "
                            f"for (int k = 0; k < K; ++k) o[k] = v;", ok="true")


class Counting:
    """``lazy::counting(start)``: start, start + 1, ... (u32)."""

    def __init__(self, start: int):
        self.start = int(start)

    def take(self, n: int) -> Lazy:
        return _Counting(self.start, _checked_len(int(n)))


class _Counting(Lazy):
    dtype = _U32

    def __init__(self, start: int, n: int):
        self.start, self.n = start, n

    def length(self) -> int:
        return self.n

    def _emit(self, g: _Gen) -> str:
        j = g.param(self.start)
        g.info[id(self)] = (j,)
        return g.fn(self, f"return (uint32_t)(la.q[{j}] + i);")

    def _emit_v(self, g: _Gen) -> str:
        (j,) = g.info[id(self)]
        return g.fn_v(self, f"for (int k = 0; k < K; ++k) o[k] = (uint32_t)(la.q[{j}] + i0 + k);",
                      ok="true")


class Stride:
    """``lazy::stride(start, step)``: start + i * step (u32, wrapping)."""

    def __init__(self, start: int, step: int):
        self.start, self.step = int(start), int(step)

    def take(self, n: int) -> Lazy:
        return _Stride(self.start, self.step, _checked_len(int(n)))


class _Stride(Lazy):
    dtype = _U32

    def __init__(self, start: int, step: int, n: int):
        self.start, self.step, self.n = start, step, n

    def length(self) -> int:
        return self.n

    def _emit(self, g: _Gen) -> str:
        a, b = g.param(self.start), g.param(self.step)
        g.info[id(self)] = (a, b)
        return g.fn(self, f"return (uint32_t)(la.q[{a}] + (uint64_t)i * la.q[{b}]);")

    def _emit_v(self, g: _Gen) -> str:
        a, b = g.info[id(self)]
        return g.fn_v(self, f"for (int k = 0; k < K; ++k) "
                            f"o[k] = (uint32_t)(la.q[{a}] + (uint64_t)(i0 + k) * la.q[{b}]);", ok="true")


def constant(value, dtype=None) -> Constant:
    return Constant(value, dtype)


def counting(start: int) -> Counting:
    return Counting(start)


def stride(start: int, step: int) -> Stride:
    return Stride(start, step)


# ---- adaptors (lazy.rs:114-560, 745-922) ------------------------------------

class _Map(Lazy):
    def __init__(self, inner: Lazy, op: UnaryOp):
        self.inner, self.op = inner, op
        self.dtype = op.output if op.output is not None else inner.dtype

    def length(self) -> int:
        return self.inner.length()

    def _emit(self, g: _Gen) -> str:
        c = g.emit(self.inner)
        ti, to = _CTYPE[self.inner.dtype], _CTYPE[self.dtype]
        return g.fn(self, f"const {ti} x = {c}(la, i); return ({to})({self.op.expr});")

    def _emit_v(self, g: _Gen) -> str:
        c = g.emit_v(self.inner)
        ti, to = _CTYPE[self.inner.dtype], _CTYPE[self.dtype]
        return g.fn_v(self, f"{ti} t[K]; {c}<K, F>(la, i0, t); for (int k = 0; k < K; ++k) "
                            f"{{ const size_t i = i0 + k; (void)i; const {ti} x = t[k]; "
                            f"o[k] = ({to})({self.op.expr}); }}", ok=f"{c}_ok(la, lo, hi)")


class _ZipMap(Lazy):
    """map over zipped inputs (``lazy::map(zip2(a, b), op)``, api/iter.rs:1169-1300):
    the expression reads x0, x1, ... (the elements of each input at i)."""

    def __init__(self, inners: list, op: UnaryOp):
        if not 1 <= len(inners) <= 12:
            raise ValueError("zip takes 1 to 12 inputs")
        lens = {n.length() for n in inners}
        if len(lens) != 1:
            a, b = sorted(lens)[:2]
            raise LengthMismatch(a, b)
        if op.output is None:
            raise TypeError("a map over zipped inputs needs the output dtype")
        self.inners, self.op, self.dtype = inners, op, op.output

    def length(self) -> int:
        return self.inners[0].length()

    def _emit(self, g: _Gen) -> str:
        names = [g.emit(n) for n in self.inners]
        reads = " ".join(f"const {_CTYPE[n.dtype]} x{k} = {c}(la, i);"
                         for k, (n, c) in enumerate(zip(self.inners, names)))
        return g.fn(self, f"{reads} return ({_CTYPE[self.dtype]})({self.op.expr});")

    def _emit_v(self, g: _Gen) -> str:
        names = [g.emit_v(n) for n in self.inners]
        arrays = " ".join(f"{_CTYPE[n.dtype]} t{k}[K]; {c}<K, F>(la, i0, t{k});"
                          for k, (n, c) in enumerate(zip(self.inners, names)))
        reads = " ".join(f"const {_CTYPE[n.dtype]} x{k} = t{k}[j];" for k, n in enumerate(self.inners))
        ok = " && ".join(f"{c}_ok(la, lo, hi)" for c in names)
        return g.fn_v(self, f"{arrays} for (int j = 0; j < K; ++j) {{ const size_t i = i0 + j; (void)i; "
                            f"{reads} o[j] = ({_CTYPE[self.dtype]})({self.op.expr}); }}", ok=ok)


class WithTable:
    """``lazy::with_table(contexts, table)`` (lazy.rs:20-110, 720-747): each
    row is a context element plus a view of the whole table. Consumed by
    ``map``, whose expression reads ``x`` (the context element) and
    ``table.at(k)`` / ``table.len()`` (``at`` reads 0 past the table's end)."""

    def __init__(self, contexts, table):
        if not isinstance(table, (DeviceVec, DeviceSlice)):
            raise TypeError("the table of with_table is a device vector or slice")
        self.contexts = _node(contexts)
        self.table = _Slot(as_slice(table))


def with_table(contexts, table) -> WithTable:
    return WithTable(contexts, table)


_TABLE_VIEW = ("template <typename D> struct TableView { const D* p; uint32_t n; "
               "__device__ inline D at(uint64_t k) const { return k < n ? p[k] : (D)0; } "
               "__device__ inline uint32_t len() const { return n; } };")


class _WithTableMap(Lazy):
    def __init__(self, wt: WithTable, op: UnaryOp):
        self.wt, self.op = wt, op
        self.dtype = op.output if op.output is not None else wt.contexts.dtype

    def length(self) -> int:
        return self.wt.contexts.length()

    def _emit(self, g: _Gen) -> str:
        c = g.emit(self.wt.contexts)
        k = g.slot(self.wt.table.s)
        if _TABLE_VIEW not in g.lines:
            g.lines.insert(0, _TABLE_VIEW)
        tc, tt, to = _CTYPE[self.wt.contexts.dtype], _CTYPE[self.wt.table.dtype], _CTYPE[self.dtype]
        return g.fn(self, f"const {tc} x = {c}(la, i); "
                          f"const TableView<{tt}> table{{(const {tt}*)la.p[{k}], la.n[{k}]}}; "
                          f"return ({to})({self.op.expr});")


class _Permute(Lazy):
    def __init__(self, values: Lazy, indices: Lazy):
        if indices.dtype != _U32:
            raise TypeError("permute indices must be u32 (MIndex)")
        self.values, self.indices = values, indices
        self.dtype = values.dtype

    def length(self) -> int:
        return self.indices.length()

    def _emit(self, g: _Gen) -> str:
        v, k = g.emit(self.values), g.emit(self.indices)
        j = g.param(self.values.length())
        t = _CTYPE[self.dtype]
        return g.fn(self, f"const uint32_t k = {k}(la, i); "
                          f"return k < (uint32_t)la.q[{j}] ? {v}(la, k) : ({t})0;")


class _Reverse(Lazy):
    def __init__(self, inner: Lazy):
        self.inner, self.dtype = inner, inner.dtype

    def length(self) -> int:
        return self.inner.length()

    def _emit(self, g: _Gen) -> str:
        c = g.emit(self.inner)
        j = g.param(self.inner.length())
        return g.fn(self, f"return {c}(la, (size_t)la.q[{j}] - 1 - i);")


class _RepeatEach(Lazy):
    def __init__(self, inner: Lazy, repeats: int):
        self.inner, self.r, self.dtype = inner, int(repeats), inner.dtype
        _checked_len(inner.length() * self.r)

    def length(self) -> int:
        return self.inner.length() * self.r

    def _emit(self, g: _Gen) -> str:
        c = g.emit(self.inner)
        j = g.param(max(self.r, 1))
        return g.fn(self, f"return {c}(la, i / la.q[{j}]);")


class _Tile(Lazy):
    def __init__(self, inner: Lazy, repeats: int):
        self.inner, self.r, self.dtype = inner, int(repeats), inner.dtype
        _checked_len(inner.length() * self.r)

    def length(self) -> int:
        return self.inner.length() * self.r

    def _emit(self, g: _Gen) -> str:
        c = g.emit(self.inner)
        j = g.param(max(self.inner.length(), 1))
        return g.fn(self, f"return {c}(la, i % la.q[{j}]);")


def _op_apply(g: _Gen, op, dtype: np.dtype) -> str:
    """The device call that combines two values with a reduction operator."""
    t = _CTYPE[dtype]
    if isinstance(op, UserReductionOp):
        name = f"lzop{len(g.lines)}"
        g.lines.append(f"struct {name} {{ __device__ static inline {t} apply({t} a, {t} b) "
                       f"{{ return ({t})({op.expr}); }} }};")
        return f"{name}::apply"
    if isinstance(op, ReductionOp) and op.code >= 0:
        if op.code == 3 and dtype.kind == "f":
            raise TypeError("Xor is defined for integer dtypes only")
        return f"Combine<{t}, {op.code}>::apply"
    raise TypeError(f"expected a ReductionOp, got {op!r}")


class _AdjacentDifference(Lazy):
    def __init__(self, inner: Lazy, op):
        self.inner, self.op, self.dtype = inner, op, inner.dtype

    def length(self) -> int:
        return self.inner.length()

    def _emit(self, g: _Gen) -> str:
        c = g.emit(self.inner)
        f = _op_apply(g, self.op, self.dtype)
        t = _CTYPE[self.dtype]
        return g.fn(self, f"const {t} cur = {c}(la, i); "
                          f"return i == 0 ? cur : {f}({c}(la, i - 1), cur);")


def map(input, op: UnaryOp) -> Lazy:  # noqa: A001 - the reference's name
    """``lazy::map(input, op)``: op(x) for each element. A tuple or list of
    equally long inputs is zipped (``zip2`` .. ``zip12``): the expression reads
    x0, x1, ... and ``op.output`` names the result dtype."""
    if isinstance(input, (tuple, list)):
        return _ZipMap([_node(x) for x in input], op)
    if isinstance(input, WithTable):
        return _WithTableMap(input, op)
    return _Map(_node(input), op)


def _rows(values, f):
    """An adaptor of zipped rows (a tuple or list of equally long columns)
    is the tuple of the adapted columns; a single input is adapted as is."""
    if isinstance(values, (tuple, list)):
        cols = [_node(v) for v in values]
        lens = {c.length() for c in cols}
        if len(lens) > 1:
            a, b = sorted(lens)[:2]
            raise LengthMismatch(a, b)
        return tuple(f(c) for c in cols)
    return f(_node(values))


def identity(input):
    """``lazy::identity(input)`` (``map`` with ``op::Identity``)."""
    return _rows(input, lambda c: _Map(c, UnaryOp("x")))


def permute(values, indices):
    """``lazy::permute(values, indices)``: values[indices[i]] (zipped values:
    every column through the same indices)."""
    idx = _node(indices)
    return _rows(values, lambda c: _Permute(c, idx))


def reverse(values):
    """``lazy::reverse(values)``: values[len - 1 - i]."""
    return _rows(values, _Reverse)


def repeat_each(values, repeats: int):
    """``lazy::repeat_each(values, r)``: each element r times in a row."""
    return _rows(values, lambda c: _RepeatEach(c, repeats))


def tile(values, repeats: int):
    """``lazy::tile(values, r)``: the whole input r times."""
    return _rows(values, lambda c: _Tile(c, repeats))


def slice_rows(values, start: int = 0, end: int | None = None):
    """``MIter::slice`` of a lazy input or of zipped lazy rows."""
    return _rows(values, lambda c: c.slice(start, end))


def adjacent_difference(input, op) -> Lazy:
    """``lazy::adjacent_difference(input, op)``: x[0], then op(x[i-1], x[i])."""
    return _AdjacentDifference(_node(input), op)


# ---- consumers --------------------------------------------------------------

_KIND = {"reduction": 1, "predicate": 2, "map": 3}


def compiled(exec: Executor, kind: str, dtype: np.dtype, src: str, op=None) -> C.c_void_p:
    """The executor's compiled operator for (kind, dtype, read source, op),
    made on first use (msv_lazy_op_create)."""
    builtin, expr, ident = -1, None, None
    if kind == "reduction":
        if isinstance(op, UserReductionOp):
            expr, ident = op.expr, op.identity
        else:
            builtin = op.code
    elif kind == "predicate" and op is not None:
        expr = op.expr
    key = (kind, np.dtype(dtype).str, src, builtin, expr, ident)
    h = exec._jit_ops.get(key)
    if h is None:
        h = C.c_void_p()
        check(L.lib().msv_lazy_op_create(
            exec.handle, _KIND[kind], L.DTYPE_OF_NP[np.dtype(dtype)], src.encode(), builtin,
            None if expr is None else expr.encode(), None if ident is None else ident.encode(),
            C.byref(h)))
        exec._jit_ops[key] = h
    return h


def check_compiles(kind: str, dtype, src: str | None, op=None) -> None:
    """Raises the error msv_lazy_op_create would give, without a device."""
    builtin, expr, ident = -1, None, None
    if kind == "reduction":
        if isinstance(op, UserReductionOp):
            expr, ident = op.expr, op.identity
        else:
            builtin = op.code
    elif kind == "predicate" and op is not None:
        expr = op.expr
    check(L.lib().msv_lazy_op_check(
        _KIND[kind], L.DTYPE_OF_NP[np.dtype(dtype)], None if src is None else src.encode(), builtin,
        None if expr is None else expr.encode(), None if ident is None else ident.encode()))


def map_into_vec(exec: Executor, lazy: Lazy) -> DeviceVec:
    if isinstance(lazy, tuple):  # zipped rows: one vector per column (into_columns)
        return tuple(map_into_vec(exec, c) for c in lazy)
    src, args = lazy.lower()
    h = C.c_void_p()
    check(L.lib().msv_map_lazy(exec.handle, C.byref(args),
                               compiled(exec, "map", lazy.dtype, src), C.byref(h)))
    return DeviceVec(exec, h)


def as_input(exec: Executor, x) -> DeviceSlice:
    """A device slice of `x`: lazy inputs are written out first."""
    if isinstance(x, Lazy):
        return map_into_vec(exec, x).slice()
    return as_slice(x)
