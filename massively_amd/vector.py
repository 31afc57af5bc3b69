"""``massively::vector`` algorithms on the gfx950 kernels of libmsv.

Signatures follow the reference's free functions: the executor first, then
device slices (or vectors), then the operator. Owned results are returned as
new :class:`DeviceVec` objects; ``reduce`` returns a host scalar.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .error import LengthMismatch, check
from .op import (BinaryPredicateOp, PredicateOp, ReductionOp, UnaryOp, UserEqual, UserLess,
                 UserPredicate, UserReductionOp)
from . import lazy
from .runtime import DeviceSlice, DeviceVec, Executor, as_slice

_lib = L.lib


def _in(exec: Executor, x) -> DeviceSlice:
    """A device slice of an input; a lazy input is written out first (the
    algorithms that fuse lazy inputs dispatch before this)."""
    return lazy.as_input(exec, x)


def _cols(x):
    """The columns of a zipped input (a tuple or list, the reference's zip2 ..
    zip12 rows), or None for a single column."""
    if isinstance(x, (tuple, list)):
        if not 1 <= len(x) <= 12:
            raise ValueError("a zipped input has 1 to 12 columns")
        return list(x)
    return None


def _col_len(exec: Executor, x) -> int:
    s = _in(exec, x) if not isinstance(x, lazy.Lazy) else None
    if s is None:
        return len(x)
    return s.vec.len() - s.offset if s.length == L.WHOLE else s.length


def _col_slices(exec: Executor, cols):
    """A C array of the columns' slices (lazy columns written out first)."""
    sl = [_in(exec, c)._c() for c in cols]
    return (L.MsvSlice * len(sl))(*sl)


def _columns_out(exec: Executor, outs) -> tuple:
    return tuple(_out(exec, C.c_void_p(h)) for h in outs)


def _same_lengths(exec: Executor, cols) -> int:
    lens = [_col_len(exec, c) for c in cols]
    for n in lens[1:]:
        if n != lens[0]:
            raise LengthMismatch(lens[0], n)
    return lens[0]


def _iota(exec: Executor, n: int, start: int = 0) -> DeviceVec:
    v = exec.alloc(np.uint32, n)
    if n:
        check(_lib().msv_iota_u32(exec.handle, start, v.slice()._c()))
    return v


def _out(exec: Executor, h: C.c_void_p) -> DeviceVec:
    return DeviceVec(exec, h)


def _scalar(value, dtype) -> np.ndarray:
    return np.array([value], dtype=np.dtype(dtype))


def inclusive_scan(exec: Executor, input, op: ReductionOp) -> DeviceVec:
    """massively/src/api/algorithm/scan.rs:96-113. A lazy input is computed
    inside the scan kernels."""
    if isinstance(input, lazy.Lazy):
        src, args = input.lower()
        h = C.c_void_p()
        check(_lib().msv_inclusive_scan_lazy(exec.handle, C.byref(args),
                                             lazy.compiled(exec, "reduction", input.dtype, src, op),
                                             C.byref(h)))
        return _out(exec, h)
    s = _in(exec, input)
    h = C.c_void_p()
    if isinstance(op, UserReductionOp):
        check(_lib().msv_inclusive_scan_user(exec.handle, s._c(), op.handle(exec, s.dtype), C.byref(h)))
        return _out(exec, h)
    check(_lib().msv_inclusive_scan(exec.handle, s._c(), op.code, C.byref(h)))
    return _out(exec, h)


def exclusive_scan(exec: Executor, input, init, op: ReductionOp) -> DeviceVec:
    """massively/src/api/algorithm/scan.rs:216-257: out[0] = init."""
    if isinstance(input, lazy.Lazy):
        src, args = input.lower()
        init_buf = _scalar(init, input.dtype)
        h = C.c_void_p()
        check(_lib().msv_exclusive_scan_lazy(exec.handle, C.byref(args),
                                             init_buf.ctypes.data_as(C.c_void_p),
                                             lazy.compiled(exec, "reduction", input.dtype, src, op),
                                             C.byref(h)))
        return _out(exec, h)
    s = _in(exec, input)
    init_buf = _scalar(init, s.dtype)
    h = C.c_void_p()
    if isinstance(op, UserReductionOp):
        check(_lib().msv_exclusive_scan_user(exec.handle, s._c(), init_buf.ctypes.data_as(C.c_void_p),
                                             op.handle(exec, s.dtype), C.byref(h)))
        return _out(exec, h)
    check(
        _lib().msv_exclusive_scan(
            exec.handle, s._c(), init_buf.ctypes.data_as(C.c_void_p), op.code, C.byref(h)
        )
    )
    return _out(exec, h)


def reduce(exec: Executor, input, init, op: ReductionOp):
    """massively/src/api/algorithm/reduce.rs:32-50: op(init, fold(input)). A
    lazy input is computed inside the reduction kernel."""
    if isinstance(input, lazy.Lazy):
        src, args = input.lower()
        init_buf = _scalar(init, input.dtype)
        out = np.zeros(1, dtype=input.dtype)
        check(_lib().msv_reduce_lazy(exec.handle, C.byref(args), init_buf.ctypes.data_as(C.c_void_p),
                                     lazy.compiled(exec, "reduction", input.dtype, src, op),
                                     out.ctypes.data_as(C.c_void_p)))
        return out[0]
    s = _in(exec, input)
    init_buf = _scalar(init, s.dtype)
    out = np.zeros(1, dtype=s.dtype)
    if isinstance(op, UserReductionOp):
        check(_lib().msv_reduce_user(exec.handle, s._c(), init_buf.ctypes.data_as(C.c_void_p),
                                     op.handle(exec, s.dtype), out.ctypes.data_as(C.c_void_p)))
        return out[0]
    check(
        _lib().msv_reduce(
            exec.handle, s._c(), init_buf.ctypes.data_as(C.c_void_p), op.code,
            out.ctypes.data_as(C.c_void_p),
        )
    )
    return out[0]


def copy_where(exec: Executor, input, stencil) -> DeviceVec:
    """massively/src/api/algorithm/selection.rs:146-165 (device-resident length).
    A zipped input (tuple of columns) gives the tuple of compacted columns, from
    one launch that reads the stencil once (msv_copy_where_columns)."""
    cols = _cols(input)
    if cols is not None:
        _same_lengths(exec, cols)
        outs = (C.c_void_p * len(cols))()
        check(_lib().msv_copy_where_columns(exec.handle, _col_slices(exec, cols), len(cols),
                                            _in(exec, stencil)._c(), None, 0, outs))
        return _columns_out(exec, outs)
    h = C.c_void_p()
    check(_lib().msv_copy_where(exec.handle, _in(exec, input)._c(), _in(exec, stencil)._c(),
                                C.byref(h)))
    return _out(exec, h)


def remove_where(exec: Executor, input, stencil) -> DeviceVec:
    """massively/src/api/algorithm/selection.rs:203-222 (zipped inputs as copy_where)."""
    cols = _cols(input)
    if cols is not None:
        _same_lengths(exec, cols)
        outs = (C.c_void_p * len(cols))()
        check(_lib().msv_copy_where_columns(exec.handle, _col_slices(exec, cols), len(cols),
                                            _in(exec, stencil)._c(), None, 1, outs))
        return _columns_out(exec, outs)
    h = C.c_void_p()
    check(_lib().msv_remove_where(exec.handle, _in(exec, input)._c(), _in(exec, stencil)._c(),
                                  C.byref(h)))
    return _out(exec, h)


def copy_where_if(exec: Executor, input, stencil, pred: PredicateOp, invert: bool = False):
    """copy_where with ``lazy::map(stencil, pred)`` as the stencil, fused. A
    :class:`UserPredicate` is evaluated on the input itself (``stencil`` must be
    the input). A lazy input with ``stencil`` the same object is computed
    inside the compaction kernel, and the predicate (built-in or user) is
    applied to it. A zipped input (tuple of columns) with a separate stencil
    gives the tuple of compacted columns."""
    cols = _cols(input)
    if cols is not None:
        if isinstance(pred, UserPredicate):
            raise ValueError("a UserPredicate applies to a single input column")
        _same_lengths(exec, cols)
        st = _in(exec, stencil)
        p = pred.c_struct(st.dtype)
        outs = (C.c_void_p * len(cols))()
        check(_lib().msv_copy_where_columns(exec.handle, _col_slices(exec, cols), len(cols), st._c(),
                                            C.byref(p), 1 if invert else 0, outs))
        return _columns_out(exec, outs)
    if isinstance(input, lazy.Lazy) and stencil is input:
        if invert:
            raise ValueError("remove_where_if of a lazy input: use the negated predicate")
        src, args = input.lower()
        user = isinstance(pred, UserPredicate)
        h = C.c_void_p()
        p = None if user else pred.c_struct(input.dtype)
        check(_lib().msv_copy_where_lazy(exec.handle, C.byref(args),
                                         lazy.compiled(exec, "predicate", input.dtype, src,
                                                       pred if user else None),
                                         None if user else C.byref(p), C.byref(h)))
        return _out(exec, h)
    st = _in(exec, stencil)
    if isinstance(pred, UserPredicate):
        src = _in(exec, input)
        if invert or src.vec is not st.vec or src.offset != st.offset:
            raise ValueError("a UserPredicate is applied to the input itself, without invert")
        h = C.c_void_p()
        check(_lib().msv_copy_where_user(exec.handle, src._c(), pred.handle(exec, src.dtype), C.byref(h)))
        return _out(exec, h)
    p = pred.c_struct(st.dtype)
    h = C.c_void_p()
    check(_lib().msv_copy_where_pred(exec.handle, _in(exec, input)._c(), st._c(), C.byref(p),
                                     1 if invert else 0, C.byref(h)))
    return _out(exec, h)


def remove_where_if(exec: Executor, input, stencil, pred: PredicateOp):
    return copy_where_if(exec, input, stencil, pred, invert=True)


def partition(exec: Executor, input, pred: PredicateOp):
    """massively/src/api/algorithm/selection.rs:245-288: stable partition,
    passing items first. Returns (output, boundary)."""
    src = _in(exec, input)
    p = pred.c_struct(src.dtype)
    h = C.c_void_p()
    boundary = C.c_uint32()
    check(_lib().msv_partition(exec.handle, src._c(), C.byref(p), C.byref(h), C.byref(boundary)))
    return _out(exec, h), int(boundary.value)


def gather(exec: Executor, values, indices) -> DeviceVec:
    """massively/src/api/algorithm/indexed.rs:157-208: out[i] = values[indices[i]].
    Zipped values (tuple of columns) give the tuple of gathered columns, from one
    launch that reads the indices once (msv_gather_columns)."""
    cols = _cols(values)
    if cols is not None:
        _same_lengths(exec, cols)
        outs = (C.c_void_p * len(cols))()
        check(_lib().msv_gather_columns(exec.handle, _col_slices(exec, cols), len(cols),
                                        _in(exec, indices)._c(), outs))
        return _columns_out(exec, outs)
    h = C.c_void_p()
    check(_lib().msv_gather(exec.handle, _in(exec, values)._c(), _in(exec, indices)._c(),
                            C.byref(h)))
    return _out(exec, h)


def scatter(exec: Executor, values, indices, output) -> None:
    """massively/src/api/algorithm/scatter.rs:248-265: output[indices[i]] = values[i].
    Zipped values scatter into the zipped output in one launch that reads the
    indices once (msv_scatter_columns)."""
    cols = _cols(values)
    if cols is not None:
        outs = _cols(output)
        if outs is None or len(outs) != len(cols):
            raise ValueError("zipped values need a zipped output of as many columns")
        _same_lengths(exec, cols)
        _same_lengths(exec, outs)
        osl = [as_slice(o)._c() for o in outs]
        check(_lib().msv_scatter_columns(exec.handle, _col_slices(exec, cols), len(cols),
                                         _in(exec, indices)._c(), (L.MsvSlice * len(osl))(*osl)))
        return
    check(_lib().msv_scatter(exec.handle, _in(exec, values)._c(), _in(exec, indices)._c(),
                             as_slice(output)._c()))


def scatter_reduce(exec: Executor, values, indices, init, op: ReductionOp, output) -> None:
    """massively/src/api/algorithm/scatter.rs:337-420: output[d] = op(output[d],
    init (+) the values aimed at d); ``init`` must be op's identity."""
    v = _in(exec, values)
    i = np.array([init], dtype=v.dtype)
    check(_lib().msv_scatter_reduce(exec.handle, v._c(), _in(exec, indices)._c(),
                                    i.ctypes.data_as(C.c_void_p), op.code, as_slice(output)._c()))


def radix_sort_by_key(exec: Executor, keys, values, return_keys: bool = False):
    """massively/src/api/algorithm/by_key.rs:252-297: stable, integer/IEEE total order.

    Returns the permuted values (and the sorted keys when ``return_keys``).
    ``keys`` and ``values`` may also be tuples of columns (the reference's
    ``zip2``/``zip3`` keys and up to ``zip12`` values, core/radix.rs:499-531):
    keys sort lexicographically, leftmost column first, and the result is the
    tuple of permuted value columns.
    """
    if isinstance(keys, (tuple, list)) or isinstance(values, (tuple, list)):
        if return_keys:
            raise ValueError("return_keys is not available for zipped columns")
        kc = [_in(exec, k)._c() for k in (keys if isinstance(keys, (tuple, list)) else (keys,))]
        vc = [_in(exec, v)._c() for v in (values if isinstance(values, (tuple, list)) else (values,))]
        karr = (L.MsvSlice * len(kc))(*kc)
        varr = (L.MsvSlice * len(vc))(*vc)
        outs = (C.c_void_p * len(vc))()
        check(_lib().msv_radix_sort_by_key_columns(exec.handle, karr, len(kc), varr, len(vc), outs))
        res = tuple(_out(exec, C.c_void_p(h)) for h in outs)
        return res if isinstance(values, (tuple, list)) else res[0]
    hv, hk = C.c_void_p(), C.c_void_p()
    check(_lib().msv_radix_sort_by_key(exec.handle, _in(exec, keys)._c(), _in(exec, values)._c(),
                                       C.byref(hv), C.byref(hk) if return_keys else None))
    v = _out(exec, hv)
    return (v, _out(exec, hk)) if return_keys else v


def sort_by_key(exec: Executor, keys, values, less: BinaryPredicateOp, return_keys=False):
    """massively/src/api/algorithm/by_key.rs:193-250: stable comparison sort.
    Zipped values (tuple of columns): the keys are sorted once with their row
    indices and each column is gathered by that permutation."""
    cols = _cols(values)
    if cols is not None:
        n = _same_lengths(exec, cols)
        perm, sk = sort_by_key(exec, keys, _iota(exec, n), less, return_keys=True)
        res = gather(exec, tuple(cols), perm)
        return (res, sk) if return_keys else res
    hv, hk = C.c_void_p(), C.c_void_p()
    k = _in(exec, keys)
    if isinstance(less, UserLess):
        check(_lib().msv_sort_by_key_user(exec.handle, k._c(), _in(exec, values)._c(),
                                          less.handle(exec, k.dtype), C.byref(hv),
                                          C.byref(hk) if return_keys else None))
    else:
        check(_lib().msv_sort_by_key(exec.handle, k._c(), _in(exec, values)._c(),
                                     less.code, C.byref(hv), C.byref(hk) if return_keys else None))
    v = _out(exec, hv)
    return (v, _out(exec, hk)) if return_keys else v


def sort(exec: Executor, input, less: BinaryPredicateOp) -> DeviceVec:
    """massively/src/api/algorithm/ordering.rs:64-76."""
    h = C.c_void_p()
    x = _in(exec, input)
    if isinstance(less, UserLess):
        check(_lib().msv_sort_user(exec.handle, x._c(), less.handle(exec, x.dtype), C.byref(h)))
    else:
        check(_lib().msv_sort(exec.handle, x._c(), less.code, C.byref(h)))
    return _out(exec, h)


def merge(exec: Executor, left, right, less: BinaryPredicateOp) -> DeviceVec:
    """massively/src/api/algorithm/merge.rs:109-190: stable, ties take left."""
    h = C.c_void_p()
    a, b = _in(exec, left), _in(exec, right)
    if isinstance(less, UserLess):
        check(_lib().msv_merge_user(exec.handle, a._c(), b._c(), less.handle(exec, a.dtype),
                                    C.byref(h)))
    else:
        check(_lib().msv_merge(exec.handle, a._c(), b._c(), less.code, C.byref(h)))
    return _out(exec, h)


def merge_by_key(exec: Executor, left_keys, left_values, right_keys, right_values,
                 less: BinaryPredicateOp):
    """massively/src/api/algorithm/merge.rs:195-282. Returns (keys, values).
    Zipped values (tuples of as many columns on both sides): the keys are merged
    once with their row indices (left rows first, then right rows) and each
    column is gathered from the two sides by that order."""
    lc, rc = _cols(left_values), _cols(right_values)
    if lc is not None or rc is not None:
        if lc is None or rc is None or len(lc) != len(rc):
            raise ValueError("zipped values need as many columns on both sides")
        nl, nr = _same_lengths(exec, lc), _same_lengths(exec, rc)
        keys, pos = merge_by_key(exec, left_keys, _iota(exec, nl), right_keys, _iota(exec, nr, nl),
                                 less)
        outs = []
        for a, b in zip(lc, rc):
            both = exec.alloc(_in(exec, a).dtype, nl + nr)
            if nl:
                scatter(exec, a, lazy.counting(0).take(nl), both)
            if nr:
                scatter(exec, b, lazy.counting(nl).take(nr), both)
            outs.append(gather(exec, both, pos))
        return keys, tuple(outs)
    hk, hv = C.c_void_p(), C.c_void_p()
    lk = _in(exec, left_keys)
    args = (exec.handle, lk._c(), _in(exec, left_values)._c(), _in(exec, right_keys)._c(),
            _in(exec, right_values)._c())
    if isinstance(less, UserLess):
        check(_lib().msv_merge_by_key_user(*args, less.handle(exec, lk.dtype), C.byref(hk),
                                           C.byref(hv)))
    else:
        check(_lib().msv_merge_by_key(*args, less.code, C.byref(hk), C.byref(hv)))
    return _out(exec, hk), _out(exec, hv)


def _set(kind: int, exec: Executor, left, right, less: BinaryPredicateOp) -> DeviceVec:
    h = C.c_void_p()
    a, b = _in(exec, left), _in(exec, right)
    if isinstance(less, UserLess):
        check(_lib().msv_set_op_user(exec.handle, kind, a._c(), b._c(),
                                     less.handle(exec, a.dtype), C.byref(h)))
    else:
        check(_lib().msv_set_op(exec.handle, kind, a._c(), b._c(), less.code, C.byref(h)))
    return _out(exec, h)


def set_union(exec: Executor, left, right, less: BinaryPredicateOp) -> DeviceVec:
    """massively/src/api/algorithm/set.rs:44-110 (multiset union, capacity L+R)."""
    return _set(0, exec, left, right, less)


def set_intersection(exec: Executor, left, right, less: BinaryPredicateOp) -> DeviceVec:
    """massively/src/api/algorithm/set.rs:112-145."""
    return _set(1, exec, left, right, less)


def set_difference(exec: Executor, left, right, less: BinaryPredicateOp) -> DeviceVec:
    """massively/src/api/algorithm/set.rs:147-195."""
    return _set(2, exec, left, right, less)


# ---- by-key family -------------------------------------------------------------
# The reference's argument order is (keys, values, equal, [init,] op); the
# equality may be left out for the built-in Equal (numeric == on the key
# dtype). A UserEqual runs the same algorithms on its own runs.
def _equal_and_rest(args, n_rest):
    if len(args) == n_rest + 1 and isinstance(args[0], BinaryPredicateOp):
        return args[0], args[1:]
    if len(args) == n_rest:
        return None, args
    raise TypeError(f"expected [equal,] and {n_rest} more arguments, got {len(args)}")


def _user_equal(equal):
    return equal if isinstance(equal, UserEqual) else None


def inclusive_scan_by_key(exec: Executor, keys, values, *args) -> DeviceVec:
    """massively/src/api/algorithm/by_key.rs:399-427: ``(keys, values, [equal,] op)``.
    Zipped values: op applies to each column (a row-wise operator made of one
    built-in operator per component)."""
    equal, (op,) = _equal_and_rest(args, 1)
    cols = _cols(values)
    if cols is not None:
        _same_lengths(exec, cols)
        return tuple(inclusive_scan_by_key(exec, keys, c, *args) for c in cols)
    h = C.c_void_p()
    k, v = _in(exec, keys), _in(exec, values)
    ue = _user_equal(equal)
    if ue is not None:
        check(_lib().msv_inclusive_scan_by_key_user(exec.handle, k._c(), v._c(),
                                                    ue.handle(exec, k.dtype), op.code, C.byref(h)))
    else:
        check(_lib().msv_inclusive_scan_by_key(exec.handle, k._c(), v._c(), op.code, C.byref(h)))
    return _out(exec, h)


def exclusive_scan_by_key(exec: Executor, keys, values, *args) -> DeviceVec:
    """massively/src/api/algorithm/by_key.rs:497-546: ``(keys, values, [equal,] init, op)``.
    Zipped values: per column, ``init`` a tuple of one value per column."""
    equal, (init, op) = _equal_and_rest(args, 2)
    cols = _cols(values)
    if cols is not None:
        _same_lengths(exec, cols)
        inits = init if isinstance(init, (tuple, list)) else (init,) * len(cols)
        pre = (equal,) if equal is not None else ()
        return tuple(exclusive_scan_by_key(exec, keys, c, *pre, i0, op) for c, i0 in zip(cols, inits))
    k, v = _in(exec, keys), _in(exec, values)
    i = _scalar(init, v.dtype)
    h = C.c_void_p()
    ue = _user_equal(equal)
    if ue is not None:
        check(_lib().msv_exclusive_scan_by_key_user(exec.handle, k._c(), v._c(),
                                                    ue.handle(exec, k.dtype),
                                                    i.ctypes.data_as(C.c_void_p), op.code, C.byref(h)))
    else:
        check(_lib().msv_exclusive_scan_by_key(exec.handle, k._c(), v._c(),
                                               i.ctypes.data_as(C.c_void_p), op.code, C.byref(h)))
    return _out(exec, h)


def reduce_by_key(exec: Executor, keys, values, *args):
    """massively/src/api/algorithm/by_key.rs:618-682: ``(keys, values, [equal,] init, op)``
    gives (run keys, init (+) run values), both with a device-resident length.
    Zipped values: per column (``init`` a tuple of one value per column), with
    the run keys once."""
    equal, (init, op) = _equal_and_rest(args, 2)
    cols = _cols(values)
    if cols is not None:
        _same_lengths(exec, cols)
        inits = init if isinstance(init, (tuple, list)) else (init,) * len(cols)
        pre = (equal,) if equal is not None else ()
        pairs = [reduce_by_key(exec, keys, c, *pre, i0, op) for c, i0 in zip(cols, inits)]
        return pairs[0][0], tuple(v for _, v in pairs)
    k, v = _in(exec, keys), _in(exec, values)
    i = _scalar(init, v.dtype)
    hk, hv = C.c_void_p(), C.c_void_p()
    ue = _user_equal(equal)
    if ue is not None:
        check(_lib().msv_reduce_by_key_user(exec.handle, k._c(), v._c(), ue.handle(exec, k.dtype),
                                            i.ctypes.data_as(C.c_void_p), op.code, C.byref(hk),
                                            C.byref(hv)))
    else:
        check(_lib().msv_reduce_by_key(exec.handle, k._c(), v._c(), i.ctypes.data_as(C.c_void_p),
                                       op.code, C.byref(hk), C.byref(hv)))
    return _out(exec, hk), _out(exec, hv)


def unique_by_key(exec: Executor, keys, values, equal: BinaryPredicateOp | None = None) -> DeviceVec:
    """massively/src/api/algorithm/by_key.rs:759-789: the first value of each key run
    (zipped values: per column)."""
    cols = _cols(values)
    if cols is not None:
        _same_lengths(exec, cols)
        return tuple(unique_by_key(exec, keys, c, equal) for c in cols)
    h = C.c_void_p()
    k, v = _in(exec, keys), _in(exec, values)
    ue = _user_equal(equal)
    if ue is not None:
        check(_lib().msv_unique_by_key_user(exec.handle, k._c(), v._c(), ue.handle(exec, k.dtype),
                                            C.byref(h)))
    else:
        check(_lib().msv_unique_by_key(exec.handle, k._c(), v._c(), C.byref(h)))
    return _out(exec, h)


def map(exec: Executor, input, op: UnaryOp) -> DeviceVec:  # noqa: A001 - the reference's name
    """massively/src/api/algorithm/transform.rs:91-108: op(x) for each element,
    computed from the (possibly lazy) input in one kernel. Zipped rows with
    ``op::Identity`` give one vector per column (``MStorage::into_columns``)."""
    if isinstance(input, tuple) and op.expr.strip() == "x" and op.output is None:
        return tuple(lazy.map_into_vec(exec, lazy.map(c, op)) for c in input)
    return lazy.map_into_vec(exec, lazy.map(input, op))


def adjacent_difference(exec: Executor, input, op: ReductionOp) -> DeviceVec:
    """massively/src/api/algorithm/scan.rs:156-175: x[0], then op(x[i-1], x[i])."""
    return lazy.map_into_vec(exec, lazy.adjacent_difference(input, op))
