"""Lazy inputs on the host, without a GPU: the lowering of every constructor
of massively/src/api/lazy.rs to a read source, the reference's length rules
and errors, and that each generated source compiles (hiprtc, no device) into
every consuming kernel for every dtype."""
import numpy as np
import pytest

import massively_amd as m
from massively_amd import _lib as L
from massively_amd import lazy
from massively_amd.op import UnaryOp, UserPredicate, UserReductionOp
from massively_amd.runtime import DeviceSlice

DTYPES = [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64]


class _FakeVec:
    """Stands in for a device vector: the lowering reads only dtype / len."""

    handle = None

    def __init__(self, dtype, n):
        self.dtype, self.n = np.dtype(dtype), n

    def len(self):
        return self.n


def _slot(dtype, n, offset=0, length=None):
    return DeviceSlice(_FakeVec(dtype, n), offset, L.WHOLE if length is None else length)


def _everything(dtype):
    """One tree that uses every constructor."""
    x = _slot(dtype, 100, offset=1)
    idx = _slot(np.uint32, 50)
    t = lazy.permute(lazy.map(x, UnaryOp("x")), idx)
    t = lazy.tile(lazy.repeat_each(lazy.reverse(t), 2), 3)
    t = lazy.adjacent_difference(t, m.op.Add).slice(1, 250)
    n = len(t)
    parts = (t, lazy.constant(np.array(3, dtype)[()]).take(n),
             lazy.map(lazy.counting(5).take(n), UnaryOp("x", dtype)),
             lazy.identity(lazy.map(lazy.stride(1, 2).take(n), UnaryOp("x", dtype))))
    return lazy.map(parts, UnaryOp("x0 + x1 + x2 + x3", output=dtype))


def test_lengths_follow_the_reference():
    x = _slot(np.uint32, 10)
    assert len(lazy.map(x, UnaryOp("x"))) == 10
    assert len(lazy.permute(x, _slot(np.uint32, 4))) == 4
    assert len(lazy.reverse(x)) == 10
    assert len(lazy.repeat_each(x, 3)) == 30 and len(lazy.tile(x, 3)) == 30
    assert len(lazy.repeat_each(x, 0)) == 0
    assert len(lazy.adjacent_difference(x, m.op.Add)) == 10
    assert len(lazy.counting(7).take(5).slice(1, 4)) == 3
    assert len(_everything(np.float32)) == 249


def test_length_and_type_errors():
    with pytest.raises(m.LengthTooLarge):
        lazy.repeat_each(_slot(np.uint32, 1 << 31), 2)
    with pytest.raises(m.LengthTooLarge):
        lazy.counting(0).take(1 << 32)
    with pytest.raises(m.LengthMismatch):
        lazy.map((_slot(np.float32, 3), _slot(np.float32, 4)), UnaryOp("x0", np.float32))
    with pytest.raises(TypeError):
        lazy.map(lazy.counting(0), UnaryOp("x"))  # no .take
    with pytest.raises(TypeError):
        lazy.permute(_slot(np.uint32, 3), _slot(np.int64, 3))  # indices are MIndex
    with pytest.raises(TypeError):
        lazy.adjacent_difference(_slot(np.float32, 3), m.op.Xor).lower()
    with pytest.raises(ValueError):
        lazy.counting(0).take(5).slice(2, 9)
    many = [_slot(np.uint32, 4) for _ in range(14)]
    with pytest.raises(ValueError):
        lazy.map(many[:13], UnaryOp("x0", np.uint32))  # zip takes at most 12 inputs
    t = lazy.map(many[:12], UnaryOp("x0", np.uint32))
    t = lazy.map((t, many[12], many[13]), UnaryOp("x0", np.uint32))
    with pytest.raises(ValueError):
        t.lower()  # 14 slots > 13


def test_lowering_records_slots_params_and_length():
    x = _slot(np.float32, 100, offset=3, length=40)
    src, args = lazy.adjacent_difference(lazy.reverse(x), m.op.Max).lower()
    assert "msv_read" in src and "MSV_SLOT(0, float, i)" in src
    assert "Combine<float, 2>::apply" in src
    assert args.nslots == 1 and args.len == 40
    assert args.slots[0].offset == 3 and args.slots[0].len == 40
    assert args.params[0] == 40  # reverse reads len - 1 - i


def test_source_depends_on_shape_not_values():
    """Parameters are run-time arguments, so one compiled operator serves
    every call of the same shape."""
    a = lazy.map(lazy.counting(3).take(10), UnaryOp("x * 2u")).lower()
    b = lazy.map(lazy.counting(900).take(77), UnaryOp("x * 2u")).lower()
    assert a[0] == b[0] and a[1].params[0] != b[1].params[0]
    c = lazy.constant(np.float32(1.5)).take(3).lower()
    d = lazy.constant(np.float32(-2.0)).take(3).lower()
    assert c[0] == d[0] and c[1].params[0] == np.float32(1.5).view(np.uint32)


@pytest.mark.parametrize("dtype", DTYPES)
def test_generated_sources_compile(dtype):
    """Every constructor, for every dtype, into the map kernel and a built-in
    reduction; the remaining consumer kernels for the narrowest and widest
    dtypes (their instantiation does not otherwise depend on the dtype)."""
    src, _ = _everything(dtype).lower()
    lazy.check_compiles("map", dtype, src)
    if dtype in (np.uint32, np.float64):
        lazy.check_compiles("reduction", dtype, src, m.op.Add)
        lazy.check_compiles("reduction", dtype, src, UserReductionOp("a > b ? a : b", "0"))
        lazy.check_compiles("predicate", dtype, src)  # the built-in predicates
        lazy.check_compiles("predicate", dtype, src, UserPredicate("x > 1"))
    else:
        lazy.check_compiles("reduction", dtype, src, m.op.Xor if np.dtype(dtype).kind != "f" else m.op.Min)


def test_user_adjacent_op_and_compile_errors():
    x = _slot(np.uint32, 10)
    src, _ = lazy.adjacent_difference(x, UserReductionOp("b - a", "0")).lower()
    assert "apply(uint32_t a, uint32_t b)" in src
    lazy.check_compiles("map", np.uint32, src)
    bad, _ = lazy.map(x, UnaryOp("x +")).lower()
    with pytest.raises(m.error.InvalidArgument) as e:
        lazy.check_compiles("map", np.uint32, bad)
    assert "compile" in str(e.value)
    with pytest.raises(m.error.InvalidArgument):  # a map needs a read source
        lazy.check_compiles("map", np.uint32, None)
    with pytest.raises(m.error.InvalidArgument):  # xor of floats
        lazy.check_compiles("reduction", np.float32, src, m.op.Xor)


@pytest.mark.parametrize("dtype", [np.uint32, np.float64])
def test_user_ops_on_device_vectors_compile(dtype):
    """msv_user_op_create's kernels (pointer inputs, no read source) under
    hiprtc, whose standard headers differ from the library build's."""
    lazy.check_compiles("reduction", dtype, None, UserReductionOp("a > b ? a : b", "0"))
    lazy.check_compiles("predicate", dtype, None, UserPredicate("x > 1"))


@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.float32])
def test_user_comparators_compile(dtype):
    """msv_user_op_create's comparator kernels (block sort, merge passes,
    merge, set walk; run heads) under hiprtc, without a device."""
    lib = L.load()
    dt = L.DTYPE_OF_NP[np.dtype(dtype)]
    assert lib.msv_lazy_op_check(4, dt, None, -1, b"a < b", None) == 0, L.last_error().message
    assert lib.msv_lazy_op_check(5, dt, None, -1, b"a == b", None) == 0, L.last_error().message
    assert lib.msv_lazy_op_check(4, dt, None, -1, b"a <", None) == L.ERR_INVALID_ARGUMENT
    assert b"does not compile" in L.last_error().message
    assert lib.msv_lazy_op_check(4, dt, b"x", -1, b"a < b", None) == L.ERR_INVALID_ARGUMENT


def test_with_table_lowering():
    t = lazy.map(lazy.with_table(_slot(np.uint32, 7), _slot(np.float64, 3)),
                 UnaryOp("table.at(x) * 2.0", np.float64))
    src, args = t.lower()
    assert len(t) == 7 and args.nslots == 2 and "TableView<double>" in src
    lazy.check_compiles("map", np.float64, src)
    with pytest.raises(TypeError):
        lazy.with_table(_slot(np.uint32, 3), lazy.counting(0).take(3))  # the table is device data


def test_adaptors_over_zipped_rows():
    a, b = _slot(np.uint32, 3), _slot(np.float32, 3)
    r = lazy.reverse((a, b))
    assert isinstance(r, tuple) and [len(c) for c in r] == [3, 3]
    assert [c.dtype for c in lazy.tile((a, b), 2)] == [np.dtype(np.uint32), np.dtype(np.float32)]
    assert [len(c) for c in lazy.slice_rows(lazy.repeat_each((a, b), 3), 2, 8)] == [6, 6]
    with pytest.raises(m.LengthMismatch):
        lazy.reverse((a, _slot(np.uint32, 4)))
