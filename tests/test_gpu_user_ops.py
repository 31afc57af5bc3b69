"""GPU parity of user-defined operators compiled at run time (msv_user_op_*).

* A user reduction equal to a built-in one (a + b, max, xor) gives the
  oracle's bytes for reduce and the scans (integer types; max also for f32).
* A non-commutative associative operator (the larger magnitude, the left one
  on ties) must be combined in index order: scans and reduce equal the
  sequential left fold, which a reversed combination would not.
* A user predicate gives copy_where's oracle bytes; a compile error is an
  InvalidArgument carrying the compiler's diagnostic.
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd.op import UserPredicate, UserReductionOp

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(401)
SIZES = [0, 1, 2, 4095, 4096, 4097, 65537, (1 << 22) + 3]

ADD = UserReductionOp("a + b", "0")
MAX_F = UserReductionOp("fmaxf(a, b)", "-INFINITY")
XOR = UserReductionOp("a ^ b", "0")
# |b| > |a| ? b : a -- associative (the first of the largest magnitudes), not commutative.
BIGGER = UserReductionOp("((b < 0 ? -b : b) > (a < 0 ? -a : a)) ? b : a", "0")


def _bigger(a, b):
    return b if abs(b) > abs(a) else a


def _fold_scan(x, op, init=None):
    out = np.empty_like(x)
    acc = init
    for i, v in enumerate(x.tolist()):
        if init is None:
            acc = v if i == 0 else op(acc, v)
            out[i] = acc
        else:
            out[i] = acc
            acc = op(acc, v)
    return out


def test_user_add_and_xor_match_builtin(exec_):
    for n in SIZES:
        x = RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        d = exec_.to_device(x) if n else exec_.alloc(np.uint32, 0)
        got = exec_.to_host(m.vector.inclusive_scan(exec_, d, ADD))
        assert got.tobytes() == oracle.inclusive_scan(x, "add").tobytes(), n
        got = exec_.to_host(m.vector.exclusive_scan(exec_, d, 5, ADD))
        assert got.tobytes() == oracle.exclusive_scan(x, 5, "add").tobytes(), n
        assert int(m.vector.reduce(exec_, d, 9, ADD)) == int(oracle.reduce(x, 9, "add")), n
    x = RNG.integers(0, 2**63, 300001, dtype=np.uint64)
    d = exec_.to_device(x)
    assert int(m.vector.reduce(exec_, d, 3, XOR)) == int(oracle.reduce(x, 3, "xor"))
    got = exec_.to_host(m.vector.inclusive_scan(exec_, d, XOR))
    assert got.tobytes() == oracle.inclusive_scan(x, "xor").tobytes()


def test_user_float_max(exec_):
    for n in (1, 4097, 300001):
        x = RNG.standard_normal(n).astype(np.float32)
        d = exec_.to_device(x)
        got = exec_.to_host(m.vector.inclusive_scan(exec_, d, MAX_F))
        assert got.tobytes() == np.maximum.accumulate(x).tobytes(), n
        assert m.vector.reduce(exec_, d, np.float32(-np.inf), MAX_F) == x.max()


def test_user_noncommutative_op_keeps_index_order(exec_):
    for n in (3, 4095, 4097, 100003):
        x = RNG.integers(-3, 4, n).astype(np.int32)  # many ties of magnitude
        d = exec_.to_device(x)
        want = _fold_scan(x, _bigger)
        assert exec_.to_host(m.vector.inclusive_scan(exec_, d, BIGGER)).tolist() == want.tolist(), n
        want = _fold_scan(x, _bigger, init=0)
        assert exec_.to_host(m.vector.exclusive_scan(exec_, d, 0, BIGGER)).tolist() == want.tolist(), n
        red = 0
        for v in x.tolist():
            red = _bigger(red, v)
        assert int(m.vector.reduce(exec_, d, 0, BIGGER)) == red, n


def test_user_predicate_copy_where(exec_):
    p = UserPredicate("(x % 7u) == 3u")
    for n in SIZES:
        x = RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        d = exec_.to_device(x) if n else exec_.alloc(np.uint32, 0)
        got = exec_.to_host(m.vector.copy_where_if(exec_, d, d, p))
        assert got.tobytes() == x[x % 7 == 3].tobytes(), n
    pf = UserPredicate("x > 0.5f && x < 0.75f")
    x = RNG.random(70001).astype(np.float32)
    d = exec_.to_device(x)
    got = exec_.to_host(m.vector.copy_where_if(exec_, d, d, pf))
    assert got.tobytes() == x[(x > 0.5) & (x < 0.75)].tobytes()


def test_user_op_compile_error_and_dtype_check(exec_):
    bad = UserReductionOp("a +", "0")
    d = exec_.to_device(np.arange(10, dtype=np.uint32))
    with pytest.raises(m.error.InvalidArgument) as e:
        m.vector.inclusive_scan(exec_, d, bad)
    assert "compile" in str(e.value)
    u = UserReductionOp("a + b", "0")
    u.handle(exec_, np.uint32)
    f = exec_.to_device(np.arange(10, dtype=np.float32))
    # a separate compiled form is made per dtype
    assert float(m.vector.reduce(exec_, f, 0.0, u)) == 45.0
