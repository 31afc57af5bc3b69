"""GPU parity of lazy inputs fused into the consuming kernels
(msv_lazy_op_create, msv_*_lazy), through the C ABI.

* The reference's doc examples and tests for every constructor
  (massively/src/api/lazy.rs, tests/vector/api/runtime.rs) as golden vectors.
* Fused reduce and scans over computed inputs: closed forms (sum of a
  counting stream, wrapping mod 2^32; 4e9 lazy constants, the reference's
  4G-scale test tests/vector/api/reduce.rs:60-75), and the oracle's bytes for
  scans of mapped / reversed / permuted columns.
* Every constructor written out (vector.map) against numpy indexing, for
  integer and float dtypes and unaligned views.
* copy_where with built-in and user predicates applied to a lazy input, and
  lazy inputs given to algorithms that write them out first (gather, sort).
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import lazy
from massively_amd.op import UnaryOp, UserPredicate, UserReductionOp
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(777)
SIZES = [0, 1, 2, 4095, 4096, 4097, 65537, (1 << 22) + 3]


def _host(exec_, v):
    return exec_.to_host(v)


def _u32(exec_, a):
    return exec_.to_device(np.asarray(a, np.uint32))


# ---- golden -----------------------------------------------------------------------

def test_lazy_golden(exec_):
    ident = UnaryOp("x")
    g = golden("lazy_stride_doc")
    out = m.vector.map(exec_, lazy.stride(g["start"], g["step"]).take(g["take"]), ident)
    assert _host(exec_, out).tolist() == g["expected"]
    g = golden("lazy_map_double_doc")
    out = m.vector.map(exec_, lazy.map(_u32(exec_, g["input"]), UnaryOp(g["unary"])), ident)
    assert _host(exec_, out).tolist() == g["expected"]
    g = golden("lazy_permute_doc")
    out = m.vector.map(exec_, lazy.permute(_u32(exec_, g["values"]), _u32(exec_, g["indices"])), ident)
    assert _host(exec_, out).tolist() == g["expected"]
    g = golden("lazy_reverse_doc")
    assert _host(exec_, m.vector.map(exec_, lazy.reverse(_u32(exec_, g["input"])), ident)).tolist() == g["expected"]
    g = golden("lazy_repeat_each_doc")
    out = m.vector.map(exec_, lazy.repeat_each(_u32(exec_, g["input"]), g["repeats"]), ident)
    assert _host(exec_, out).tolist() == g["expected"]
    g = golden("lazy_tile_doc")
    out = m.vector.map(exec_, lazy.tile(_u32(exec_, g["input"]), g["repeats"]), ident)
    assert _host(exec_, out).tolist() == g["expected"]
    g = golden("lazy_adjacent_difference_doc")
    op = UserReductionOp(g["reduction"], "0")
    out = m.vector.map(exec_, lazy.adjacent_difference(_u32(exec_, g["input"]), op), ident)
    assert _host(exec_, out).tolist() == g["expected"]
    g = golden("lazy_identity_doc")
    assert _host(exec_, m.vector.map(exec_, lazy.identity(_u32(exec_, g["input"])), ident)).tolist() == g["expected"]
    g = golden("lazy_counting_slice")
    t = lazy.counting(g["start"]).take(g["take"]).slice(*g["slice"])
    assert _host(exec_, m.vector.map(exec_, t, ident)).tolist() == g["expected"]
    g = golden("lazy_constant_gather")
    vals = lazy.constant(g["value"]).take(g["take"])
    idx = lazy.constant(g["indices_value"]).take(g["indices_take"])
    assert _host(exec_, m.vector.gather(exec_, vals, idx)).tolist() == g["expected_gather"]
    assert _host(exec_, m.vector.map(exec_, vals, ident)).tolist() == g["expected_map"]
    g = golden("vector_map_doc")
    assert _host(exec_, m.vector.map(exec_, _u32(exec_, g["input"]), UnaryOp(g["unary"]))).tolist() == g["expected"]
    g = golden("adjacent_difference_doc")
    out = m.vector.adjacent_difference(exec_, _u32(exec_, g["input"]), m.op.Add)
    assert _host(exec_, out).tolist() == g["expected"]


# ---- fused reduce and scans ---------------------------------------------------------

def test_reduce_counting_closed_form(exec_):
    for n in SIZES + [(1 << 26) + 5]:
        got = int(m.vector.reduce(exec_, lazy.counting(0).take(n), 0, m.op.Add))
        assert got == (n * (n - 1) // 2) % (1 << 32), n
        got = int(m.vector.reduce(exec_, lazy.counting(7).take(n), 3, m.op.Max))
        assert got == (n + 6 if n else 3), n
        assert int(m.vector.reduce(exec_, lazy.constant(1).take(n), 0, m.op.Add)) == n


def test_reduce_four_billion_lazy_constants(exec_):
    """tests/vector/api/reduce.rs:60-75 (ignored there as 4G-scale)."""
    n = 4_000_000_000
    assert int(m.vector.reduce(exec_, lazy.constant(1).take(n), 0, m.op.Add)) == n
    # the last element of a counting stream of maximal length
    assert int(m.vector.reduce(exec_, lazy.counting(0).take(0xFFFFFFFF), 0, m.op.Max)) == 0xFFFFFFFE


def test_scans_of_computed_inputs(exec_):
    for n in SIZES:
        x = RNG.integers(0, 2**32, n + 1, dtype=np.uint64).astype(np.uint32)
        d = exec_.to_device(x).slice(1, n + 1)  # unaligned view
        xs = x[1:]
        t = lazy.map(d, UnaryOp("x * 3u + 1u"))
        y = (xs * np.uint32(3) + np.uint32(1)).astype(np.uint32)
        assert _host(exec_, m.vector.inclusive_scan(exec_, t, m.op.Add)).tobytes() == \
            oracle.inclusive_scan(y, "add").tobytes(), n
        assert _host(exec_, m.vector.exclusive_scan(exec_, t, 9, m.op.Xor)).tobytes() == \
            oracle.exclusive_scan(y, 9, "xor").tobytes(), n
        assert int(m.vector.reduce(exec_, t, 5, m.op.Add)) == int(oracle.reduce(y, 5, "add")), n
        r = lazy.reverse(d)
        assert _host(exec_, m.vector.inclusive_scan(exec_, r, m.op.Min)).tobytes() == \
            oracle.inclusive_scan(xs[::-1].copy(), "min").tobytes(), n


def test_scan_user_op_over_lazy_keeps_index_order(exec_):
    big = UserReductionOp("((b < 0 ? -b : b) > (a < 0 ? -a : a)) ? b : a", "0")
    for n in (5, 4097, 100003):
        x = RNG.integers(-3, 4, n).astype(np.int64)
        idx = RNG.permutation(n).astype(np.uint32)
        t = lazy.permute(exec_.to_device(x), exec_.to_device(idx))
        y = x[idx]
        want, acc = [], None
        for v in y.tolist():
            acc = v if acc is None else (v if abs(v) > abs(acc) else acc)
            want.append(acc)
        assert _host(exec_, m.vector.inclusive_scan(exec_, t, big)).tolist() == want, n
        assert int(m.vector.reduce(exec_, t, 0, big)) == want[-1], n


def test_float_reduce_and_scan_of_map(exec_):
    # integer values whose squares and every partial sum stay exact (below
    # 2^24 for f32), so any association gives the sequential result
    for dtype, ns, lim in ((np.float32, (1, 4097), 30), (np.float64, (1, 4097, 300001), 100)):
        for n in ns:
            x = RNG.integers(-lim, lim, n).astype(dtype)
            t = lazy.map(exec_.to_device(x), UnaryOp("x * x"))
            assert float(m.vector.reduce(exec_, t, 0.0, m.op.Add)) == float((x.astype(np.float64) ** 2).sum())
            got = _host(exec_, m.vector.inclusive_scan(exec_, t, m.op.Add))
            assert got.tobytes() == np.cumsum(x * x, dtype=dtype).tobytes(), (dtype, n)


def test_monte_carlo_pi_from_zipped_map(exec_):
    """tests/vector/api/reduce.rs:25-41: count points inside the unit circle."""
    n = 100_000
    x = RNG.random(n).astype(np.float32)
    y = RNG.random(n).astype(np.float32)
    hits = lazy.map((exec_.to_device(x), exec_.to_device(y)),
                    UnaryOp("x0 * x0 + x1 * x1 <= 1.0f ? 1u : 0u", output=np.uint32))
    count = int(m.vector.reduce(exec_, hits, 0, m.op.Add))
    assert 3.0 < count / n * 4 < 3.3
    want = int(np.count_nonzero(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2 <= 1.0))
    assert abs(count - want) <= 2  # points on the circle may round either way


# ---- every constructor written out ----------------------------------------------------

@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.float32, np.float64])
def test_constructors_written_out(exec_, dtype):
    dt = np.dtype(dtype)
    for n in (1, 5, 4097, 70001):
        x = (RNG.integers(-1000, 1000, n + 1)).astype(dt)
        d = exec_.to_device(x).slice(1, n + 1)
        xs = x[1:]
        idx = RNG.integers(0, n + n // 8 + 1, n // 2 + 1).astype(np.uint32)  # some out of range
        di = exec_.to_device(idx)
        cases = [
            (lazy.reverse(d), xs[::-1]),
            (lazy.repeat_each(d, 3), np.repeat(xs, 3)),
            (lazy.tile(d, 2), np.tile(xs, 2)),
            (lazy.permute(d, di), np.where(idx < n, xs[np.minimum(idx, n - 1)], 0).astype(dt)),
            (lazy.adjacent_difference(d, m.op.Max),
             np.concatenate([xs[:1], np.maximum(xs[:-1], xs[1:])]).astype(dt)),
            (lazy.identity(d).slice(n // 3, n), xs[n // 3:]),
            (lazy.map(lazy.counting(5).take(n), UnaryOp("x", dt)), np.arange(5, 5 + n).astype(dt)),
            (lazy.constant(np.array(-2).astype(dt)[()]).take(n), np.full(n, np.array(-2).astype(dt))),
        ]
        for k, (t, want) in enumerate(cases):
            got = _host(exec_, t.materialize(exec_))
            assert got.dtype == dt and got.tobytes() == np.ascontiguousarray(want).tobytes(), (dt, n, k)


def test_map_changes_dtype(exec_):
    x = RNG.integers(0, 1 << 20, 10001).astype(np.uint32)
    out = m.vector.map(exec_, exec_.to_device(x), UnaryOp("(double)x * 0.5", output=np.float64))
    assert _host(exec_, out).tobytes() == (x.astype(np.float64) * 0.5).tobytes()


# ---- copy_where of a lazy input, and lazy inputs to other algorithms ----------------------

def test_copy_where_of_lazy_input(exec_):
    for n in SIZES:
        x = RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        d = exec_.to_device(x) if n else exec_.alloc(np.uint32, 0)
        t = lazy.map(d, UnaryOp("x ^ 0x5a5a5a5au"))
        y = x ^ np.uint32(0x5A5A5A5A)
        got = _host(exec_, m.vector.copy_where_if(exec_, t, t, m.op.Lt(1 << 30)))
        assert got.tobytes() == y[y < (1 << 30)].tobytes(), n
        got = _host(exec_, m.vector.copy_where_if(exec_, t, t, UserPredicate("(x & 7u) == 3u")))
        assert got.tobytes() == y[(y & 7) == 3].tobytes(), n
    f = lazy.map(lazy.counting(0).take(100000), UnaryOp("(float)x * 0.25f", np.float32))
    got = _host(exec_, m.vector.copy_where_if(exec_, f, f, m.op.Ge(24000.0)))
    assert got.tobytes() == (np.arange(96000, 100000) * 0.25).astype(np.float32).tobytes()


def test_lazy_inputs_to_other_algorithms(exec_):
    vals = exec_.to_device(np.array([10, 20, 30, 40], np.uint32))
    assert _host(exec_, m.vector.gather(exec_, vals, lazy.counting(1).take(3))).tolist() == [20, 30, 40]
    n = 50000
    x = RNG.integers(0, 1 << 30, n).astype(np.uint32)
    t = lazy.map(exec_.to_device(x), UnaryOp("x >> 3"))
    got = _host(exec_, m.vector.sort(exec_, t, m.op.Less))
    assert got.tobytes() == np.sort(x >> 3).tobytes()


def test_lazy_op_errors(exec_):
    with pytest.raises(m.error.InvalidArgument):
        m.vector.map(exec_, lazy.counting(0).take(4), UnaryOp("x +"))
    t = lazy.counting(0).take(4)
    with pytest.raises(ValueError):
        m.vector.remove_where_if(exec_, t, t, m.op.Lt(2))
    other = m.Executor(0)
    try:
        v = other.to_device(np.arange(4, dtype=np.uint32))
        with pytest.raises(m.ForeignExecutor):
            m.vector.reduce(exec_, lazy.map(v, UnaryOp("x")), 0, m.op.Add)
    finally:
        other.close()


def test_with_table(exec_):
    g = golden("lazy_with_table_doc")
    t = lazy.map(lazy.with_table(_u32(exec_, g["contexts"]), _u32(exec_, g["table"])),
                 UnaryOp(g["unary"], np.uint32))
    assert _host(exec_, m.vector.map(exec_, t, UnaryOp("x"))).tolist() == g["expected"]
    n, size = 200003, 5000
    ctx = RNG.integers(0, size + 100, n).astype(np.uint32)  # some past the table's end
    table = RNG.standard_normal(size)
    t = lazy.map(lazy.with_table(exec_.to_device(ctx), exec_.to_device(table)),
                 UnaryOp("table.at(x) + (double)table.len()", np.float64))
    want = np.where(ctx < size, table[np.minimum(ctx, size - 1)], 0.0) + size
    assert _host(exec_, m.vector.map(exec_, t, UnaryOp("x"))).tobytes() == want.tobytes()
    assert float(m.vector.reduce(exec_, t, 0.0, m.op.Max)) == want.max()
