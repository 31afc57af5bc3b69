"""GPU parity: copy_where / remove_where (stencil + built-in predicates),
gather / scatter, and the seeded generators, bit-exact vs the oracle."""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import random as mr
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(11)
LENGTHS = [0, 1, 2, 31, 32, 33, 255, 256, 257, 1023, 1024, 1025, 4095, 4096, 4097, 8193,
           65537, 300001]


def test_golden(exec_):
    v = golden("copy_where_doc")
    x, s = exec_.to_device(np.array(v["input"], np.uint32)), exec_.to_device(np.array(v["stencil"], np.uint32))
    assert exec_.to_host(m.vector.copy_where(exec_, x, s)).tolist() == v["expected"]
    assert exec_.to_host(m.vector.remove_where(exec_, x, s)).tolist() == golden("remove_where_doc")["expected"]
    v = golden("gather_doc")
    vals = exec_.to_device(np.array(v["values"], np.uint32))
    idx = exec_.to_device(np.array(v["indices"], np.uint32))
    assert exec_.to_host(m.vector.gather(exec_, vals, idx)).tolist() == v["expected"]
    out = exec_.to_device(np.zeros(3, np.uint32))
    m.vector.scatter(exec_, vals, idx, out)
    assert exec_.to_host(out).tolist() == golden("scatter_doc")["expected"]


@pytest.mark.parametrize("dtype", [np.uint32, np.float32, np.uint64, np.float64])
def test_copy_where_stencil(exec_, dtype):
    for n in LENGTHS:
        x = RNG.integers(0, 2**31, n).astype(dtype)
        for p in (0.0, 0.1, 0.5, 1.0):
            st = (RNG.random(n) < p).astype(np.uint32) * RNG.integers(1, 7, n, dtype=np.uint32)
            dx, ds = exec_.to_device(x), exec_.to_device(st)
            y = m.vector.copy_where(exec_, dx, ds)
            assert y.has_device_len and y.capacity == n
            assert np.array_equal(exec_.to_host(y), oracle.copy_where(x, st)), (n, p)
            assert len(y) == int((st != 0).sum())
            z = m.vector.remove_where(exec_, dx, ds)
            assert np.array_equal(exec_.to_host(z), oracle.remove_where(x, st)), (n, p)


def test_copy_where_predicates(exec_):
    n = 100003
    x = RNG.integers(0, 2**32, n, dtype=np.uint32)
    dx = exec_.to_device(x)
    t = 0x19999999  # 10% threshold
    for pred, kind, a, b in [(m.op.Lt(t), 2, t, 0), (m.op.Le(t), 3, t, 0), (m.op.Gt(t), 4, t, 0),
                             (m.op.Ge(t), 5, t, 0), (m.op.NonZero, 0, 0, 0), (m.op.Zero, 1, 0, 0),
                             (m.op.BitsAny(0x10), 6, 0x10, 0), (m.op.BitsEq(0x30, 0x10), 7, 0x30, 0x10)]:
        y = exec_.to_host(m.vector.copy_where_if(exec_, dx, dx, pred))
        assert np.array_equal(y, oracle.copy_where_pred(x, x, kind, a, b)), kind
        z = exec_.to_host(m.vector.remove_where_if(exec_, dx, dx, pred))
        assert np.array_equal(z, oracle.copy_where_pred(x, x, kind, a, b, invert=True)), kind
    # Predicate on a separate float stencil and on signed stencils.
    f = RNG.standard_normal(n).astype(np.float32)
    y = exec_.to_host(m.vector.copy_where_if(exec_, dx, exec_.to_device(f), m.op.Lt(np.float32(0.25))))
    assert np.array_equal(y, x[f < np.float32(0.25)])
    i64 = RNG.integers(-100, 100, n).astype(np.int64)
    y = exec_.to_host(m.vector.copy_where_if(exec_, dx, exec_.to_device(i64), m.op.Ge(-3)))
    assert np.array_equal(y, x[i64 >= -3])


def test_copy_where_rejects_mismatch(exec_):
    x = exec_.to_device(np.arange(5, dtype=np.uint32))
    s = exec_.to_device(np.ones(4, dtype=np.uint32))
    with pytest.raises(m.LengthMismatch) as e:
        m.vector.copy_where(exec_, x, s)
    assert e.value == m.LengthMismatch(5, 4)


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
def test_gather_scatter_permutation(exec_, dtype):
    for n in LENGTHS:
        v = RNG.integers(0, 2**32, n).astype(dtype)
        idx = RNG.permutation(n).astype(np.uint32)
        dv, di = exec_.to_device(v), exec_.to_device(idx)
        assert np.array_equal(exec_.to_host(m.vector.gather(exec_, dv, di)), oracle.gather(v, idx))
        out = exec_.to_device(np.zeros(n, dtype))
        m.vector.scatter(exec_, dv, di, out)
        assert np.array_equal(exec_.to_host(out), oracle.scatter(v, idx, np.zeros(n, dtype)))


def test_gather_out_of_range_reads_zero(exec_):
    v = exec_.to_device(np.array([5, 6, 7], np.uint32))
    idx = exec_.to_device(np.array([2, 3, 0xFFFFFFFF, 0], np.uint32))
    assert exec_.to_host(m.vector.gather(exec_, v, idx)).tolist() == [7, 0, 0, 5]


def test_generators_bit_exact(exec_):
    for lo, hi, seed in [(0, 1023, 2), (0, 2**32 - 1, 4), (10, 20, 123), (4, 4, 1)]:
        d = exec_.alloc(np.uint32, 100000)
        mr.uniform_u32(exec_, lo, hi, seed, d)
        assert np.array_equal(exec_.to_host(d), oracle.uniform_u32(lo, hi, seed, 100000))
    for lo, hi, seed in [(0, 2**64 - 1, 6), (100, 200, 123), (9, 9, 1)]:
        d = exec_.alloc(np.uint64, 100000)
        mr.uniform_u64(exec_, lo, hi, seed, d)
        assert np.array_equal(exec_.to_host(d), oracle.uniform_u64(lo, hi, seed, 100000))
    d = exec_.alloc(np.uint32, 5000)
    mr.iota_u32(exec_, d, 7)
    assert np.array_equal(exec_.to_host(d), np.arange(7, 5007, dtype=np.uint32))


def test_copy_where_full_size(exec_):
    """2^28 u32 at 10% and 50% selectivity (stencil form): count and content."""
    n = 1 << 28
    x = exec_.alloc(np.uint32, n)
    mr.uniform_u32(exec_, 0, 2**32 - 1, 4, x)
    flags = exec_.alloc(np.uint32, n)
    for p in (10, 50):
        mr.uniform_u32(exec_, 0, 99, 5, flags)
        st = m.vector.copy_where_if(exec_, flags, flags, m.op.Lt(p))  # selected flag values
        # Stencil form with flags = (u < p): build on host from the same stream.
        hf = oracle.uniform_u32(0, 99, 5, n)
        stencil = exec_.to_device((hf < p).astype(np.uint32))
        y = m.vector.copy_where(exec_, x, stencil)
        hx = exec_.to_host(x)
        exp = oracle.copy_where(hx, (hf < p).astype(np.uint32))
        got = exec_.to_host(y)
        assert got.size == exp.size and np.array_equal(got, exp)
        assert len(st) == exp.size
