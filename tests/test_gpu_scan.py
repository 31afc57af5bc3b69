"""GPU parity: vector::inclusive_scan / exclusive_scan / reduce vs the oracle.

Integer results are bit-exact (u32/u64 addition wraps). Float scans and
reductions use fixed-shape trees: they must be bitwise identical run to run,
and f32/f64 sums must lie within (ceil(log2 n)+2) * eps * sum|x| of the FP64
sum (eps = 2^-24 for f32, 2^-53 for f64).
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from tests.conftest import NP, golden, golden_of

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(7)
INT_DT = [np.uint32, np.int32, np.uint64, np.int64]
# Boundary lengths of the reference's property tests
# (verification/oracle/tests/vector/common.rs:12-24) plus tile edges (4096 u32 / 2048 u64).
LENGTHS = [0, 1, 2, 31, 32, 33, 255, 256, 257, 511, 512, 513, 767, 768, 769, 1023, 1024, 1025,
           2047, 2048, 2049, 4095, 4096, 4097, 8191, 8192, 8193, 65535, 65536, 65537, 70001,
           262145]
OPS = [("add", m.op.Add), ("min", m.op.Min), ("max", m.op.Max), ("xor", m.op.Xor)]


def rand(dtype, n):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return RNG.standard_normal(n).astype(dt)
    info = np.iinfo(dt)
    return RNG.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)


def test_golden_scan_and_reduce(exec_):
    for v in golden_of("inclusive_scan") + golden_of("exclusive_scan"):
        dt = NP[v["dtype"]]
        x = np.array(v["input"], dt) if "input" in v else np.full(v["len"], v["fill"], dt)
        d = exec_.to_device(x)
        if v["op"] == "inclusive_scan":
            y = exec_.to_host(m.vector.inclusive_scan(exec_, d, m.op.Add))
        else:
            y = exec_.to_host(m.vector.exclusive_scan(exec_, d, v["init"], m.op.Add))
        if "expected" in v:
            assert y.tolist() == v["expected"], v["name"]
        else:
            assert y[v["check_indices"]].tolist() == v["expected_at"], v["name"]
    v = golden("reduce_add_doc")
    assert m.vector.reduce(exec_, exec_.to_device(np.array(v["input"], np.uint32)), 0, m.op.Add) == 10


@pytest.mark.parametrize("dtype", INT_DT)
def test_int_scans_bit_exact(exec_, dtype):
    for n in LENGTHS:
        x = rand(dtype, n)
        d = exec_.to_device(x)
        for name, op in OPS:
            y = exec_.to_host(m.vector.inclusive_scan(exec_, d, op))
            assert np.array_equal(y, oracle.inclusive_scan(x, name)), (n, name)
            init = rand(dtype, 1)[0]
            y = exec_.to_host(m.vector.exclusive_scan(exec_, d, init, op))
            assert np.array_equal(y, oracle.exclusive_scan(x, init, name)), (n, name, "excl")


@pytest.mark.parametrize("dtype", INT_DT)
def test_int_reduce_bit_exact(exec_, dtype):
    for n in LENGTHS:
        x = rand(dtype, n)
        d = exec_.to_device(x)
        for name, op in OPS:
            init = rand(dtype, 1)[0]
            assert m.vector.reduce(exec_, d, init, op) == oracle.reduce(x, init, name), (n, name)


def test_scan_of_unaligned_slices(exec_):
    x = rand(np.uint32, 20000)
    d = exec_.to_device(x)
    for off, end in [(1, 19999), (3, 8195), (5, 5), (17, 4113)]:
        y = exec_.to_host(m.vector.inclusive_scan(exec_, d.slice(off, end), m.op.Add))
        assert np.array_equal(y, oracle.inclusive_scan(x[off:end], "add")), (off, end)
        r = m.vector.reduce(exec_, d.slice(off, end), 0, m.op.Add)
        assert r == oracle.reduce(x[off:end], 0, "add")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_reduce_within_fp64_bound_and_deterministic(exec_, dtype):
    eps = 2.0**-24 if dtype == np.float32 else 2.0**-53
    for n in [1, 1000, 65537, 1 << 20, 3_000_001]:
        x = rand(dtype, n)
        d = exec_.to_device(x)
        r1 = m.vector.reduce(exec_, d, 0.0, m.op.Add)
        r2 = m.vector.reduce(exec_, d, 0.0, m.op.Add)
        assert np.array([r1]).view(np.uint8).tobytes() == np.array([r2]).view(np.uint8).tobytes()
        s = float(np.sum(x.astype(np.float64)))
        a = float(np.sum(np.abs(x.astype(np.float64))))
        import math
        bound = (math.ceil(math.log2(max(n, 2))) + 2) * eps * a
        assert abs(float(r1) - s) <= bound, (n, float(r1), s, bound)
        assert m.vector.reduce(exec_, d, 0.0, m.op.Max) == max(0.0, x.max())
        assert m.vector.reduce(exec_, d, 0.0, m.op.Min) == min(0.0, x.min())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_scan_deterministic_and_bounded(exec_, dtype):
    eps = 2.0**-24 if dtype == np.float32 else 2.0**-53
    for n in [1, 4096, 4097, 100003, 1 << 20]:
        x = rand(dtype, n)
        d = exec_.to_device(x)
        y1 = exec_.to_host(m.vector.inclusive_scan(exec_, d, m.op.Add))
        y2 = exec_.to_host(m.vector.inclusive_scan(exec_, d, m.op.Add))
        assert y1.tobytes() == y2.tobytes()
        ref = np.cumsum(x.astype(np.float64))
        absc = np.cumsum(np.abs(x.astype(np.float64)))
        import math
        bound = (math.ceil(math.log2(max(n, 2))) + 16) * eps * absc
        assert np.all(np.abs(y1.astype(np.float64) - ref) <= bound + 1e-30)
        ye = exec_.to_host(m.vector.exclusive_scan(exec_, d, 1.5, m.op.Add))
        assert ye[0] == dtype(1.5)
        assert np.all(np.abs(ye[1:].astype(np.float64) - (1.5 + ref[:-1])) <= bound[:-1] + 4 * eps)
        for name, op in (("min", m.op.Min), ("max", m.op.Max)):
            y = exec_.to_host(m.vector.inclusive_scan(exec_, d, op))
            assert np.array_equal(y, oracle.inclusive_scan(x, name))


def test_scan_full_size_u32_wraps(exec_):
    """2^28 u32 uniform in [0,1024), wrapping add: size-independent properties
    plus a direct oracle comparison."""
    n = 1 << 28
    x = exec_.alloc(np.uint32, n)
    from massively_amd import random as mr
    mr.uniform_u32(exec_, 0, 1023, 2, x)
    y = m.vector.inclusive_scan(exec_, x, m.op.Add)
    hx = exec_.to_host(x)
    hy = exec_.to_host(y)
    assert np.array_equal(hy, oracle.inclusive_scan(hx, "add"))
    total = m.vector.reduce(exec_, x, 0, m.op.Add)
    assert int(hy[-1]) == int(total)
