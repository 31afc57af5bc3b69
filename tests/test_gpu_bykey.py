"""GPU parity of the by-key family (inclusive/exclusive_scan_by_key,
reduce_by_key, unique_by_key) against the oracle, through the C ABI.

Integer values are bit-exact. f32 values: each scan output and each run's
reduction lies within (2 * ceil(log2 L) + 40) * eps * sum|x| of the FP64 value
(L = run length), the segmented-scan bound; results are bitwise reproducible.
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(31)
SIZES = [0, 1, 2, 100, 4095, 4096, 4097, 8191, 8192, 8193, 70000, 1 << 20]


def _keys(kdtype, n, runs):
    k = np.sort(RNG.integers(0, max(1, runs), n))
    if np.dtype(kdtype).kind == "f":
        k = (k - runs // 2).astype(kdtype)
        z = k == 0
        k[z] = np.where(RNG.random(int(z.sum())) < 0.5, -0.0, 0.0).astype(kdtype)
        if n > 10:
            k[n // 3] = np.nan  # a NaN key equals nothing, so it is a run of its own
    return k.astype(kdtype)


def test_by_key_golden(exec_):
    v = golden("by_key_owned")
    k = exec_.to_device(np.array(v["keys"], np.uint32))
    x = exec_.to_device(np.array(v["values"], np.uint32))
    assert exec_.to_host(m.vector.inclusive_scan_by_key(exec_, k, x, m.op.Add)).tolist() == v["expected_inclusive"]
    assert exec_.to_host(m.vector.exclusive_scan_by_key(exec_, k, x, 0, m.op.Add)).tolist() == v["expected_exclusive"]
    rk, rv = m.vector.reduce_by_key(exec_, k, x, 0, m.op.Add)
    assert exec_.to_host(rk).tolist() == v["expected_reduce_keys"]
    assert exec_.to_host(rv).tolist() == v["expected_reduce"]
    assert exec_.to_host(m.vector.unique_by_key(exec_, k, x)).tolist() == v["expected_unique"]
    for name in ("inclusive_scan_by_key_doc", "exclusive_scan_by_key_doc", "reduce_by_key_doc"):
        g = golden(name)
        k = exec_.to_device(np.array(g["keys"], np.uint32))
        x = exec_.to_device(np.array(g["values"], np.uint32))
        if name.startswith("inclusive"):
            assert exec_.to_host(m.vector.inclusive_scan_by_key(exec_, k, x, m.op.Add)).tolist() == g["expected"]
        elif name.startswith("exclusive"):
            assert exec_.to_host(m.vector.exclusive_scan_by_key(exec_, k, x, 0, m.op.Add)).tolist() == g["expected"]
        else:
            rk, rv = m.vector.reduce_by_key(exec_, k, x, 0, m.op.Add)
            assert exec_.to_host(rk).tolist() == g["expected_keys"]
            assert exec_.to_host(rv).tolist() == g["expected"]


@pytest.mark.parametrize("kdtype", [np.uint32, np.int64, np.float32])
@pytest.mark.parametrize("op,name", [(m.op.Add, "add"), (m.op.Max, "max"), (m.op.Xor, "xor")])
def test_by_key_int_values_exact(exec_, kdtype, op, name):
    for n in SIZES:
        for runs in (1, max(1, n // 3), max(1, n // 40)):
            k = _keys(kdtype, n, runs)
            x = RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
            dk, dx = exec_.to_device(k), exec_.to_device(x)
            got = exec_.to_host(m.vector.inclusive_scan_by_key(exec_, dk, dx, op))
            assert got.tobytes() == oracle.inclusive_scan_by_key(k, x, name).tobytes(), (n, runs)
            got = exec_.to_host(m.vector.exclusive_scan_by_key(exec_, dk, dx, 7, op))
            assert got.tobytes() == oracle.exclusive_scan_by_key(k, x, 7, name).tobytes(), (n, runs)
            rk, rv = m.vector.reduce_by_key(exec_, dk, dx, 7, op)
            wk, wv = oracle.reduce_by_key(k, x, 7, name)
            assert exec_.to_host(rk).tobytes() == wk.tobytes(), (n, runs)
            assert exec_.to_host(rv).tobytes() == wv.tobytes(), (n, runs)
            u = exec_.to_host(m.vector.unique_by_key(exec_, dk, dx))
            assert u.tobytes() == oracle.unique_by_key(k, x).tobytes(), (n, runs)


def test_by_key_u64_values_and_f32_values(exec_):
    n = 300000
    k = _keys(np.uint32, n, 5000)
    x = RNG.integers(0, 2**63, n, dtype=np.uint64)
    dk, dx = exec_.to_device(k), exec_.to_device(x)
    assert exec_.to_host(m.vector.unique_by_key(exec_, dk, dx)).tobytes() == oracle.unique_by_key(k, x).tobytes()
    rk, rv = m.vector.reduce_by_key(exec_, dk, dx, 3, m.op.Add)
    wk, wv = oracle.reduce_by_key(k, x, 3, "add")
    assert exec_.to_host(rv).tobytes() == wv.tobytes()
    # f32 values: bounded error per run and bitwise reproducible
    f = RNG.standard_normal(n).astype(np.float32)
    df = exec_.to_device(f)
    y1 = exec_.to_host(m.vector.inclusive_scan_by_key(exec_, dk, df, m.op.Add))
    y2 = exec_.to_host(m.vector.inclusive_scan_by_key(exec_, dk, df, m.op.Add))
    assert y1.tobytes() == y2.tobytes()
    starts = np.flatnonzero(np.r_[True, k[1:] != k[:-1]])
    ends = np.r_[starts[1:], n]
    f64 = f.astype(np.float64)
    for b, e in zip(starts[:200], ends[:200]):
        exact = np.cumsum(f64[b:e])
        mag = np.cumsum(np.abs(f64[b:e]))
        L = e - b
        assert np.all(np.abs(y1[b:e] - exact) <= (2 * np.ceil(np.log2(max(L, 2))) + 40) * 2.0**-24 * mag + 1e-30)
    rk, rv = m.vector.reduce_by_key(exec_, dk, df, 0.0, m.op.Add)
    rvh = exec_.to_host(rv)
    sums = np.add.reduceat(f64, starts)
    mags = np.add.reduceat(np.abs(f64), starts)
    L = ends - starts
    assert np.all(np.abs(rvh - sums) <= (2 * np.ceil(np.log2(np.maximum(L, 2))) + 40) * 2.0**-24 * mags + 1e-30)


def test_by_key_length_mismatch(exec_):
    k = exec_.to_device(np.arange(5, dtype=np.uint32))
    x = exec_.to_device(np.arange(4, dtype=np.uint32))
    with pytest.raises(m.LengthMismatch):
        m.vector.inclusive_scan_by_key(exec_, k, x, m.op.Add)
    with pytest.raises(m.LengthMismatch):
        m.vector.unique_by_key(exec_, k, x)


@pytest.mark.parametrize("kdtype,vdtype", [(np.uint64, np.uint64), (np.float32, np.int64),
                                           (np.float64, np.uint32), (np.int32, np.uint64)])
def test_by_key_mixed_widths_and_unaligned_views(exec_, kdtype, vdtype):
    """Key and value widths differ (the keys are read per value vector), and
    views that start one element in (not 16-byte aligned) take the scalar
    loads; both give the oracle's bytes."""
    for n in (5, 2049, 4100, 20000, 300001):
        k = _keys(kdtype, n + 1, max(1, n // 7))
        x = RNG.integers(0, 2**63, n + 1, dtype=np.uint64).astype(vdtype)
        dk, dx = exec_.to_device(k), exec_.to_device(x)
        for lo in (0, 1):
            kk, xx = k[lo:lo + n], x[lo:lo + n]
            sk, sx = dk.slice(lo, lo + n), dx.slice(lo, lo + n)
            got = exec_.to_host(m.vector.inclusive_scan_by_key(exec_, sk, sx, m.op.Add))
            assert got.tobytes() == oracle.inclusive_scan_by_key(kk, xx, "add").tobytes(), (n, lo)
            got = exec_.to_host(m.vector.exclusive_scan_by_key(exec_, sk, sx, 5, m.op.Min))
            assert got.tobytes() == oracle.exclusive_scan_by_key(kk, xx, 5, "min").tobytes(), (n, lo)
            rk, rv = m.vector.reduce_by_key(exec_, sk, sx, 9, m.op.Xor)
            wk, wv = oracle.reduce_by_key(kk, xx, 9, "xor")
            assert exec_.to_host(rk).tobytes() == wk.tobytes(), (n, lo)
            assert exec_.to_host(rv).tobytes() == wv.tobytes(), (n, lo)
            u = exec_.to_host(m.vector.unique_by_key(exec_, sk, sx))
            assert u.tobytes() == oracle.unique_by_key(kk, xx).tobytes(), (n, lo)
