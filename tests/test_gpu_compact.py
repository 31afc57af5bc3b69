"""GPU parity of vector::partition, vector::scatter_reduce and the compacting
segment operations seg::{Filter, Unique, Take}, through the C ABI.

Integer and byte results are bit-exact against the oracle. Float
scatter_reduce sums take the reference's shape (stable sort by index, one
reduction per run, one combine per destination), so they are bitwise
reproducible; with integer-valued inputs whose sums stay below 2^24 every
association is exact and they equal the oracle's bytes, and min / max are
exact for any values.
"""
import math

import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import seg
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(91)
SIZES = [0, 1, 5, 4095, 4096, 4097, 8191, 8192, 8193, 70001, (1 << 22) + 3]


def _rand(dtype, n):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return RNG.standard_normal(n).astype(dt)
    if dt.itemsize == 4:
        return RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(dt)
    return RNG.integers(0, 2**63, n, dtype=np.uint64).view(dt)


def _bits(v, dtype):
    a = np.array([v], dtype=dtype)
    return int(a.view(np.uint32 if a.itemsize == 4 else np.uint64)[0])


# ---- partition ------------------------------------------------------------------

def test_partition_golden(exec_):
    g = golden("partition_even_doc")
    out, b = m.vector.partition(exec_, exec_.to_device(np.array(g["input"], np.uint32)),
                                m.op.BitsEq(1, 0))
    assert exec_.to_host(out).tolist() == g["expected"] and b == g["expected_boundary"]


@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.float32, np.float64])
def test_partition_matches_oracle(exec_, dtype):
    for n in SIZES:
        x = _rand(dtype, n + 1)
        if np.dtype(dtype).kind == "f":
            x[::101] = np.nan
            pred, kind, a = m.op.Lt(0.25), 2, _bits(0.25, dtype)
        else:
            pred, kind, a = m.op.BitsAny(0x30), 6, 0x30
        d = exec_.to_device(x)
        for lo in (0, 1):  # the view at 1 is not 16-byte aligned
            xs = x[lo:lo + n]
            out, b = m.vector.partition(exec_, d.slice(lo, lo + n), pred)
            want, wb = oracle.partition(xs, kind, a)
            assert b == wb, (n, lo)
            assert exec_.to_host(out).tobytes() == want.tobytes(), (n, lo)


# ---- scatter_reduce -------------------------------------------------------------

def test_scatter_reduce_golden(exec_):
    for name in ("scatter_reduce_doc", "scatter_reduce_owned"):
        g = golden(name)
        out = exec_.to_device(np.array(g["output_before"], np.uint32))
        m.vector.scatter_reduce(exec_, exec_.to_device(np.array(g["values"], np.uint32)),
                                exec_.to_device(np.array(g["indices"], np.uint32)), g["init"],
                                m.op.Add, out)
        assert exec_.to_host(out).tolist() == g["expected"], name


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.uint64, np.int64])
def test_scatter_reduce_int_exact(exec_, dtype):
    info = np.iinfo(dtype)
    for n, dests in ((0, 5), (1, 1), (5000, 37), (300000, 1 << 18), (1 << 20, 1000)):
        v = _rand(dtype, n)
        idx = RNG.integers(0, dests + dests // 10 + 1, n).astype(np.uint32)  # some dropped
        base = _rand(dtype, dests)
        for op, name, init in ((m.op.Add, "add", 0), (m.op.Min, "min", info.max),
                               (m.op.Max, "max", info.min), (m.op.Xor, "xor", 0)):
            out = exec_.to_device(base)
            m.vector.scatter_reduce(exec_, exec_.to_device(v) if n else exec_.alloc(dtype, 0),
                                    exec_.to_device(idx) if n else exec_.alloc(np.uint32, 0),
                                    init, op, out)
            want = oracle.scatter_reduce(v, idx, init, name, base)
            assert exec_.to_host(out).tobytes() == want.tobytes(), (n, name)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_scatter_reduce_float(exec_, dtype):
    for n, dests in ((7, 3), (5000, 37), (300000, 1 << 16)):
        idx = RNG.integers(0, dests, n).astype(np.uint32)
        # integer-valued: every partial sum is exact, so any association agrees
        v = RNG.integers(-50, 50, n).astype(dtype)
        base = RNG.integers(-1000, 1000, dests).astype(dtype)
        out = exec_.to_device(base)
        dv, di = exec_.to_device(v), exec_.to_device(idx)
        m.vector.scatter_reduce(exec_, dv, di, 0.0, m.op.Add, out)
        assert exec_.to_host(out).tobytes() == oracle.scatter_reduce(v, idx, 0.0, "add", base).tobytes()
        # general values: min / max exact; add reproducible run to run
        g = _rand(dtype, n)
        dg = exec_.to_device(g)
        for op, name, init in ((m.op.Min, "min", np.inf), (m.op.Max, "max", -np.inf)):
            out = exec_.to_device(base)
            m.vector.scatter_reduce(exec_, dg, di, init, op, out)
            assert exec_.to_host(out).tobytes() == oracle.scatter_reduce(g, idx, init, name, base).tobytes()
        runs = []
        for _ in range(2):
            out = exec_.to_device(base)
            m.vector.scatter_reduce(exec_, dg, di, 0.0, m.op.Add, out)
            runs.append(exec_.to_host(out))
        assert runs[0].tobytes() == runs[1].tobytes()
        want = oracle.scatter_reduce(g.astype(np.float64), idx, 0.0, "add", base.astype(np.float64))
        mag = np.zeros(dests)
        np.add.at(mag, idx, np.abs(g.astype(np.float64)))
        eps = np.finfo(dtype).eps
        L = np.bincount(idx, minlength=dests)
        bound = (2 * np.ceil(np.log2(np.maximum(L, 2))) + 40) * eps * (mag + np.abs(base)) + 1e-300
        assert np.all(np.abs(runs[0].astype(np.float64) - want) <= bound)


# ---- seg::{Filter, Unique, Take} ----------------------------------------------------

def _offsets(lengths):
    return np.concatenate([[0], np.cumsum(np.asarray(lengths, np.int64))]).astype(np.uint32)


def _seg_cases():
    yield [3, 2]
    yield [0, 1, 2, 31, 32, 33, 0, 127, 128, 129, 255, 256, 257]
    yield [0, 0, 5, 0, 8192, 0, 1, 8191, 8193, 0, 0, 20000]
    yield [0]
    yield [0, 0, 0]
    L = np.minimum(np.round(RNG.lognormal(math.log(32), 1.1774, 40000)), 1 << 16).astype(np.int64)
    L[RNG.random(L.size) < 0.02] = 0
    yield L.tolist()


def _run(exec_, alg, d, off_dev):
    r = seg.ForEachSegment(alg).run(exec_, seg.SegmentIterator(d, off_dev))
    vals = exec_.to_host(r.values)
    offs = exec_.to_host(r.offsets.offsets())
    return vals, offs, r.offsets


def test_seg_compact_golden(exec_):
    g = golden("seg_unique_owned")
    vals, offs, _ = _run(exec_, seg.Unique(), exec_.to_device(np.array(g["input"], np.uint32)),
                         exec_.to_device(np.array(g["offsets"], np.uint32)))
    assert vals.tolist() == g["expected"] and offs.tolist() == g["expected_offsets"]
    g = golden("seg_take_empty_segments")
    d = exec_.to_device(np.array(g["input"], np.uint32))
    o = exec_.to_device(np.array(g["offsets"], np.uint32))
    vals, offs, _ = _run(exec_, seg.Take(g["take"]), d, o)
    assert vals.tolist() == g["expected"] and offs.tolist() == g["expected_offsets"]
    vals, offs, _ = _run(exec_, seg.Take(0), d, o)
    assert vals.tolist() == g["expected_take0"] and offs.tolist() == g["expected_offsets_take0"]


@pytest.mark.parametrize("dtype", [np.uint32, np.float32, np.int64])
def test_seg_compact_matches_oracle(exec_, dtype):
    for lengths in _seg_cases():
        off = _offsets(lengths)
        n = int(off[-1])
        if np.dtype(dtype).kind == "f":
            x = RNG.integers(-2, 3, n + 1).astype(dtype)  # many equal neighbours, +-0
            x[RNG.random(n + 1) < 0.1] *= -1
            x[RNG.random(n + 1) < 0.01] = np.nan
            pred, kind, a = m.op.Gt(0.5), 4, _bits(0.5, dtype)
        else:
            x = (RNG.integers(0, 3, n + 1)).astype(dtype)
            pred, kind, a = m.op.Lt(2), 2, _bits(2, dtype)
        dx = exec_.to_device(x)
        do = exec_.to_device(off)
        for lo in (0, 1):  # the view at 1 is not 16-byte aligned
            xs = x[lo:lo + n]
            dv = dx.slice(lo, lo + n)
            for alg, want in ((seg.Filter(pred), oracle.seg_filter(xs, off, kind, a)),
                              (seg.Unique(), oracle.seg_unique(xs, off)),
                              (seg.Take(0), oracle.seg_take(xs, off, 0)),
                              (seg.Take(3), oracle.seg_take(xs, off, 3)),
                              (seg.Take(40), oracle.seg_take(xs, off, 40)),
                              (seg.Take(1 << 30), oracle.seg_take(xs, off, 1 << 30))):
                vals, offs, sg = _run(exec_, alg, dv, do)
                assert vals.tobytes() == want[0].tobytes(), (alg, len(lengths), lo)
                assert offs.tobytes() == want[1].tobytes(), (alg, len(lengths), lo)
                assert sg.value_count() == want[0].size and sg.segment_count() == len(lengths)


def test_seg_compact_result_feeds_seg_ops(exec_):
    """The new Segmentation is a valid input of the other seg:: operations."""
    lengths = [5, 0, 300, 9000, 1]
    off = _offsets(lengths)
    x = RNG.integers(0, 4, int(off[-1])).astype(np.uint32)
    r = seg.ForEachSegment(seg.Take(100)).run(exec_, seg.SegmentIterator(exec_.to_device(x),
                                                                         exec_.to_device(off)))
    red = exec_.to_host(seg.ForEachSegment(seg.Reduce(m.op.Add, 0)).run(exec_, r))
    assert red.tolist() == [int(x[off[i]:off[i] + min(100, lengths[i])].sum()) for i in range(len(lengths))]
