"""GPU parity: seg:: scans, reduce and sort, and the Segmentation constructors.

Integer results are bit-exact. f32/f64 segmented reduce/scan use fixed-shape
passes: bitwise reproducible, and each segment's result lies within
(2 * ceil(log2 L) + 40) * eps * sum|x| of the FP64 value of that segment
(L = segment length), a bound covering the tile-local tree depth plus the
chain of tile carries.
"""
import math

import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import seg
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(23)
BOUNDARY_LENGTHS = [0, 1, 2, 31, 32, 33, 127, 128, 129, 255, 256, 257]


def offsets_of(lengths):
    return np.concatenate([[0], np.cumsum(np.asarray(lengths, np.int64))]).astype(np.uint32)


def lognormal_lengths(s, median=32, sigma=1.1774, cap=1 << 16, empty=0.02):
    L = np.minimum(np.round(RNG.lognormal(math.log(median), sigma, s)), cap).astype(np.int64)
    L[RNG.random(s) < empty] = 0
    return L


def cases():
    yield [3, 2]
    yield BOUNDARY_LENGTHS
    yield [0, 0, 5, 0, 4096, 0, 1, 4095, 4097, 0, 0]
    yield [0]
    yield [0, 0, 0]
    yield list(RNG.integers(0, 6, 3000))
    yield list(lognormal_lengths(4000))
    yield [70001]
    yield [1] * 9000
    yield [20000, 0, 30000, 1, 2, 3]


def test_golden(exec_):
    v = golden("seg_owned_results")
    vals = exec_.to_device(np.array(v["values"], np.uint32))
    offs = exec_.to_device(np.array(v["offsets"], np.uint32))
    r = seg.ForEachSegment(seg.Reduce(m.op.Add, 0)).run(exec_, seg.SegmentIterator(vals, offs))
    assert exec_.to_host(r).tolist() == v["expected"]
    v = golden("seg_exclusive_from_reduce")
    vals = exec_.to_device(np.array(v["values"], np.uint32))
    offs = exec_.to_device(np.array(v["offsets"], np.uint32))
    it = seg.ForEachSegment(seg.ExclusiveScan(m.op.Add, 0)).run(exec_, seg.SegmentIterator(vals, offs))
    assert exec_.to_host(it.values).tolist() == v["expected"]
    r = seg.ForEachSegment(seg.Reduce(m.op.Add, 0)).run(exec_, seg.SegmentIterator(vals, offs))
    assert exec_.to_host(r).tolist() == v["expected_reduce"]
    v = golden("segmentation_from_lengths")
    sg = seg.Segmentation.from_lengths(exec_, exec_.to_device(np.array(v["lengths"], np.uint32)))
    assert exec_.to_host(sg.offsets()).tolist() == v["expected_offsets"]
    assert sg.value_count() == 6 and sg.segment_count() == 3


def test_segmentation_validation(exec_):
    with pytest.raises(m.InvalidSegmentation):
        seg.Segmentation.from_offsets(exec_, exec_.to_device(np.array([1, 2], np.uint32)))
    with pytest.raises(m.InvalidSegmentation):
        seg.Segmentation.from_offsets(exec_, exec_.to_device(np.array([0, 5, 3], np.uint32)))
    with pytest.raises(m.InvalidSegmentation):
        seg.Segmentation.from_offsets(exec_, exec_.to_device(np.zeros(0, np.uint32)))
    with pytest.raises(m.LengthTooLarge):
        seg.Segmentation.from_lengths(exec_, exec_.to_device(np.array([2**31, 2**31], np.uint32)))
    sg = seg.Segmentation.from_lengths(exec_, exec_.to_device(np.zeros(0, np.uint32)))
    assert sg.segment_count() == 0 and sg.value_count() == 0
    sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(np.array([0, 3, 3, 7], np.uint32)))
    assert sg.value_count() == 7 and sg.segment_count() == 3
    with pytest.raises(m.LengthMismatch):
        seg.ForEachSegment(seg.Reduce(m.op.Add, 0)).run(
            exec_, seg.SegmentIterator(exec_.to_device(np.zeros(6, np.uint32)), sg))


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.uint64, np.int64])
def test_int_seg_scans_and_reduce(exec_, dtype):
    for lengths in cases():
        off = offsets_of(lengths)
        n = int(off[-1])
        info = np.iinfo(dtype)
        x = RNG.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
        sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
        dx = exec_.to_device(x)
        for name, op in (("add", m.op.Add), ("min", m.op.Min), ("max", m.op.Max), ("xor", m.op.Xor)):
            y = seg.ForEachSegment(seg.InclusiveScan(op)).run(exec_, seg.SegmentIterator(dx, sg))
            assert np.array_equal(exec_.to_host(y.values), oracle.seg_inclusive_scan(x, off, name)), (len(lengths), name)
            init = dtype(7)
            y = seg.ForEachSegment(seg.ExclusiveScan(op, init)).run(exec_, seg.SegmentIterator(dx, sg))
            assert np.array_equal(exec_.to_host(y.values), oracle.seg_exclusive_scan(x, off, init, name)), (len(lengths), name)
            r = seg.ForEachSegment(seg.Reduce(op, init)).run(exec_, seg.SegmentIterator(dx, sg))
            assert np.array_equal(exec_.to_host(r), oracle.seg_reduce(x, off, init, name)), (len(lengths), name)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_seg_reduce_long_segments(exec_, dtype):
    """Segments spanning hundreds of tiles (the cross-tile partials are
    combined pairwise): within (2 log2 L + 40) eps sum|x| of the FP64 sum, the
    same bytes on a second run, and integer sums exact."""
    eps = 2.0**-24 if dtype == np.float32 else 2.0**-53
    L = [3, (1 << 22) + 5, 100, 0, 0, (1 << 20) - 1, 8192, 8191, 1, 1 << 21]
    off = offsets_of(L)
    n = int(off[-1])
    x = (RNG.standard_normal(n) + 0.25).astype(dtype)
    sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
    dx = exec_.to_device(x)
    r1 = exec_.to_host(seg.ForEachSegment(seg.Reduce(m.op.Add, 1.5)).run(exec_, seg.SegmentIterator(dx, sg)))
    r2 = exec_.to_host(seg.ForEachSegment(seg.Reduce(m.op.Add, 1.5)).run(exec_, seg.SegmentIterator(dx, sg)))
    assert r1.tobytes() == r2.tobytes()
    x64 = x.astype(np.float64)
    for s, ln in enumerate(L):
        a, b = off[s], off[s + 1]
        exact = 1.5 + float(np.sum(x64[a:b]))
        bound = (2 * math.ceil(math.log2(max(ln, 2))) + 40) * eps * (1.5 + float(np.sum(np.abs(x64[a:b]))))
        assert abs(float(r1[s]) - exact) <= bound, (s, ln)
    xi = RNG.integers(0, 2**32, n, dtype=np.uint32)
    ri = exec_.to_host(seg.ForEachSegment(seg.Reduce(m.op.Add, 0)).run(
        exec_, seg.SegmentIterator(exec_.to_device(xi), sg)))
    assert np.array_equal(ri, oracle.seg_reduce(xi, off, np.uint32(0), "add"))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_seg_reduce_and_scan(exec_, dtype):
    eps = 2.0**-24 if dtype == np.float32 else 2.0**-53
    for lengths in cases():
        off = offsets_of(lengths)
        n = int(off[-1])
        x = RNG.standard_normal(n).astype(dtype)
        sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
        dx = exec_.to_device(x)
        r1 = exec_.to_host(seg.ForEachSegment(seg.Reduce(m.op.Add, 0.0)).run(exec_, seg.SegmentIterator(dx, sg)))
        r2 = exec_.to_host(seg.ForEachSegment(seg.Reduce(m.op.Add, 0.0)).run(exec_, seg.SegmentIterator(dx, sg)))
        assert r1.tobytes() == r2.tobytes()
        x64 = x.astype(np.float64)
        for s in range(len(lengths)):
            a, b = off[s], off[s + 1]
            L = int(b - a)
            exact = float(np.sum(x64[a:b]))
            bound = (2 * math.ceil(math.log2(max(L, 2))) + 40) * eps * float(np.sum(np.abs(x64[a:b])))
            assert abs(float(r1[s]) - exact) <= bound, (s, L)
        for name, op in (("min", m.op.Min), ("max", m.op.Max)):
            y = seg.ForEachSegment(seg.InclusiveScan(op)).run(exec_, seg.SegmentIterator(dx, sg))
            assert np.array_equal(exec_.to_host(y.values), oracle.seg_inclusive_scan(x, off, name))
            r = exec_.to_host(seg.ForEachSegment(seg.Reduce(op, 0.0)).run(exec_, seg.SegmentIterator(dx, sg)))
            assert np.array_equal(r, oracle.seg_reduce(x, off, 0.0, name))


@pytest.mark.parametrize("kdtype", [np.uint32, np.float32, np.int64])
def test_seg_sort(exec_, kdtype):
    for lengths in cases():
        off = offsets_of(lengths)
        n = int(off[-1])
        if np.dtype(kdtype).kind == "f":
            keys = np.round(RNG.standard_normal(n) * 3).astype(kdtype)
        else:
            keys = RNG.integers(0, 9, n).astype(kdtype)
        vals = RNG.standard_normal(n).astype(np.float32)
        sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
        for mode, cmp in (("less", m.op.Less), ("greater", m.op.Greater)):
            out = seg.ForEachSegment(seg.Sort(cmp)).run(
                exec_, seg.SegmentIterator((exec_.to_device(keys), exec_.to_device(vals)), sg))
            k, v = out.values
            perm = oracle.seg_sort_perm(keys, off, mode)
            assert exec_.to_host(k).tobytes() == keys[perm].tobytes(), (len(lengths), mode)
            assert exec_.to_host(v).tobytes() == vals[perm].tobytes(), (len(lengths), mode)


def test_batched_seg_config_2p26(exec_):
    """2^26 elements in 2^20 log-normal segments (median 32, cap 2^16, 2% empty):
    u32 scan and f32 reduce vs the oracle."""
    S = 1 << 20
    L = lognormal_lengths(S)
    total = int(L.sum())
    target = 1 << 26
    L = np.floor(L * (target / total)).astype(np.int64)
    L[-1] += target - int(L.sum())
    off = offsets_of(L)
    n = int(off[-1])
    assert n == target
    x = RNG.integers(0, 2**32, n, dtype=np.uint32)
    sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
    y = seg.ForEachSegment(seg.InclusiveScan(m.op.Add)).run(exec_, seg.SegmentIterator(exec_.to_device(x), sg))
    assert np.array_equal(exec_.to_host(y.values), oracle.seg_inclusive_scan(x, off, "add"))
    f = RNG.standard_normal(n).astype(np.float32)
    r = exec_.to_host(seg.ForEachSegment(seg.Reduce(m.op.Add, 0.0)).run(exec_, seg.SegmentIterator(exec_.to_device(f), sg)))
    exact = np.add.reduceat(f.astype(np.float64), off[:-1].astype(np.int64)) if n else np.zeros(S)
    exact[L == 0] = 0.0
    absum = np.add.reduceat(np.abs(f.astype(np.float64)), off[:-1].astype(np.int64))
    absum[L == 0] = 0.0
    bound = (2 * np.ceil(np.log2(np.maximum(L, 2))) + 40) * 2.0**-24 * absum
    assert np.all(np.abs(r.astype(np.float64) - exact) <= bound)


def _mixed_lengths():
    # Tiny and empty segments, every power-of-two size class boundary of the
    # wave sorts (1 .. 1025), segments around every candidate's half tile and
    # tile (512 .. 8192), and segments longer than any tile.
    L = list(RNG.integers(0, 40, 2000))
    for b in range(11):
        L += [(1 << b) - 1, 1 << b, (1 << b) + 1]
    for t in (512, 1024, 2048, 4096, 8192):
        L += [t // 2 - 1, t // 2, t // 2 + 1, 0, t - 1, t, t + 1, 3]
    L += [20000, 1, 0, 65536, 2, 9000]
    L += list(lognormal_lengths(3000))
    return L


@pytest.mark.parametrize("wave", [1, 0])
@pytest.mark.parametrize("launch", [m.Executor.SEG_LAUNCH_PERSISTENT, m.Executor.SEG_LAUNCH_PER_CLASS])
@pytest.mark.parametrize("kdtype,vdtype", [(np.uint32, np.float32), (np.int32, None), (np.float64, np.uint64),
                                           (np.uint64, np.uint32)])
def test_seg_sort_every_tile_config(exec_, wave, launch, kdtype, vdtype):
    """Every tile-shape candidate and both launch forms give the oracle's order,
    with short segments sorted inside a wave (wave=1) or packed into windows
    (wave=0). Half the keys come from 5 values, so ties (stability) are common."""
    off = offsets_of(_mixed_lengths())
    n = int(off[-1])
    if np.dtype(kdtype).kind == "f":
        keys = np.round(RNG.standard_normal(n) * 50).astype(kdtype)
        keys[::7] = -0.0
        keys[1::2] = np.round(keys[1::2] / 40)
    else:
        keys = RNG.integers(0, 1 << 20, n).astype(kdtype)
        keys[1::2] = RNG.integers(0, 5, keys[1::2].size).astype(kdtype)
    vals = None if vdtype is None else RNG.integers(0, 2**32, n).astype(vdtype)
    sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
    dk = exec_.to_device(keys)
    cols = dk if vals is None else (dk, exec_.to_device(vals))
    ran = 0
    try:
        exec_.set_option(m.Executor.OPT_SEGSORT_WAVE, wave)
        exec_.set_option(m.Executor.OPT_SEG_LAUNCH, launch)
        for k in range(len(m.segsort_configs())):
            exec_.set_option(m.Executor.OPT_SEGSORT_CONFIG, k)
            for mode, cmp in (("less", m.op.Less), ("greater", m.op.Greater)):
                try:
                    out = seg.ForEachSegment(seg.Sort(cmp)).run(exec_, seg.SegmentIterator(cols, sg))
                except m.Unsupported:
                    break  # this tile shape does not fit the element width
                perm = oracle.seg_sort_perm(keys, off, mode)
                got = out.values if vals is None else out.values[0]
                assert exec_.to_host(got).tobytes() == keys[perm].tobytes(), (k, mode)
                if vals is not None:
                    assert exec_.to_host(out.values[1]).tobytes() == vals[perm].tobytes(), (k, mode)
                ran += 1
    finally:
        exec_.set_option(m.Executor.OPT_SEGSORT_CONFIG, -1)
        exec_.set_option(m.Executor.OPT_SEG_LAUNCH, m.Executor.SEG_LAUNCH_PERSISTENT)
        exec_.set_option(m.Executor.OPT_SEGSORT_WAVE, 1)
    assert ran >= 2


@pytest.mark.parametrize("wave", [1, 0])
@pytest.mark.parametrize("kdtype,vdtype", [(np.uint32, np.float32), (np.int64, None), (np.float64, np.uint64)])
def test_seg_sort_long_segments(exec_, wave, kdtype, vdtype):
    """Segments longer than a block's tile, sorted by many blocks each: many
    long segments between short ones, one of 300000 elements, keys that vary
    in the low byte only (the other digit passes are skipped), and a segment
    set whose keys are all equal (every pass skipped but the data still moves)."""
    L = []
    for _ in range(40):
        L += list(RNG.integers(0, 50, 5)) + [int(RNG.integers(4097, 30000))]
    L += [300000, 7, 4096, 4097, 8192, 8193, 0]
    off = offsets_of(L)
    n = int(off[-1])
    kinds = {"random": None, "low_byte": None, "equal": None}
    for kind in kinds:
        if np.dtype(kdtype).kind == "f":
            keys = np.round(RNG.standard_normal(n) * 1e3).astype(kdtype)
            keys[::5] = -0.0
            keys[1::5] = 0.0
            if kind == "low_byte":
                keys = RNG.integers(0, 4, n).astype(kdtype)
        else:
            keys = RNG.integers(-(1 << 40) if np.dtype(kdtype).kind == "i" else 0, 1 << 31, n).astype(kdtype)
            if kind == "low_byte":
                keys = RNG.integers(0, 256, n).astype(kdtype)
        if kind == "equal":
            keys = np.full(n, 3, dtype=kdtype)
        vals = None if vdtype is None else RNG.integers(0, 2**32, n).astype(vdtype)
        sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
        dk = exec_.to_device(keys)
        cols = dk if vals is None else (dk, exec_.to_device(vals))
        try:
            exec_.set_option(m.Executor.OPT_SEGSORT_WAVE, wave)
            for mode, cmp in (("less", m.op.Less), ("greater", m.op.Greater)):
                out = seg.ForEachSegment(seg.Sort(cmp)).run(exec_, seg.SegmentIterator(cols, sg))
                perm = oracle.seg_sort_perm(keys, off, mode)
                got = out.values if vals is None else out.values[0]
                assert exec_.to_host(got).tobytes() == keys[perm].tobytes(), (kind, mode)
                if vals is not None:
                    assert exec_.to_host(out.values[1]).tobytes() == vals[perm].tobytes(), (kind, mode)
        finally:
            exec_.set_option(m.Executor.OPT_SEGSORT_WAVE, 1)


def test_seg_sort_config_2p26_autotuned(exec_):
    """The batched-segment configuration: 2^26 u32 keys + f32 values in 2^20
    log-normal segments, autotuned tile shape, persistent launch."""
    S = 1 << 20
    L = lognormal_lengths(S)
    L = np.floor(L * ((1 << 26) / int(L.sum()))).astype(np.int64)
    L[-1] += (1 << 26) - int(L.sum())
    off = offsets_of(L)
    n = int(off[-1])
    keys = RNG.integers(0, 2**32, n, dtype=np.uint32)
    vals = RNG.standard_normal(n).astype(np.float32)
    sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
    out = seg.ForEachSegment(seg.Sort(m.op.Less)).run(
        exec_, seg.SegmentIterator((exec_.to_device(keys), exec_.to_device(vals)), sg))
    perm = oracle.seg_sort_perm(keys, off, "less")
    assert exec_.to_host(out.values[0]).tobytes() == keys[perm].tobytes()
    assert exec_.to_host(out.values[1]).tobytes() == vals[perm].tobytes()


def test_tuning_cache_round_trip(exec_, tmp_path):
    """The seg::Sort autotuner's choice is saved, cleared and loaded back (the
    lookup then returns it without timing again), and sorts stay exact."""
    m.tuning_clear()
    L = RNG.integers(0, 200, 1 << 12)
    off = offsets_of(L)
    n = int(off[-1])
    keys = RNG.integers(0, 2**32, n, dtype=np.uint32)
    vals = RNG.standard_normal(n).astype(np.float32)
    sg = seg.Segmentation.from_offsets(exec_, exec_.to_device(off))
    it = seg.SegmentIterator((exec_.to_device(keys), exec_.to_device(vals)), sg)
    launch = m.Executor.SEG_LAUNCH_PERSISTENT
    out = seg.ForEachSegment(seg.Sort(m.op.Less)).run(exec_, it)
    k = m.segsort_tuned_config(4, 4, launch, n, len(L))
    assert 0 <= k < len(m.segsort_configs())
    path = tmp_path / "tune.txt"
    m.tuning_save(path)
    m.tuning_clear()
    assert m.segsort_tuned_config(4, 4, launch, n, len(L)) == -1
    m.tuning_load(path)
    assert m.segsort_tuned_config(4, 4, launch, n, len(L)) == k
    again = seg.ForEachSegment(seg.Sort(m.op.Less)).run(exec_, it)
    perm = oracle.seg_sort_perm(keys, off, "less")
    for o in (out, again):
        assert exec_.to_host(o.values[0]).tobytes() == keys[perm].tobytes()
        assert exec_.to_host(o.values[1]).tobytes() == vals[perm].tobytes()
