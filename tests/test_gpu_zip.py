"""GPU parity of zipped columns (the reference's zip2 .. zip12 rows) in the
algorithms beyond radix_sort_by_key, and of lazy slicing chains.

Zipped gather, scatter and copy_where / remove_where move all columns in one
launch that reads the indices or the stencil once (msv_*_columns); the
other algorithms process the columns through the single-column kernels
(sort_by_key and merge_by_key order the keys once, with row indices, and
gather every column by that order). Each column must equal the single-column
result and numpy's; the columns of one call share lengths (LengthMismatch
otherwise)."""
import numpy as np
import pytest

import massively_amd as m
from massively_amd import lazy
from massively_amd.op import UnaryOp
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(515)


def _cols(n):
    return (RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
            RNG.standard_normal(n),
            RNG.integers(-2**40, 2**40, n).astype(np.int64))


def _dev(exec_, cols):
    return tuple(exec_.to_device(c) for c in cols)


def _host(exec_, vs):
    return [exec_.to_host(v) for v in vs]


def test_lazy_slicing_golden(exec_):
    ident = UnaryOp("x")
    g = golden("lazy_reverse_sliced")
    t = lazy.reverse(exec_.to_device(np.array(g["input"], np.uint32)))
    for a, b in g["slices"]:
        t = t.slice(a, b)
    assert exec_.to_host(m.vector.map(exec_, t, ident)).tolist() == g["expected"]
    for name, ctor in (("lazy_repeat_each_sliced", lazy.repeat_each), ("lazy_tile_sliced", lazy.tile)):
        g = golden(name)
        for col, want in (("input", "expected"), ("input_second", "expected_second")):
            t = ctor(exec_.to_device(np.array(g[col], np.uint32)), g["repeats"])
            for a, b in g["slices"]:
                t = t.slice(a, b)
            assert exec_.to_host(m.vector.map(exec_, t, ident)).tolist() == g[want], name
    g = golden("zip3_map_sliced")
    t = lazy.map(tuple(exec_.to_device(np.array(c, np.uint32)) for c in g["columns"]),
                 UnaryOp("x0 + x1 + x2", np.uint32))
    for a, b in g["slices"]:
        t = t.slice(a, b)
    assert exec_.to_host(m.vector.map(exec_, t, ident)).tolist() == g["expected"]


def test_zipped_gather_scatter(exec_):
    for n in (1, 4097, 100003):
        cols = _cols(n)
        d = _dev(exec_, cols)
        idx = RNG.integers(0, n, n // 2 + 1).astype(np.uint32)
        got = _host(exec_, m.vector.gather(exec_, d, exec_.to_device(idx)))
        for g, c in zip(got, cols):
            assert g.tobytes() == c[idx].tobytes(), n
        perm = RNG.permutation(n).astype(np.uint32)
        outs = tuple(exec_.alloc(c.dtype, n) for c in cols)
        m.vector.scatter(exec_, d, exec_.to_device(perm), outs)
        for o, c in zip(_host(exec_, outs), cols):
            want = np.empty_like(c)
            want[perm] = c
            assert o.tobytes() == want.tobytes(), n


def test_zipped_copy_where(exec_):
    for n in (0, 1, 4097, 300001):
        cols = _cols(n)
        d = _dev(exec_, cols) if n else tuple(exec_.alloc(c.dtype, 0) for c in cols)
        st = RNG.integers(0, 2, n).astype(np.uint32)
        ds = exec_.to_device(st) if n else exec_.alloc(np.uint32, 0)
        for got, c in zip(_host(exec_, m.vector.copy_where(exec_, d, ds)), cols):
            assert got.tobytes() == c[st != 0].tobytes(), n
        for got, c in zip(_host(exec_, m.vector.remove_where(exec_, d, ds)), cols):
            assert got.tobytes() == c[st == 0].tobytes(), n
        key = cols[0]
        for got, c in zip(_host(exec_, m.vector.copy_where_if(exec_, d, d[0], m.op.Lt(1 << 31))), cols):
            assert got.tobytes() == c[key < (1 << 31)].tobytes(), n


def test_twelve_columns_one_pass(exec_):
    """12 columns of mixed widths (u32, i32, f32, u64, i64, f64), read through
    views that start one element in (unaligned rows), against numpy and the
    single-column calls: gather with out-of-range indices (0), scatter with
    out-of-range indices (dropped) into small and into 2^24-element outputs
    (the latter column by column in the binned form), copy_where /
    remove_where and a predicate on a separate stencil column."""
    dts = [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64] * 2
    for n in (5, 8193, 70001):
        cols = [RNG.integers(0, 2**31, n + 1).astype(dt) if np.dtype(dt).kind != "f"
                else RNG.standard_normal(n + 1).astype(dt) for dt in dts]
        base = [exec_.to_device(c) for c in cols]
        views = tuple(b.slice(1, n + 1) for b in base)
        host = [c[1:] for c in cols]
        idx = RNG.integers(0, n + 7, n // 3 + 3).astype(np.uint32)
        got = _host(exec_, m.vector.gather(exec_, views, exec_.to_device(idx)))
        ok = idx < n
        for g, c in zip(got, host):
            want = np.zeros(idx.size, dtype=c.dtype)
            want[ok] = c[idx[ok]]
            assert g.tobytes() == want.tobytes(), n
        single = exec_.to_host(m.vector.gather(exec_, views[4], exec_.to_device(idx)))
        assert single.tobytes() == got[4].tobytes()
        st = (RNG.random(n) < 0.3).astype(np.uint32)
        ds = exec_.to_device(st)
        for got, c in zip(_host(exec_, m.vector.copy_where(exec_, views, ds)), host):
            assert got.tobytes() == c[st != 0].tobytes(), n
        for got, c in zip(_host(exec_, m.vector.remove_where(exec_, views, ds)), host):
            assert got.tobytes() == c[st == 0].tobytes(), n
        key = host[5]
        res = m.vector.copy_where_if(exec_, views, views[5], m.op.Gt(2**30))
        for got, c in zip(_host(exec_, res), host):
            assert got.tobytes() == c[key > 2**30].tobytes(), n
        for olen in (n, 1 << 24):
            if olen == 1 << 24 and n < 8193:
                continue
            perm = RNG.permutation(max(olen, n))[:n].astype(np.uint32)
            perm[::97] = olen + 3  # dropped
            outs = tuple(exec_.alloc(c.dtype, olen) for c in host[:3])
            m.vector.scatter(exec_, views[:3], exec_.to_device(perm), outs)
            keep = perm < olen  # destinations are distinct: check the written ones
            for o, c in zip(_host(exec_, outs), host[:3]):
                assert o[perm[keep]].tobytes() == c[keep].tobytes(), (n, olen)


def test_zipped_sort_and_merge_by_key(exec_):
    for n in (1, 5000, 200001):
        keys = RNG.integers(0, 1000, n).astype(np.uint32)  # many ties: stability matters
        cols = _cols(n)
        for less, order in ((m.op.Less, np.argsort(keys, kind="stable")),
                            (m.op.Greater, np.argsort(-keys.astype(np.int64), kind="stable"))):
            res, sk = m.vector.sort_by_key(exec_, exec_.to_device(keys), _dev(exec_, cols), less,
                                           return_keys=True)
            assert exec_.to_host(sk).tobytes() == keys[order].tobytes()
            for got, c in zip(_host(exec_, res), cols):
                assert got.tobytes() == c[order].tobytes(), (n, less)
    for nl, nr in ((0, 3), (4, 0), (3000, 5001)):
        lk = np.sort(RNG.integers(0, 500, nl)).astype(np.uint32)
        rk = np.sort(RNG.integers(0, 500, nr)).astype(np.uint32)
        lc, rc = _cols(nl), _cols(nr)
        dl = _dev(exec_, lc) if nl else tuple(exec_.alloc(c.dtype, 0) for c in lc)
        dr = _dev(exec_, rc) if nr else tuple(exec_.alloc(c.dtype, 0) for c in rc)
        dlk = exec_.to_device(lk) if nl else exec_.alloc(np.uint32, 0)
        drk = exec_.to_device(rk) if nr else exec_.alloc(np.uint32, 0)
        keys, vals = m.vector.merge_by_key(exec_, dlk, dl, drk, dr, m.op.Less)
        order = np.argsort(np.concatenate([lk, rk]), kind="stable")  # ties: left rows first
        assert exec_.to_host(keys).tobytes() == np.concatenate([lk, rk])[order].tobytes()
        for got, a, b in zip(_host(exec_, vals), lc, rc):
            assert got.tobytes() == np.concatenate([a, b])[order].tobytes(), (nl, nr)


def test_zipped_by_key(exec_):
    n = 100003
    keys = np.repeat(np.arange(n // 7 + 1, dtype=np.uint32), 7)[:n]
    ints = (RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
            RNG.integers(-2**40, 2**40, n).astype(np.int64))
    dk, d = exec_.to_device(keys), _dev(exec_, ints)
    for got, c in zip(_host(exec_, m.vector.inclusive_scan_by_key(exec_, dk, d, m.op.Add)), ints):
        want = exec_.to_host(m.vector.inclusive_scan_by_key(exec_, dk, exec_.to_device(c), m.op.Add))
        assert got.tobytes() == want.tobytes()
    got = _host(exec_, m.vector.exclusive_scan_by_key(exec_, dk, d, (3, -5), m.op.Xor))
    for g, c, i0 in zip(got, ints, (3, -5)):
        want = exec_.to_host(m.vector.exclusive_scan_by_key(exec_, dk, exec_.to_device(c), i0, m.op.Xor))
        assert g.tobytes() == want.tobytes()
    rk, rv = m.vector.reduce_by_key(exec_, dk, d, 0, m.op.Add)
    assert exec_.to_host(rk).tobytes() == np.unique(keys).tobytes()
    for g, c in zip(_host(exec_, rv), ints):
        want = np.add.reduceat(c, np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]]))
        assert g.tobytes() == want.astype(c.dtype).tobytes()
    heads = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
    for g, c in zip(_host(exec_, m.vector.unique_by_key(exec_, dk, d)), ints):
        assert g.tobytes() == c[heads].tobytes()


def test_zipped_length_mismatch(exec_):
    a = exec_.to_device(np.arange(5, dtype=np.uint32))
    b = exec_.to_device(np.arange(6, dtype=np.uint32))
    with pytest.raises(m.LengthMismatch):
        m.vector.gather(exec_, (a, b), exec_.to_device(np.zeros(2, np.uint32)))
    with pytest.raises(m.LengthMismatch):
        m.vector.copy_where(exec_, (a, b), a)


def test_lazy_adaptors_over_zipped_rows(exec_):
    """tests/vector/api/lazy.rs:301-341: reverse / repeat_each / tile of zip2,
    sliced, written out column by column."""
    first = exec_.to_device(np.array([1, 2, 3], np.uint32))
    second = exec_.to_device(np.array([10, 20, 30], np.uint32))
    a, b = m.vector.map(exec_, lazy.reverse((first, second)), UnaryOp("x"))
    assert exec_.to_host(a).tolist() == [3, 2, 1] and exec_.to_host(b).tolist() == [30, 20, 10]
    for name, ctor in (("lazy_repeat_each_sliced", lazy.repeat_each), ("lazy_tile_sliced", lazy.tile)):
        g = golden(name)
        rows = ctor((exec_.to_device(np.array(g["input"], np.uint32)),
                     exec_.to_device(np.array(g["input_second"], np.uint32))), g["repeats"])
        for lo, hi in g["slices"]:
            rows = lazy.slice_rows(rows, lo, hi)
        a, b = m.vector.map(exec_, rows, UnaryOp("x"))
        assert exec_.to_host(a).tolist() == g["expected"], name
        assert exec_.to_host(b).tolist() == g["expected_second"], name
    idx = exec_.to_device(np.array([2, 0, 7], np.uint32))
    a, b = m.vector.map(exec_, lazy.permute((first, second), idx), UnaryOp("x"))
    assert exec_.to_host(a).tolist() == [3, 1, 0] and exec_.to_host(b).tolist() == [30, 10, 0]
