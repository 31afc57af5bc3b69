"""Pins the CPU oracle (oracle/msv_oracle.c) before it is trusted as the checker.

Each oracle function is checked against at least one of: the reference's own
known-answer vectors (tests/golden/vectors.json), an independent library
routine (numpy's cumulative ufuncs, stable argsort, fancy indexing), a closed
form, or brute force on tiny inputs. The checks are chosen so a dropped term,
a wrong sign/index or a swapped operand fails one of them.
"""
from collections import Counter
from functools import reduce as pyreduce

import numpy as np
import pytest

import oracle
from tests.conftest import GOLDEN, NP, golden, golden_of, keys_of

RNG = np.random.default_rng(1234)
INT_DTYPES = [np.uint32, np.int32, np.uint64, np.int64]
LENGTHS = [0, 1, 2, 31, 32, 33, 255, 256, 257, 1023, 1024, 1025, 4097]


def rand(dtype, n, lo=None, hi=None):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return RNG.standard_normal(n).astype(dt)
    info = np.iinfo(dt)
    lo = info.min if lo is None else lo
    hi = info.max if hi is None else hi
    return RNG.integers(lo, hi, size=n, dtype=dt, endpoint=True)


# ---------------------------------------------------------------- golden vectors
def test_golden_scans():
    for v in golden_of("inclusive_scan"):
        dt = NP[v["dtype"]]
        x = np.array(v["input"], dtype=dt) if "input" in v else np.full(v["len"], v["fill"], dt)
        y = oracle.inclusive_scan(x, v["reduce"])
        if "expected" in v:
            assert y.tolist() == v["expected"], v["name"]
        else:
            assert y[v["check_indices"]].tolist() == v["expected_at"]
            assert int(y[-1]) == v["expected_last"]
    for v in golden_of("exclusive_scan"):
        dt = NP[v["dtype"]]
        x = np.array(v["input"], dtype=dt) if "input" in v else np.full(v["len"], v["fill"], dt)
        y = oracle.exclusive_scan(x, v["init"], v["reduce"])
        if "expected" in v:
            assert y.tolist() == v["expected"], v["name"]
        else:
            assert y[v["check_indices"]].tolist() == v["expected_at"], v["name"]
    v = golden("exclusive_scan_init_once")
    out = np.array(v["output_before"], dtype=np.uint32)
    lo, hi = v["output_range"]
    out[lo:hi] = oracle.exclusive_scan(np.array(v["input"], np.uint32), v["init"], "add")
    assert out.tolist() == v["expected"]


def test_golden_reduce_select_move():
    v = golden("reduce_add_doc")
    assert oracle.reduce(np.array(v["input"], np.uint32), v["init"], "add") == v["expected"]
    for name, fn in (("copy_where_doc", oracle.copy_where), ("remove_where_doc", oracle.remove_where)):
        v = golden(name)
        assert fn(np.array(v["input"], np.uint32), np.array(v["stencil"], np.uint32)).tolist() == v["expected"]
    v = golden("gather_doc")
    assert oracle.gather(np.array(v["values"], np.uint32), v["indices"]).tolist() == v["expected"]
    v = golden("scatter_doc")
    out = oracle.scatter(np.array(v["values"], np.uint32), v["indices"], v["output_before"])
    assert out.tolist() == v["expected"]


def test_golden_ordering():
    v = golden("merge_stable")
    k, val = oracle.merge_by_key(np.array(v["left_keys"], np.uint32), np.array(v["left_values"], np.uint32),
                                 np.array(v["right_keys"], np.uint32), np.array(v["right_values"], np.uint32))
    assert k.tolist() == v["expected_keys"] and val.tolist() == v["expected_values"]
    v = golden("set_multiplicity")
    a, b = np.array(v["left"], np.uint32), np.array(v["right"], np.uint32)
    for kind in ("union", "intersection", "difference"):
        assert oracle.set_op(kind, a, b).tolist() == v[f"expected_{kind}"], kind
    for v in golden_of("radix_sort_by_key"):
        keys = keys_of(v)
        values = np.array(v["values"], dtype=np.uint32)
        _, out = oracle.radix_sort_by_key(keys, values)
        if v["expected"] == "stable_argsort":
            exp = values[np.argsort(keys, kind="stable")]
            assert out.tolist() == exp.tolist(), v["name"]
        else:
            assert out.tolist() == v["expected"], v["name"]


def test_golden_seg():
    v = golden("seg_owned_results")
    assert oracle.seg_reduce(np.array(v["values"], np.uint32), v["offsets"], v["init"], "add").tolist() == v["expected"]
    v = golden("seg_exclusive_from_reduce")
    x = np.array(v["values"], np.uint32)
    assert oracle.seg_exclusive_scan(x, v["offsets"], v["init"], "add").tolist() == v["expected"]
    assert oracle.seg_reduce(x, v["offsets"], v["init"], "add").tolist() == v["expected_reduce"]


# ---------------------------------------------------------------- library routines
@pytest.mark.parametrize("dtype", INT_DTYPES)
@pytest.mark.parametrize("n", [0, 1, 7, 1000, 70001])
def test_int_scans_match_numpy_accumulate(dtype, n):
    x = rand(dtype, n)
    with np.errstate(over="ignore"):
        assert np.array_equal(oracle.inclusive_scan(x, "add"), np.add.accumulate(x, dtype=dtype))
    assert np.array_equal(oracle.inclusive_scan(x, "min"), np.minimum.accumulate(x))
    assert np.array_equal(oracle.inclusive_scan(x, "max"), np.maximum.accumulate(x))
    assert np.array_equal(oracle.inclusive_scan(x, "xor"), np.bitwise_xor.accumulate(x))
    # exclusive(init) = inclusive over [init] ++ x[:-1]   (invariant linking the two)
    init = rand(dtype, 1)[0]
    for op in ("add", "min", "max", "xor"):
        ex = oracle.exclusive_scan(x, init, op)
        if n:
            shifted = np.concatenate([[init], x[:-1]]).astype(dtype)
            assert np.array_equal(ex, oracle.inclusive_scan(shifted, op)), op


def test_u32_add_wraps():
    x = np.full(5, 0xFFFFFFFF, dtype=np.uint32)
    assert oracle.inclusive_scan(x, "add").tolist() == [(k * 0xFFFFFFFF) % 2**32 for k in range(1, 6)]
    assert oracle.reduce(x, 3, "add") == (3 + 5 * 0xFFFFFFFF) % 2**32


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_float_scan_is_sequential_left_fold(dtype):
    x = rand(dtype, 3000)
    # np.add.accumulate is a sequential left fold in the element dtype.
    assert np.array_equal(oracle.inclusive_scan(x, "add"), np.add.accumulate(x, dtype=dtype))
    assert np.array_equal(oracle.inclusive_scan(x, "min"), np.minimum.accumulate(x))


@pytest.mark.parametrize("dtype", INT_DTYPES + [np.float32, np.float64])
def test_reduce_brute_force(dtype):
    x = rand(dtype, 257)
    init = rand(dtype, 1)[0]
    def add(a, b):
        with np.errstate(over="ignore"):
            return dtype(a + b) if np.dtype(dtype).kind == "f" else np.array([a, b], dtype).sum(dtype=dtype)
    assert oracle.reduce(x, init, "add") == pyreduce(add, x, init)
    assert oracle.reduce(x, init, "min") == min([init] + list(x))
    assert oracle.reduce(x, init, "max") == max([init] + list(x))


def test_fp64_sum_and_bound():
    x = RNG.standard_normal(100000).astype(np.float32)
    s, a = oracle.sum_f32_f64(x)
    import math
    assert abs(s - math.fsum(x.astype(np.float64))) < 1e-9
    assert abs(a - math.fsum(np.abs(x.astype(np.float64)))) < 1e-9
    assert oracle.f32_sum_bound(2**28, 1.0) == 30 * 2.0**-24


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64, np.float32])
def test_selection_matches_boolean_mask(dtype):
    x = rand(dtype, 5000)
    st = RNG.integers(0, 3, 5000, dtype=np.uint32)
    assert np.array_equal(oracle.copy_where(x, st), x[st != 0])
    assert np.array_equal(oracle.remove_where(x, st), x[st == 0])


def test_predicates_match_numpy():
    x = rand(np.uint32, 4000)
    t = np.uint32(1 << 31)
    cases = [(0, 0, 0, x != 0), (1, 0, 0, x == 0), (2, int(t), 0, x < t), (3, int(t), 0, x <= t),
             (4, int(t), 0, x > t), (5, int(t), 0, x >= t),
             (6, 0x10, 0, (x & 0x10) != 0), (7, 0x30, 0x10, (x & 0x30) == 0x10)]
    for kind, a, b, mask in cases:
        assert np.array_equal(oracle.copy_where_pred(x, x, kind, a, b), x[mask]), kind
        assert np.array_equal(oracle.copy_where_pred(x, x, kind, a, b, invert=True), x[~mask]), kind
    f = rand(np.float32, 4000)
    th = np.float32(0.25)
    bits = int(np.array([th]).view(np.uint32)[0])
    assert np.array_equal(oracle.copy_where_pred(x, f, 2, bits), x[f < th])
    i = rand(np.int32, 4000)
    tb = int(np.array([-5], np.int32).view(np.uint32)[0])
    assert np.array_equal(oracle.copy_where_pred(x, i, 4, tb), x[i > -5])


def test_gather_scatter_match_indexing():
    v = rand(np.uint64, 3000)
    idx = RNG.permutation(3000).astype(np.uint32)
    assert np.array_equal(oracle.gather(v, idx), v[idx])
    out = np.zeros(3000, np.uint64)
    out[idx] = v
    assert np.array_equal(oracle.scatter(v, idx, np.zeros(3000, np.uint64)), out)


@pytest.mark.parametrize("dtype", INT_DTYPES + [np.float32, np.float64])
@pytest.mark.parametrize("n", [0, 1, 33, 513, 4099])
def test_sort_perm_matches_stable_argsort(dtype, n):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        keys = np.round(RNG.standard_normal(n) * 4).astype(dt)  # many ties incl. -0/+0
        keys[::7] = -0.0
    else:
        keys = rand(dtype, n, 0, 9) if dt.kind == "u" else rand(dtype, n, -4, 4)
    perm = oracle.sort_perm(keys, "less")
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
    # Greater: stable descending = reversed stable-ascending of the reversed input.
    rev = (n - 1 - np.argsort(keys[::-1], kind="stable"))[::-1]
    assert np.array_equal(oracle.sort_perm(keys, "greater"), rev)
    if dt.kind != "f":
        assert np.array_equal(oracle.sort_perm(keys, "total"), np.argsort(keys, kind="stable"))


def test_total_order_separates_zeros_and_orders_nans():
    f = np.array([0.0, -0.0, 1.0, -1.0], np.float32)
    assert oracle.sort_perm(f, "total").tolist() == [3, 1, 0, 2]
    assert oracle.sort_perm(f, "less").tolist() == [3, 0, 1, 2]
    big = rand(np.float32, 2000)
    assert np.array_equal(oracle.sort_perm(big, "total"), np.argsort(big, kind="stable"))


@pytest.mark.parametrize("na,nb", [(0, 0), (0, 5), (5, 0), (100, 77), (1000, 3000)])
def test_merge_matches_stable_argsort_of_concat(na, nb):
    a = np.sort(rand(np.uint32, na, 0, 50))
    b = np.sort(rand(np.uint32, nb, 0, 50))
    src = oracle.merge_src(a, b)
    assert np.array_equal(src, np.argsort(np.concatenate([a, b]), kind="stable"))


def _multiset(kind, a, b):
    ca, cb = Counter(a.tolist()), Counter(b.tolist())
    out = []
    for k in sorted(set(ca) | set(cb)):
        m, n = ca.get(k, 0), cb.get(k, 0)
        out += [k] * {"union": max(m, n), "intersection": min(m, n), "difference": max(m - n, 0)}[kind]
    return out


@pytest.mark.parametrize("kind", ["union", "intersection", "difference"])
@pytest.mark.parametrize("na,nb,hi", [(0, 0, 5), (0, 9, 5), (9, 0, 5), (200, 300, 20), (2000, 1500, 3000)])
def test_set_ops_match_multiset_counts(kind, na, nb, hi):
    a = np.sort(rand(np.uint32, na, 0, hi))
    b = np.sort(rand(np.uint32, nb, 0, hi))
    assert oracle.set_op(kind, a, b).tolist() == _multiset(kind, a, b)


def _rank_rule_src(kind, a, b):
    """Sources of a set operation by the reference's rank rules
    (core/set.rs:18-46): with m = count of a key on the left and n on the
    right, and r the rank of an element inside its equal run, intersection
    keeps left elements with r < n, difference left elements with r >= n, and
    union all left elements plus right elements with r >= m, in merge order."""
    ca, cb = Counter(a.tolist()), Counter(b.tolist())
    ra, rb = Counter(), Counter()
    out = []  # (key, side, position)
    for i, k in enumerate(a.tolist()):
        r = ra[k]
        ra[k] += 1
        n = cb.get(k, 0)
        if kind == "union" or (kind == "intersection" and r < n) or (kind == "difference" and r >= n):
            out.append((k, 0, i))
    if kind == "union":
        for j, k in enumerate(b.tolist()):
            r = rb[k]
            rb[k] += 1
            if r >= ca.get(k, 0):
                out.append((k, 1, len(a) + j))
    out.sort()
    return [p for _, _, p in out]


@pytest.mark.parametrize("kind", ["union", "intersection", "difference"])
@pytest.mark.parametrize("na,nb,hi", [(0, 7, 4), (7, 0, 4), (300, 200, 12), (1500, 2500, 900)])
def test_set_op_src_follows_rank_rules(kind, na, nb, hi):
    a = np.sort(rand(np.uint32, na, 0, hi))
    b = np.sort(rand(np.uint32, nb, 0, hi))
    src = oracle.set_op_src(kind, a, b)
    assert src.tolist() == _rank_rule_src(kind, a, b)
    assert np.array_equal(np.concatenate([a, b])[src], oracle.set_op(kind, a, b))


def _offsets(lengths):
    return np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint32)


def test_seg_ops_match_per_segment_numpy():
    lengths = [0, 1, 2, 31, 32, 33, 127, 128, 129, 0, 255, 256, 257, 0]
    off = _offsets(lengths)
    x = rand(np.uint32, int(off[-1]), 0, 1000)
    segs = [x[off[i]:off[i + 1]] for i in range(len(lengths))]
    assert np.array_equal(oracle.seg_inclusive_scan(x, off, "add"),
                          np.concatenate([np.add.accumulate(s, dtype=np.uint32) for s in segs]))
    ex = oracle.seg_exclusive_scan(x, off, 7, "add")
    exp = np.concatenate([np.concatenate([[7], 7 + np.add.accumulate(s, dtype=np.uint32)[:-1]]).astype(np.uint32)
                          if len(s) else s for s in segs])
    assert np.array_equal(ex, exp)
    assert oracle.seg_reduce(x, off, 5, "max").tolist() == [max([5] + s.tolist()) for s in segs]
    keys = rand(np.uint32, int(off[-1]), 0, 6)
    perm = oracle.seg_sort_perm(keys, off)
    exp = np.concatenate([off[i] + np.argsort(keys[off[i]:off[i + 1]], kind="stable")
                          for i in range(len(lengths))]).astype(np.uint32)
    assert np.array_equal(perm, exp)


def test_regression_pairs_sort_stably():
    for r in GOLDEN["regressions"]:
        if r["kind"] != "pairs":
            continue
        pairs = np.array(r["data"], dtype=np.uint32)
        k, v = pairs[:, 0].copy(), pairs[:, 1].copy()
        _, out = oracle.sort_by_key(k, v)
        assert np.array_equal(out, v[np.argsort(k, kind="stable")])


def test_generators_follow_reference_properties():
    # massively/tests/vector/api/random.rs:204-226 and :60-64, :118-121.
    assert oracle.uniform_u32(4, 4, 1, 8).tolist() == [4] * 8
    assert oracle.uniform_u64(9, 9, 1, 8).tolist() == [9] * 8
    a = oracle.uniform_u32(10, 20, 123, 1000)
    assert a.min() >= 10 and a.max() <= 20
    assert np.array_equal(a, oracle.uniform_u32(10, 20, 123, 1000))
    assert not np.array_equal(a, oracle.uniform_u32(10, 20, 124, 1000))
    # Windows of one stream are consistent (slicing a lazy stream, random.rs:88-101).
    assert np.array_equal(oracle.uniform_u32(0, 2**32 - 1, 5, 10, start=100),
                          oracle.uniform_u32(0, 2**32 - 1, 5, 200)[100:110])
    # pcg_hash32 is a bijection: no collisions over a window of 2^20 inputs.
    h = oracle.uniform_u32(0, 2**32 - 1, 0, 1 << 20)
    assert np.unique(h).size == h.size


# ---- by-key family --------------------------------------------------------
def _np_runs(keys):
    """Run starts under the built-in Equal (numeric ==; NaN equals nothing)."""
    k = np.asarray(keys)
    if k.size == 0:
        return np.zeros(0, np.int64)
    return np.flatnonzero(np.r_[True, k[1:] != k[:-1]])


def test_by_key_golden():
    v = golden("inclusive_scan_by_key_doc")
    k, x = np.array(v["keys"], np.uint32), np.array(v["values"], np.uint32)
    assert oracle.inclusive_scan_by_key(k, x, "add").tolist() == v["expected"]
    v = golden("exclusive_scan_by_key_doc")
    assert oracle.exclusive_scan_by_key(k, x, v["init"], "add").tolist() == v["expected"]
    v = golden("reduce_by_key_doc")
    rk, rv = oracle.reduce_by_key(k, x, v["init"], "add")
    assert rk.tolist() == v["expected_keys"] and rv.tolist() == v["expected"]
    v = golden("unique_by_key_doc")
    assert oracle.unique_by_key(np.array(v["keys"], np.uint32),
                                np.array(v["values"], np.uint32)).tolist() == v["expected"]
    v = golden("by_key_owned")
    k, x = np.array(v["keys"], np.uint32), np.array(v["values"], np.uint32)
    assert oracle.inclusive_scan_by_key(k, x, "add").tolist() == v["expected_inclusive"]
    assert oracle.exclusive_scan_by_key(k, x, v["init"], "add").tolist() == v["expected_exclusive"]
    rk, rv = oracle.reduce_by_key(k, x, v["init"], "add")
    assert rk.tolist() == v["expected_reduce_keys"] and rv.tolist() == v["expected_reduce"]
    assert oracle.unique_by_key(k, x).tolist() == v["expected_unique"]


@pytest.mark.parametrize("kdtype", [np.uint32, np.int64, np.float32])
@pytest.mark.parametrize("op", ["add", "max", "xor"])
def test_by_key_vs_numpy_runs(kdtype, op):
    """Against numpy: per-run ufunc accumulate / reduceat on run boundaries
    from `k[1:] != k[:-1]`, the first value of each run for unique."""
    rng = np.random.default_rng(9)
    for n in (0, 1, 2, 7, 300, 5000):
        if np.dtype(kdtype).kind == "f":
            k = np.sort(rng.integers(-20, 20, n)).astype(kdtype)
            k[k == 0] = np.where(rng.random((k == 0).sum()) < 0.5, -0.0, 0.0).astype(kdtype)
        else:
            k = np.sort(rng.integers(0, max(1, n // 7), n)).astype(kdtype)
        x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        starts = _np_runs(k)
        uf = {"add": np.add, "max": np.maximum, "xor": np.bitwise_xor}[op]
        inc = np.empty_like(x)
        exc = np.empty_like(x)
        red = []
        bounds = list(starts) + [n]
        for b, e in zip(bounds[:-1], bounds[1:]):
            inc[b:e] = uf.accumulate(x[b:e], dtype=np.uint32)
            ex = np.concatenate([[np.uint32(5)], x[b:e - 1]]).astype(np.uint32)
            exc[b:e] = uf.accumulate(ex, dtype=np.uint32)
            red.append(uf.reduce(np.concatenate([[np.uint32(5)], x[b:e]]).astype(np.uint32),
                                 dtype=np.uint32))
        assert np.array_equal(oracle.inclusive_scan_by_key(k, x, op), inc)
        assert np.array_equal(oracle.exclusive_scan_by_key(k, x, 5, op), exc)
        rk, rv = oracle.reduce_by_key(k, x, 5, op)
        assert rk.tobytes() == k[starts].tobytes()
        assert rv.tolist() == [int(r) for r in red]
        assert np.array_equal(oracle.unique_by_key(k, x), x[starts])


# ---- partition, scatter_reduce, compacting segment operations ---------------

def test_partition_golden_and_masks():
    g = golden("partition_even_doc")
    out, b = oracle.partition(np.array(g["input"], np.uint32), 7, 1, 0)  # (x & 1) == 0
    assert out.tolist() == g["expected"] and b == g["expected_boundary"]
    for dt in (np.uint32, np.int64, np.float32):
        x = rand(dt, 3001)
        if np.dtype(dt).kind == "f":
            x[::97] = np.nan  # NaN < t is false: NaNs go to the failing side
            t = np.float32(0.1)
            kind, a = 2, int(np.array([t]).view(np.uint32)[0])
            mask = x < t
        else:
            kind, a = 6, 0x10
            mask = (x.view(np.uint32 if x.itemsize == 4 else np.uint64) & 0x10) != 0
        out, b = oracle.partition(x, kind, a)
        exp = np.concatenate([x[mask], x[~mask]])
        assert b == int(mask.sum())
        assert out.tobytes() == exp.tobytes()


def test_scatter_reduce_golden():
    for name in ("scatter_reduce_doc", "scatter_reduce_owned"):
        g = golden(name)
        out = oracle.scatter_reduce(np.array(g["values"], np.uint32), g["indices"], g["init"], "add",
                                    np.array(g["output_before"], np.uint32))
        assert out.tolist() == g["expected"], name


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.uint64, np.int64])
def test_scatter_reduce_matches_ufunc_at(dtype):
    """Integer results do not depend on the order of combination, so numpy's
    unbuffered ufunc.at (sequential in index order) is the same answer."""
    n, m = 5000, 700
    v = rand(dtype, n)
    idx = RNG.integers(0, m + 50, n).astype(np.uint32)  # some out of range: dropped
    base = rand(dtype, m)
    keep = idx < m
    for op, uf, init in (("add", np.add, 0), ("min", np.minimum, np.iinfo(dtype).max),
                         ("max", np.maximum, np.iinfo(dtype).min), ("xor", np.bitwise_xor, 0)):
        exp = base.copy()
        uf.at(exp, idx[keep], v[keep])
        got = oracle.scatter_reduce(v, idx, init, op, base)
        assert got.tobytes() == exp.tobytes(), op


def test_scatter_reduce_float_folds_in_input_order():
    """op(out[d], init (+) v1 (+) v2 ...), left to right: with values whose sums
    are exact the order does not matter; with 2^24 + 1 + 1 it does (a fold that
    started from out[d] would round twice)."""
    v = np.array([1.0, 1.0, 3.0], np.float32)
    out = oracle.scatter_reduce(v, [0, 0, 1], 0.0, "add", np.array([2.0**24, 5.0], np.float32))
    # proposal for 0 is 0 + 1 + 1 = 2, and 2^24 + 2 is exact in f32.
    assert out.tolist() == [2.0**24 + 2.0, 8.0]
    x = rand(np.float32, 3000)
    idx = RNG.integers(0, 400, 3000).astype(np.uint32)
    base = rand(np.float32, 400)
    exp = base.copy()
    np.minimum.at(exp, idx, x)
    assert oracle.scatter_reduce(x, idx, np.inf, "min", base).tobytes() == exp.tobytes()


def _segs(lengths, x):
    off = _offsets(lengths)
    return off, [x[off[i]:off[i + 1]] for i in range(len(lengths))]


def _np_compact(parts):
    lens = [len(p) for p in parts]
    vals = np.concatenate(parts) if parts else np.array([])
    return vals, _offsets(lens)


def test_seg_compact_golden():
    g = golden("seg_unique_owned")
    vals, off = oracle.seg_unique(np.array(g["input"], np.uint32), g["offsets"])
    assert vals.tolist() == g["expected"] and off.tolist() == g["expected_offsets"]
    g = golden("seg_take_empty_segments")
    x = np.array(g["input"], np.uint32)
    vals, off = oracle.seg_take(x, g["offsets"], g["take"])
    assert vals.tolist() == g["expected"] and off.tolist() == g["expected_offsets"]
    vals, off = oracle.seg_take(x, g["offsets"], 0)
    assert vals.tolist() == g["expected_take0"] and off.tolist() == g["expected_offsets_take0"]


def test_seg_compact_matches_per_segment_numpy():
    lengths = [0, 1, 2, 31, 32, 33, 0, 0, 127, 128, 129, 255, 256, 257, 0]
    x = rand(np.uint32, sum(lengths), 0, 4)
    off, segs = _segs(lengths, x)
    vals, o = oracle.seg_filter(x, off, 2, 2)  # x < 2
    ev, eo = _np_compact([s[s < 2] for s in segs])
    assert vals.tobytes() == ev.astype(np.uint32).tobytes() and np.array_equal(o, eo)
    vals, o = oracle.seg_unique(x, off)
    ev, eo = _np_compact([s[np.r_[True, s[1:] != s[:-1]]] if len(s) else s for s in segs])
    assert vals.tobytes() == ev.astype(np.uint32).tobytes() and np.array_equal(o, eo)
    for k in (0, 1, 3, 32, 300):
        vals, o = oracle.seg_take(x, off, k)
        ev, eo = _np_compact([s[:k] for s in segs])
        assert vals.tobytes() == ev.astype(np.uint32).tobytes() and np.array_equal(o, eo), k


def test_seg_unique_uses_numeric_equality():
    """-0.0 equals +0.0 (dropped), a NaN equals nothing (kept), and a segment
    start is always kept even when it equals the previous segment's end."""
    x = np.array([0.0, -0.0, 0.0, np.nan, np.nan, 1.0, 1.0, 1.0], np.float32)
    vals, off = oracle.seg_unique(x, [0, 3, 6, 8])
    # [0, -0, 0] -> [0]; [nan, nan, 1] -> all kept; [1, 1] -> [1] (its start is kept)
    assert vals.size == 5 and not np.signbit(vals).any()
    assert vals[0] == 0.0 and np.isnan(vals[1:3]).all() and vals[3:].tolist() == [1.0, 1.0]
    assert off.tolist() == [0, 1, 4, 5]


def test_compound_sort_perm_golden_and_tuples():
    for name in ("radix_compound_keys_seven_values", "radix_three_keys_twelve_values"):
        g = golden(name)
        cols = [np.array(k, NP[d]) for k, d in zip(g["keys"], g["key_dtypes"])]
        perm = oracle.compound_sort_perm(cols)
        for vals, exp in zip(g["values"], g["expected"]):
            assert np.array(vals)[perm].tolist() == exp, name
    # Against Python's stable sort on tuples of total-order words (i64, f32 with
    # NaN and +-0, u32): a reversed column order or an unstable sort fails.
    n = 3000
    a = rand(np.int64, n, -3, 3)
    b = rand(np.float32, n)
    b[::7] = np.float32(0.0)
    b[::11] = np.float32(-0.0)
    b[::13] = np.nan
    c = rand(np.uint32, n, 0, 3)
    wa, wb, wc = (oracle.order_words(x, "total") for x in (a, b, c))
    want = sorted(range(n), key=lambda i: (int(wa[i]), int(wb[i]), int(wc[i])))
    assert oracle.compound_sort_perm([a, b, c]).tolist() == want
