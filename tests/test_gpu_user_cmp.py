"""GPU parity: user comparators (BinaryPredicateOp compiled at run time).

A user Less drives the comparison sort (sort, sort_by_key), merge,
merge_by_key, the set operations and seg::Sort; a user Equal drives the
by-key family and seg::Unique. Every comparator here is defined by a derived
key (less(a, b) == key(a) < key(b), equal(a, b) == key(a) == key(b)), so the
expected results are the oracle's stable sort, merge walk, set walk and
by-key loops applied to the derived keys, compared bit for bit. Comparators
with many ties (a % 1000, a >> 40) check stability: equal keys keep input
order, merges and set operations take the left input first.
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(23)
# Around the block-sort tile (4096) and the merge tile (8192), and several passes.
LENGTHS = [0, 1, 2, 3, 100, 511, 4095, 4096, 4097, 8191, 8192, 8193, 12289, 70001, (1 << 20) + 3]

# (key dtype, Less expression, derived order key)
LESS_CASES = [
    (np.uint32, "a < b", lambda x: x),
    (np.uint32, "(a % 1000u) < (b % 1000u)", lambda x: x % np.uint32(1000)),
    (np.int32, "a > b", lambda x: -x.astype(np.int64)),
    (np.float32, "fabsf(a) > fabsf(b)", lambda x: -np.abs(x)),
    (np.uint64, "(a >> 40) < (b >> 40)", lambda x: x >> np.uint64(40)),
    (np.int64, "(a & 0xff) > (b & 0xff)", lambda x: -(x & 0xff)),
    (np.float64, "a < b", lambda x: x),
]
IDS = [f"{np.dtype(d).name}:{e}" for d, e, _ in LESS_CASES]


def rand(dtype, n):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return RNG.standard_normal(n).astype(dt)
    info = np.iinfo(dt)
    return RNG.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)


def sorted_by(x, key):
    return x[oracle.sort_perm(key(x), "less")]


@pytest.mark.parametrize("dtype,expr,key", LESS_CASES, ids=IDS)
def test_user_sort_and_sort_by_key(exec_, dtype, expr, key):
    less = m.op.UserLess(expr)
    for n in LENGTHS:
        x = rand(dtype, n)
        perm = oracle.sort_perm(key(x), "less")
        d = exec_.to_device(x)
        assert exec_.to_host(m.vector.sort(exec_, d, less)).tobytes() == x[perm].tobytes(), n
        vals = np.arange(n, dtype=np.uint32)
        v, k = m.vector.sort_by_key(exec_, d, exec_.to_device(vals), less, return_keys=True)
        assert np.array_equal(exec_.to_host(v), perm), n
        assert exec_.to_host(k).tobytes() == x[perm].tobytes(), n
    # 8-byte values move with their keys.
    x = rand(dtype, 20_001)
    w = rand(np.float64, 20_001)
    perm = oracle.sort_perm(key(x), "less")
    v = m.vector.sort_by_key(exec_, exec_.to_device(x), exec_.to_device(w), less)
    assert exec_.to_host(v).tobytes() == w[perm].tobytes()


@pytest.mark.parametrize("dtype,expr,key", LESS_CASES, ids=IDS)
def test_user_merge_and_merge_by_key(exec_, dtype, expr, key):
    less = m.op.UserLess(expr)
    for na, nb in [(0, 0), (0, 5), (5, 0), (1, 1), (4095, 4097), (8192, 1), (70001, 30003),
                   (300_000, 500_001)]:
        a, b = sorted_by(rand(dtype, na), key), sorted_by(rand(dtype, nb), key)
        src = oracle.merge_src(key(a), key(b), "less")
        both = np.concatenate([a, b])
        da, db = exec_.to_device(a), exec_.to_device(b)
        got = exec_.to_host(m.vector.merge(exec_, da, db, less))
        assert got.tobytes() == both[src].tobytes(), (na, nb)
        va = np.arange(na, dtype=np.uint32)
        vb = np.arange(na, na + nb, dtype=np.uint32)
        k, v = m.vector.merge_by_key(exec_, da, exec_.to_device(va), db, exec_.to_device(vb), less)
        assert np.array_equal(exec_.to_host(v), src), (na, nb)
        assert exec_.to_host(k).tobytes() == both[src].tobytes(), (na, nb)


@pytest.mark.parametrize("dtype,expr,key", LESS_CASES, ids=IDS)
def test_user_set_ops(exec_, dtype, expr, key):
    less = m.op.UserLess(expr)
    for na, nb in [(0, 0), (0, 9), (9, 0), (1000, 1500), (9000, 7000), (200_001, 150_000)]:
        a, b = sorted_by(rand(dtype, na), key), sorted_by(rand(dtype, nb), key)
        # Equal derived keys on both sides: overlap for the walk to pair.
        if na and nb:
            b = sorted_by(np.concatenate([b[: nb // 2], a[: nb - nb // 2]]), key)
        both = np.concatenate([a, b])
        da, db = exec_.to_device(a), exec_.to_device(b)
        for kind, fn in [("union", m.vector.set_union), ("intersection", m.vector.set_intersection),
                         ("difference", m.vector.set_difference)]:
            src = oracle.set_op_src(kind, key(a), key(b), "less")
            got = exec_.to_host(fn(exec_, da, db, less))
            assert got.tobytes() == both[src].tobytes(), (kind, na, nb)


def test_user_less_reference_vectors(exec_):
    """The reference's ordering tests with a user Less: boundary lengths with
    keys ((i * 1103515245 + 12345) % 37) (verification/oracle/tests/vector/
    ordering.rs:38-113), merge stability (core/merge.rs:1262-1309) and the
    multiset rules (core/set.rs:418-471)."""
    less = m.op.UserLess("a < b")
    for n in [0, 1, 2, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025, 2049]:
        keys = ((np.arange(n, dtype=np.uint64) * 1_103_515_245 + 12_345) % 37).astype(np.uint32)
        d = exec_.to_device(keys)
        assert exec_.to_host(m.vector.sort(exec_, d, less)).tolist() == sorted(keys.tolist())
        vals = exec_.to_device(np.arange(n, dtype=np.uint32))
        got = exec_.to_host(m.vector.sort_by_key(exec_, d, vals, less)).tolist()
        assert got == [i for _, i in sorted(zip(keys.tolist(), range(n)), key=lambda t: t[0])]
    v = golden("merge_stable")
    k, vals = m.vector.merge_by_key(
        exec_, exec_.to_device(np.array(v["left_keys"], np.uint32)),
        exec_.to_device(np.array(v["left_values"], np.uint32)),
        exec_.to_device(np.array(v["right_keys"], np.uint32)),
        exec_.to_device(np.array(v["right_values"], np.uint32)), less)
    assert exec_.to_host(k).tolist() == v["expected_keys"]
    assert exec_.to_host(vals).tolist() == v["expected_values"]
    v = golden("set_multiplicity")
    a = exec_.to_device(np.array(v["left"], np.uint32))
    b = exec_.to_device(np.array(v["right"], np.uint32))
    assert exec_.to_host(m.vector.set_union(exec_, a, b, less)).tolist() == v["expected_union"]
    assert exec_.to_host(m.vector.set_intersection(exec_, a, b, less)).tolist() == v["expected_intersection"]
    assert exec_.to_host(m.vector.set_difference(exec_, a, b, less)).tolist() == v["expected_difference"]


def _offsets(lengths):
    return np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint32)


@pytest.mark.parametrize("dtype,expr,key", [LESS_CASES[1], LESS_CASES[3], LESS_CASES[4]],
                         ids=[IDS[1], IDS[3], IDS[4]])
def test_user_seg_sort(exec_, dtype, expr, key):
    less = m.op.UserLess(expr)
    lengths = RNG.integers(0, 300, size=5000)
    lengths[RNG.integers(0, 5000, size=100)] = 0
    lengths[7] = 20_000
    off = _offsets(lengths)
    n = int(off[-1])
    x = rand(dtype, n)
    vals = rand(np.uint32, n)
    perm = oracle.seg_sort_perm(key(x), off, "less")
    seg = m.seg.SegmentIterator((exec_.to_device(x), exec_.to_device(vals)), exec_.to_device(off))
    out = m.seg.ForEachSegment(m.seg.Sort(less)).run(exec_, seg)
    k, v = out.values
    assert exec_.to_host(k).tobytes() == x[perm].tobytes()
    assert np.array_equal(exec_.to_host(v), vals[perm])
    # keys only
    seg = m.seg.SegmentIterator(exec_.to_device(x), exec_.to_device(off))
    k = m.seg.ForEachSegment(m.seg.Sort(less)).run(exec_, seg).values
    assert exec_.to_host(k).tobytes() == x[perm].tobytes()


# (key dtype, Equal expression, derived key with the built-in equality)
EQ_CASES = [
    (np.uint32, "a / 10u == b / 10u", lambda x: x // np.uint32(10)),
    (np.float32, "floorf(a) == floorf(b)", lambda x: np.floor(x)),
    (np.int64, "(a >> 3) == (b >> 3)", lambda x: x >> 3),
]


def run_keys(dtype, n):
    """Keys with runs: a sorted draw from a narrow range, so neighbours are
    often equal under the derived key."""
    if np.dtype(dtype).kind == "f":
        return np.sort(RNG.uniform(-50, 50, n)).astype(dtype)
    return np.sort(RNG.integers(0, max(n // 3, 1) * 10, size=n)).astype(dtype)


@pytest.mark.parametrize("dtype,expr,key", EQ_CASES, ids=[e for _, e, _ in EQ_CASES])
def test_user_equal_by_key(exec_, dtype, expr, key):
    equal = m.op.UserEqual(expr)
    for n in [0, 1, 2, 1000, 4097, 70001, 1_000_003]:
        kx = run_keys(dtype, n)
        g = key(kx)
        vals = RNG.integers(0, 1000, size=n).astype(np.uint32)
        dk, dv = exec_.to_device(kx), exec_.to_device(vals)
        got = exec_.to_host(m.vector.inclusive_scan_by_key(exec_, dk, dv, equal, m.op.Add))
        assert np.array_equal(got, oracle.inclusive_scan_by_key(g, vals, "add")), n
        got = exec_.to_host(m.vector.exclusive_scan_by_key(exec_, dk, dv, equal, 7, m.op.Max))
        assert np.array_equal(got, oracle.exclusive_scan_by_key(g, vals, 7, "max")), n
        rk, rv = m.vector.reduce_by_key(exec_, dk, dv, equal, 3, m.op.Add)
        heads = oracle.unique_by_key(g, np.arange(n, dtype=np.uint32))
        _, want = oracle.reduce_by_key(g, vals, 3, "add")
        assert exec_.to_host(rk).tobytes() == kx[heads].tobytes(), n
        assert np.array_equal(exec_.to_host(rv), want), n
        got = exec_.to_host(m.vector.unique_by_key(exec_, dk, dv, equal))
        assert np.array_equal(got, oracle.unique_by_key(g, vals)), n


def test_user_equal_reference_vectors(exec_):
    """The reference's by-key doc examples (api/algorithm/by_key.rs:366-397,
    586-617, 730-757) and seg::Unique (tests/seg/api.rs:285-313) with a user
    Equal that is the numeric equality."""
    eq = m.op.UserEqual("a == b")
    v = golden("inclusive_scan_by_key_doc")
    k, x = exec_.to_device(np.array(v["keys"], np.uint32)), exec_.to_device(np.array(v["values"], np.uint32))
    assert exec_.to_host(m.vector.inclusive_scan_by_key(exec_, k, x, eq, m.op.Add)).tolist() == v["expected"]
    v = golden("reduce_by_key_doc")
    k, x = exec_.to_device(np.array(v["keys"], np.uint32)), exec_.to_device(np.array(v["values"], np.uint32))
    rk, rv = m.vector.reduce_by_key(exec_, k, x, eq, v["init"], m.op.Add)
    assert exec_.to_host(rk).tolist() == v["expected_keys"] and exec_.to_host(rv).tolist() == v["expected"]
    v = golden("unique_by_key_doc")
    k, x = exec_.to_device(np.array(v["keys"], np.uint32)), exec_.to_device(np.array(v["values"], np.uint32))
    assert exec_.to_host(m.vector.unique_by_key(exec_, k, x, eq)).tolist() == v["expected"]
    v = golden("seg_unique_owned")
    seg = m.seg.SegmentIterator(exec_.to_device(np.array(v["input"], np.uint32)),
                                exec_.to_device(np.array(v["offsets"], np.uint32)))
    out = m.seg.ForEachSegment(m.seg.Unique(eq)).run(exec_, seg)
    assert exec_.to_host(out.values).tolist() == v["expected"]
    assert exec_.to_host(out.offsets.offsets()).tolist() == v["expected_offsets"]


@pytest.mark.parametrize("dtype,expr,key", EQ_CASES, ids=[e for _, e, _ in EQ_CASES])
def test_user_seg_unique(exec_, dtype, expr, key):
    equal = m.op.UserEqual(expr)
    lengths = RNG.integers(0, 200, size=3000)
    lengths[RNG.integers(0, 3000, size=60)] = 0
    lengths[-1] = 0
    off = _offsets(lengths)
    n = int(off[-1])
    x = run_keys(dtype, n)
    g = key(x)
    seg_id = np.repeat(np.arange(len(lengths), dtype=np.uint64), lengths)
    # Kept positions: the heads of the runs of (segment, derived key).
    keyed = (seg_id << np.uint64(32)) | (np.unique(g, return_inverse=True)[1].astype(np.uint64))
    pos = oracle.unique_by_key(keyed, np.arange(n, dtype=np.uint32))
    _, want_off = oracle.seg_unique(np.unique(g, return_inverse=True)[1].astype(np.uint32), off)
    seg = m.seg.SegmentIterator(exec_.to_device(x), exec_.to_device(off))
    out = m.seg.ForEachSegment(m.seg.Unique(equal)).run(exec_, seg)
    assert exec_.to_host(out.values).tobytes() == x[pos].tobytes()
    assert np.array_equal(exec_.to_host(out.offsets.offsets()), want_off)


def test_user_comparator_errors(exec_):
    less = m.op.UserLess("a < b")
    eq = m.op.UserEqual("a == b")
    x = exec_.to_device(np.arange(10, dtype=np.uint32))
    with pytest.raises(m.error.InvalidArgument):
        m.vector.sort(exec_, x, m.op.UserLess("a <"))  # does not compile
    with pytest.raises(m.error.InvalidArgument):
        # a user Equal is not an ordering
        m.vector.merge(exec_, x, x, eq)
    with pytest.raises(m.LengthMismatch):
        m.vector.sort_by_key(exec_, x, exec_.to_device(np.arange(9, dtype=np.uint32)), less)
    with pytest.raises(m.LengthMismatch):
        m.vector.unique_by_key(exec_, x, exec_.to_device(np.arange(9, dtype=np.uint32)), eq)
