"""GPU parity: merge / merge_by_key and the multiset operations, bit-exact."""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import random as mr
from tests.conftest import golden

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(17)
SIZES = [(0, 0), (0, 1), (1, 0), (1, 1), (5, 3), (31, 33), (255, 257), (1024, 1024),
         (2047, 1), (1, 2049), (4096, 4097), (10000, 3), (65537, 40000), (300000, 250000)]


def sorted_keys(dtype, n, hi):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        k = np.round(RNG.standard_normal(n) * hi / 4).astype(dt)
        k[::9] = -0.0
        return np.sort(k, kind="stable")
    lo = -hi if dt.kind == "i" else 0
    return np.sort(RNG.integers(lo, hi, n).astype(dt), kind="stable")


def test_golden(exec_):
    v = golden("merge_stable")
    k, val = m.vector.merge_by_key(
        exec_, exec_.to_device(np.array(v["left_keys"], np.uint32)),
        exec_.to_device(np.array(v["left_values"], np.uint32)),
        exec_.to_device(np.array(v["right_keys"], np.uint32)),
        exec_.to_device(np.array(v["right_values"], np.uint32)), m.op.Less)
    assert exec_.to_host(k).tolist() == v["expected_keys"]
    assert exec_.to_host(val).tolist() == v["expected_values"]
    v = golden("set_multiplicity")
    a, b = exec_.to_device(np.array(v["left"], np.uint32)), exec_.to_device(np.array(v["right"], np.uint32))
    assert exec_.to_host(m.vector.set_union(exec_, a, b, m.op.Less)).tolist() == v["expected_union"]
    assert exec_.to_host(m.vector.set_intersection(exec_, a, b, m.op.Less)).tolist() == v["expected_intersection"]
    assert exec_.to_host(m.vector.set_difference(exec_, a, b, m.op.Less)).tolist() == v["expected_difference"]


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.float32, np.uint64, np.float64])
def test_merge_by_key(exec_, dtype):
    for na, nb in SIZES:
        for hi in (3, 1000):
            a, b = sorted_keys(dtype, na, hi), sorted_keys(dtype, nb, hi)
            av = np.arange(na, dtype=np.uint32)
            bv = np.arange(na, na + nb, dtype=np.uint32)
            k, v = m.vector.merge_by_key(exec_, exec_.to_device(a), exec_.to_device(av),
                                         exec_.to_device(b), exec_.to_device(bv), m.op.Less)
            ek, ev = oracle.merge_by_key(a, av, b, bv)
            assert exec_.to_host(k).tobytes() == ek.tobytes(), (na, nb, hi)
            assert np.array_equal(exec_.to_host(v), ev), (na, nb, hi)
            mk = exec_.to_host(m.vector.merge(exec_, exec_.to_device(a), exec_.to_device(b), m.op.Less))
            assert mk.tobytes() == oracle.merge(a, b).tobytes()


def test_merge_greater(exec_):
    a = np.sort(RNG.integers(0, 50, 5000).astype(np.uint32))[::-1].copy()
    b = np.sort(RNG.integers(0, 50, 3000).astype(np.uint32))[::-1].copy()
    got = exec_.to_host(m.vector.merge(exec_, exec_.to_device(a), exec_.to_device(b), m.op.Greater))
    assert np.array_equal(got, oracle.merge(a, b, "greater"))


@pytest.mark.parametrize("kind", ["union", "intersection", "difference"])
@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.float32])
def test_set_ops(exec_, kind, dtype):
    fn = {"union": m.vector.set_union, "intersection": m.vector.set_intersection,
          "difference": m.vector.set_difference}[kind]
    for na, nb in SIZES:
        for hi in (1, 2, 7, 100, 10**6):
            a, b = sorted_keys(dtype, na, hi), sorted_keys(dtype, nb, hi)
            y = fn(exec_, exec_.to_device(a), exec_.to_device(b), m.op.Less)
            exp = oracle.set_op(kind, a, b)
            assert y.has_device_len
            assert y.capacity == (na + nb if kind == "union" else na)
            assert exec_.to_host(y).tobytes() == exp.tobytes(), (na, nb, hi)


def test_set_ops_full_size_25pct_overlap(exec_):
    """A = sorted pcg_hash32(i), i < 2^27; B = sorted pcg_hash32(i) for i in
    [0.75*2^27, 1.75*2^27): exactly 25% overlap (pcg_hash32 is a bijection)."""
    n = 1 << 27
    ha = np.sort(oracle.uniform_u32(0, 2**32 - 1, 0, n))
    hb = np.sort(oracle.uniform_u32(0, 2**32 - 1, 0, n, start=3 * n // 4))
    a, b = exec_.to_device(ha), exec_.to_device(hb)
    inter = m.vector.set_intersection(exec_, a, b, m.op.Less)
    assert len(inter) == n // 4
    assert np.array_equal(exec_.to_host(inter), np.intersect1d(ha, hb))
    diff = m.vector.set_difference(exec_, a, b, m.op.Less)
    assert len(diff) == 3 * n // 4
    assert np.array_equal(exec_.to_host(diff), np.setdiff1d(ha, hb))
    uni = m.vector.set_union(exec_, a, b, m.op.Less)
    assert len(uni) == 7 * n // 4
    assert np.array_equal(exec_.to_host(uni), np.union1d(ha, hb))
    mg = exec_.to_host(m.vector.merge(exec_, a, b, m.op.Less))
    assert np.array_equal(mg, np.sort(np.concatenate([ha, hb])))
