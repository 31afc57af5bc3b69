"""GPU parity at the full sizes of the vector-primitive and movement
workloads: f32 reduce of 2^28 N(0,1) values against the FP64 sum, gather and
scatter through a seeded random permutation of 2^28, and merge_by_key of two
sorted 2^27 u32 arrays with duplicates inside each array (tie order at
scale). Mirrors the reference's scale tests
(verification/oracle/tests/vector_scale/common.rs:9,131-142)."""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import random as mr

pytestmark = pytest.mark.gpu


def test_reduce_f32_normal_2e28_within_fp64_bound_and_repeatable(exec_):
    n = 1 << 28
    hx = mr.normal_f32_host(exec_, 3, n)
    d = exec_.to_device(hx)
    r1 = m.vector.reduce(exec_, d, 0.0, m.op.Add)
    r2 = m.vector.reduce(exec_, d, 0.0, m.op.Add)
    assert np.float32(r1).tobytes() == np.float32(r2).tobytes()
    s, a = oracle.sum_f32_f64(hx)
    bound = oracle.f32_sum_bound(n, a)
    assert abs(float(r1) - s) <= bound, (float(r1), s, bound)
    # The sample is N(0, 1): sum |x| / n is E|z| = sqrt(2 / pi) within a few
    # standard errors (a generator that drifted would fail here).
    assert abs(a / n - np.sqrt(2 / np.pi)) < 1e-3


def test_gather_scatter_random_permutation_2e28(exec_):
    n = 1 << 28
    keys = exec_.alloc(np.uint32, n)
    mr.uniform_u32(exec_, 0, 2**32 - 1, 9, keys)
    iota = exec_.alloc(np.uint32, n)
    mr.iota_u32(exec_, iota)
    idx = m.vector.radix_sort_by_key(exec_, keys, iota)  # a seeded random permutation
    del keys, iota
    v = exec_.alloc(np.uint32, n)
    mr.uniform_u32(exec_, 0, 2**32 - 1, 19, v)
    hv, hi = exec_.to_host(v), exec_.to_host(idx)
    # A permutation of [0, n): every index in range, and the index sum and
    # xor match those of 0..n-1.
    assert int(hi.max()) == n - 1
    assert int(hi.astype(np.uint64).sum()) == n * (n - 1) // 2
    assert int(np.bitwise_xor.reduce(hi)) == 0
    got = exec_.to_host(m.vector.gather(exec_, v, idx))
    assert np.array_equal(got, hv[hi])
    del got
    out = exec_.alloc(np.uint32, n)
    m.vector.scatter(exec_, v, idx, out)
    exp = np.empty(n, np.uint32)
    exp[hi] = hv
    assert np.array_equal(exec_.to_host(out), exp)


def test_merge_by_key_2x2e27_with_duplicates(exec_):
    n = 1 << 27
    a = exec_.alloc(np.uint32, n)
    mr.uniform_u32(exec_, 0, (1 << 26) - 1, 11, a)
    b = exec_.alloc(np.uint32, n)
    mr.uniform_u32(exec_, 0, (1 << 26) - 1, 12, b)
    sa, sb = m.vector.sort(exec_, a, m.op.Less), m.vector.sort(exec_, b, m.op.Less)
    del a, b
    av = exec_.alloc(np.uint32, n)
    mr.iota_u32(exec_, av, 0)
    bv = exec_.alloc(np.uint32, n)
    mr.iota_u32(exec_, bv, n)
    k, v = m.vector.merge_by_key(exec_, sa, av, sb, bv, m.op.Less)
    hsa, hsb = exec_.to_host(sa), exec_.to_host(sb)
    ek, ev = oracle.merge_by_key(hsa, np.arange(n, dtype=np.uint32), hsb,
                                 np.arange(n, 2 * n, dtype=np.uint32))
    assert np.array_equal(exec_.to_host(k), ek)
    assert np.array_equal(exec_.to_host(v), ev)
    # Ties: every equal-key run lists the left side's positions first, each
    # side in input order (values = source positions, so they increase).
    hk, hv = exec_.to_host(k), exec_.to_host(v)
    same = hk[1:] == hk[:-1]
    assert np.all(hv[1:][same] > hv[:-1][same])
