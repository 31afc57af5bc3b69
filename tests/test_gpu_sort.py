"""GPU parity: radix_sort_by_key / sort_by_key / sort (LSD radix), bit-exact.

Stable ascending order in integer / IEEE total order (radix_sort_by_key) or by
the built-in comparators (Less/Greater; -0.0 == +0.0). NaN keys are excluded
from the comparator forms (the reference's merge sort leaves them unordered).
"""
import numpy as np
import pytest

import massively_amd as m
import oracle
from massively_amd import random as mr
from tests.conftest import GOLDEN, NP, golden, golden_of, keys_of

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(5)
LENGTHS = [1, 2, 31, 32, 33, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025, 4095, 4096, 4097,
           8193, 65537, 1 << 20]
KEY_DT = [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64]


def rand_keys(dtype, n, distinct=None):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        k = RNG.standard_normal(n).astype(dt)
        if distinct:
            k = np.round(k * distinct).astype(dt)
            k[::5] = -0.0
        return k
    info = np.iinfo(dt)
    if distinct:
        return RNG.integers(-distinct if dt.kind == "i" else 0, distinct, size=n).astype(dt)
    return RNG.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)


def test_golden_radix(exec_):
    for v in golden_of("radix_sort_by_key"):
        keys = keys_of(v)
        vals = np.array(v["values"], np.uint32)
        out = exec_.to_host(m.vector.radix_sort_by_key(exec_, exec_.to_device(keys), exec_.to_device(vals)))
        exp = vals[np.argsort(keys, kind="stable")] if v["expected"] == "stable_argsort" else v["expected"]
        assert out.tolist() == list(exp), v["name"]
    v = golden("radix_empty_and_mismatch")
    e = exec_.to_device(np.zeros(0, np.uint32))
    assert exec_.to_host(m.vector.radix_sort_by_key(exec_, e, e)).size == 0
    with pytest.raises(m.LengthMismatch) as err:
        m.vector.radix_sort_by_key(exec_, exec_.to_device(np.array(v["mismatch_keys"], np.uint32)),
                                   exec_.to_device(np.array(v["mismatch_values"], np.uint32)))
    assert err.value == m.LengthMismatch(2, 1)


@pytest.mark.parametrize("dtype", KEY_DT)
def test_radix_sort_by_key_total_order(exec_, dtype):
    for n in LENGTHS:
        for distinct in (None, 7):
            keys = rand_keys(dtype, n, distinct)
            vals = np.arange(n, dtype=np.uint32)
            v, k = m.vector.radix_sort_by_key(exec_, exec_.to_device(keys), exec_.to_device(vals),
                                              return_keys=True)
            ok, ov = oracle.radix_sort_by_key(keys, vals)
            assert np.array_equal(exec_.to_host(v), ov), (n, distinct)
            assert exec_.to_host(k).tobytes() == ok.tobytes(), (n, distinct)


@pytest.mark.parametrize("dtype", KEY_DT)
@pytest.mark.parametrize("mode", ["less", "greater"])
def test_sort_by_key_comparators(exec_, dtype, mode):
    cmp = m.op.Less if mode == "less" else m.op.Greater
    for n in [1, 33, 4097, 100003]:
        keys = rand_keys(dtype, n, 5)
        vals = RNG.integers(0, 2**64 - 1, n, dtype=np.uint64)
        v, k = m.vector.sort_by_key(exec_, exec_.to_device(keys), exec_.to_device(vals), cmp,
                                    return_keys=True)
        ok, ov = oracle.sort_by_key(keys, vals, mode)
        assert np.array_equal(exec_.to_host(v), ov), n
        assert exec_.to_host(k).tobytes() == ok.tobytes(), n
        s = exec_.to_host(m.vector.sort(exec_, exec_.to_device(keys), cmp))
        assert s.tobytes() == oracle.sort(keys, mode).tobytes(), n


def test_pass_skipping_constant_digits(exec_):
    n = 300000
    # Top 8 bits constant, and all keys equal (every digit constant).
    k = (np.uint64(0xA5) << np.uint64(56)) | RNG.integers(0, 2**56, n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    out = exec_.to_host(m.vector.radix_sort_by_key(exec_, exec_.to_device(k), exec_.to_device(vals)))
    assert np.array_equal(out, oracle.radix_sort_by_key(k, vals)[1])
    same = np.full(n, 12345, np.uint32)
    out = exec_.to_host(m.vector.radix_sort_by_key(exec_, exec_.to_device(same), exec_.to_device(vals)))
    assert np.array_equal(out, vals)
    lowconst = (RNG.integers(0, 2**24, n, dtype=np.uint32) << np.uint32(8)) | np.uint32(7)
    out = exec_.to_host(m.vector.radix_sort_by_key(exec_, exec_.to_device(lowconst), exec_.to_device(vals)))
    assert np.array_equal(out, oracle.radix_sort_by_key(lowconst, vals)[1])


@pytest.mark.parametrize("onesweep", [1, 0])
def test_radix_paths_agree(exec_, onesweep):
    """Both radix schedules (onesweep: histogram pass + look-back pass per
    digit; reduce-then-scan: upsweep + count scan + downsweep per digit)
    against the oracle, across tile counts and key / value widths."""
    exec_.set_option(m.Executor.OPT_RADIX_ONESWEEP, onesweep)
    try:
        for dtype in (np.uint32, np.float32, np.int64, np.float64):
            for n in [1, 4097, 65537, 1_000_003]:
                keys = rand_keys(dtype, n, 1000 if n > 70000 else None)
                for vals in (np.arange(n, dtype=np.uint32), RNG.integers(0, 2**63, n, dtype=np.uint64)):
                    v, k = m.vector.radix_sort_by_key(exec_, exec_.to_device(keys), exec_.to_device(vals),
                                                      return_keys=True)
                    ok, ov = oracle.radix_sort_by_key(keys, vals)
                    assert exec_.to_host(v).tobytes() == ov.tobytes(), (dtype, n, onesweep)
                    assert exec_.to_host(k).tobytes() == ok.tobytes(), (dtype, n, onesweep)
                s = exec_.to_host(m.vector.sort(exec_, exec_.to_device(keys), m.op.Greater))
                assert s.tobytes() == oracle.sort(keys, "greater").tobytes(), (dtype, n, onesweep)
    finally:
        exec_.set_option(m.Executor.OPT_RADIX_ONESWEEP, -1)


def test_regression_pairs(exec_):
    for r in GOLDEN["regressions"]:
        if r["kind"] != "pairs":
            continue
        p = np.array(r["data"], dtype=np.uint32)
        k, v = p[:, 0].copy(), p[:, 1].copy()
        out = exec_.to_host(m.vector.sort_by_key(exec_, exec_.to_device(k), exec_.to_device(v), m.op.Less))
        assert np.array_equal(out, oracle.sort_by_key(k, v)[1])


def test_plain_sort_by_key_2p20_u32_iota(exec_):
    """2^20 u32 keys uniform in [0,2^32) (seed 1), values iota: stable argsort."""
    n = 1 << 20
    keys = exec_.alloc(np.uint32, n)
    mr.uniform_u32(exec_, 0, 2**32 - 1, 1, keys)
    vals = exec_.alloc(np.uint32, n)
    mr.iota_u32(exec_, vals)
    out = exec_.to_host(m.vector.sort_by_key(exec_, keys, vals, m.op.Less))
    hk = exec_.to_host(keys)
    assert np.array_equal(out, oracle.sort_by_key(hk, np.arange(n, dtype=np.uint32))[1])


def _check_stable_perm(keys_h, vals_out):
    n = keys_h.size
    assert vals_out.size == n
    assert np.array_equal(np.sort(vals_out), np.arange(n, dtype=vals_out.dtype))
    ks = keys_h[vals_out]
    assert np.all(ks[1:] >= ks[:-1])
    ties = ks[1:] == ks[:-1]
    assert np.all(vals_out[1:][ties] > vals_out[:-1][ties])


@pytest.mark.parametrize("dist", ["uniform", "zipf", "top8const"])
def test_sort_by_key_u64_2p27(exec_, dist):
    """2^27 u64 keys + u32 payload (iota): sorted, stable permutation."""
    n = 1 << 27
    keys = exec_.alloc(np.uint64, n)
    mr.uniform_u64(exec_, 0, 2**64 - 1, 6, keys)
    hk = exec_.to_host(keys)
    if dist == "zipf":
        ranks = np.minimum(RNG.zipf(1.1, n), 1 << 20).astype(np.uint64)
        hk = (ranks * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64(0xD1B54A32D192ED03)
        keys = exec_.to_device(hk)
    elif dist == "top8const":
        hk = (np.uint64(0xA5) << np.uint64(56)) | (hk & np.uint64((1 << 56) - 1))
        keys = exec_.to_device(hk)
    vals = exec_.alloc(np.uint32, n)
    mr.iota_u32(exec_, vals)
    out = exec_.to_host(m.vector.radix_sort_by_key(exec_, keys, vals))
    _check_stable_perm(hk, out)


# ---- compound keys: zipped key columns (up to 3) and value columns (up to 12) ----

def test_radix_columns_golden(exec_):
    from tests.conftest import NP

    for name in ("radix_compound_keys_seven_values", "radix_three_keys_twelve_values"):
        g = golden(name)
        keys = tuple(exec_.to_device(np.array(k, NP[d])) for k, d in zip(g["keys"], g["key_dtypes"]))
        vals = tuple(exec_.to_device(np.array(v, np.uint32)) for v in g["values"])
        out = m.vector.radix_sort_by_key(exec_, keys, vals)
        assert [exec_.to_host(o).tolist() for o in out] == g["expected"], name


@pytest.mark.parametrize("kdts", [(np.uint32,), (np.uint32, np.uint32), (np.int32, np.uint32, np.uint32),
                                  (np.uint64, np.float32), (np.float64, np.int64, np.uint32)])
def test_radix_columns_match_oracle(exec_, kdts):
    rng = np.random.default_rng(len(kdts) * 7 + sum(np.dtype(d).itemsize for d in kdts))
    for n in (0, 1, 1000, 70001, 300000):
        cols = []
        for d in kdts:
            dt = np.dtype(d)
            if dt.kind == "f":
                c = rng.integers(-3, 4, n).astype(dt)  # many ties; +-0 and NaN below
                c[rng.random(n) < 0.05] = -0.0
                c[rng.random(n) < 0.02] = np.nan
            else:
                c = rng.integers(0, 5, n).astype(dt)
                if dt.kind == "i":
                    c -= 2
            cols.append(c)
        vals = [np.arange(n, dtype=np.uint32), rng.standard_normal(n)]
        perm = oracle.compound_sort_perm(cols)
        dk = tuple(exec_.to_device(c) if n else exec_.alloc(c.dtype, 0) for c in cols)
        dv = tuple(exec_.to_device(v) if n else exec_.alloc(v.dtype, 0) for v in vals)
        out = m.vector.radix_sort_by_key(exec_, dk, dv)
        for o, v in zip(out, vals):
            assert exec_.to_host(o).tobytes() == v[perm].tobytes(), (kdts, n)
