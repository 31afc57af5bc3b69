"""CPU oracle for massively's vector and segment primitives.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker. The product
library never calls it.

The numeric work is a C restatement of the reference's sequential oracle
(``msv_oracle.c``, which cites verification/oracle/src/{vector,seg}.rs line by
line); this module wraps it for numpy arrays. How each function is pinned
(golden vectors of the reference's tests and docs, numpy library routines,
closed forms, brute force) is listed in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

DT = {
    np.dtype(np.uint32): 0,
    np.dtype(np.int32): 1,
    np.dtype(np.float32): 2,
    np.dtype(np.uint64): 3,
    np.dtype(np.int64): 4,
    np.dtype(np.float64): 5,
}
OPS = {"add": 0, "min": 1, "max": 2, "xor": 3}
ORD = {"total": 0, "less": 1, "greater": 2}

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P, U64, I = C.c_void_p, C.c_uint64, C.c_int
        L.msvo_inclusive_scan.argtypes = [I, I, P, P, U64]
        L.msvo_exclusive_scan.argtypes = [I, I, P, P, U64, P]
        L.msvo_reduce.argtypes = [I, I, P, U64, P, P]
        L.msvo_sum_f32_f64.argtypes = [P, U64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.msvo_copy_where.argtypes = [I, P, P, U64, I, P]
        L.msvo_copy_where.restype = U64
        L.msvo_copy_where_pred.argtypes = [I, P, I, P, U64, I, U64, U64, I, P]
        L.msvo_copy_where_pred.restype = U64
        L.msvo_gather.argtypes = [I, P, P, U64, P]
        L.msvo_scatter.argtypes = [I, P, P, U64, P]
        L.msvo_order_word.argtypes = [I, P, I]
        L.msvo_order_word.restype = U64
        L.msvo_sort_perm.argtypes = [I, I, P, U64, P]
        L.msvo_sort_by_key.argtypes = [I, I, I, P, P, U64, P, P]
        L.msvo_merge_src.argtypes = [I, I, P, U64, P, U64, P]
        L.msvo_set_op.argtypes = [I, I, I, P, U64, P, U64, P]
        L.msvo_set_op.restype = U64
        L.msvo_set_op_src.argtypes = [I, I, I, P, U64, P, U64, P]
        L.msvo_set_op_src.restype = U64
        L.msvo_seg_inclusive_scan.argtypes = [I, I, P, P, P, U64]
        L.msvo_seg_exclusive_scan.argtypes = [I, I, P, P, P, U64, P]
        L.msvo_seg_reduce.argtypes = [I, I, P, P, U64, P, P]
        L.msvo_seg_sort_perm.argtypes = [I, I, P, P, U64, P]
        L.msvo_inclusive_scan_by_key.argtypes = [I, I, I, P, P, P, U64]
        L.msvo_exclusive_scan_by_key.argtypes = [I, I, I, P, P, P, U64, P]
        L.msvo_reduce_by_key.argtypes = [I, I, I, P, P, U64, P, P, P]
        L.msvo_reduce_by_key.restype = C.c_int64
        L.msvo_unique_by_key.argtypes = [I, I, P, P, U64, P]
        L.msvo_unique_by_key.restype = C.c_int64
        L.msvo_partition.argtypes = [I, P, U64, I, U64, U64, P]
        L.msvo_partition.restype = U64
        L.msvo_scatter_reduce.argtypes = [I, I, P, P, U64, P, P, U64]
        L.msvo_seg_compact.argtypes = [I, I, P, P, U64, I, U64, U64, U64, P, P]
        L.msvo_seg_compact.restype = U64
        L.msvo_pcg_hash32.argtypes = [C.c_uint32]
        L.msvo_pcg_hash32.restype = C.c_uint32
        L.msvo_seed_key.argtypes = [U64]
        L.msvo_seed_key.restype = C.c_uint32
        L.msvo_uniform_u32.argtypes = [C.c_uint32, C.c_uint32, U64, U64, U64, P]
        L.msvo_uniform_u64.argtypes = [U64, U64, U64, U64, U64, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _c(a) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype not in DT:
        raise TypeError(f"unsupported dtype {a.dtype}")
    return a


def inclusive_scan(x, op: str) -> np.ndarray:
    x = _c(x)
    y = np.empty_like(x)
    lib().msvo_inclusive_scan(DT[x.dtype], OPS[op], _p(x), _p(y), x.size)
    return y


def exclusive_scan(x, init, op: str) -> np.ndarray:
    x = _c(x)
    y = np.empty_like(x)
    i = np.array([init], dtype=x.dtype)
    lib().msvo_exclusive_scan(DT[x.dtype], OPS[op], _p(x), _p(y), x.size, _p(i))
    return y


def reduce(x, init, op: str):
    x = _c(x)
    i = np.array([init], dtype=x.dtype)
    o = np.zeros(1, dtype=x.dtype)
    lib().msvo_reduce(DT[x.dtype], OPS[op], _p(x), x.size, _p(i), _p(o))
    return o[0]


def sum_f32_f64(x) -> tuple[float, float]:
    """FP64 sum and sum of magnitudes of an f32 array."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    s, a = C.c_double(), C.c_double()
    lib().msvo_sum_f32_f64(_p(x), x.size, C.byref(s), C.byref(a))
    return s.value, a.value


def f32_sum_bound(n: int, abs_sum: float) -> float:
    """Stated bound for an f32 sum of n terms: (ceil(log2 n) + 2) * 2^-24 * sum|x|."""
    import math

    levels = math.ceil(math.log2(max(n, 2)))
    return (levels + 2) * 2.0**-24 * abs_sum


def copy_where(x, stencil, invert: bool = False) -> np.ndarray:
    x = _c(x)
    s = np.ascontiguousarray(stencil, dtype=np.uint32)
    out = np.empty_like(x)
    k = lib().msvo_copy_where(DT[x.dtype], _p(x), _p(s), x.size, int(invert), _p(out))
    return out[:k].copy()


def remove_where(x, stencil) -> np.ndarray:
    return copy_where(x, stencil, invert=True)


def copy_where_pred(x, stencil, kind: int, a_bits: int = 0, b_bits: int = 0,
                    invert: bool = False) -> np.ndarray:
    x, s = _c(x), _c(stencil)
    out = np.empty_like(x)
    k = lib().msvo_copy_where_pred(DT[x.dtype], _p(x), DT[s.dtype], _p(s), x.size, kind,
                                   a_bits, b_bits, int(invert), _p(out))
    return out[:k].copy()


def gather(values, indices) -> np.ndarray:
    v = _c(values)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    out = np.empty(idx.size, dtype=v.dtype)
    lib().msvo_gather(DT[v.dtype], _p(v), _p(idx), idx.size, _p(out))
    return out


def scatter(values, indices, output) -> np.ndarray:
    v = _c(values)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    out = np.array(output, dtype=v.dtype, copy=True)
    lib().msvo_scatter(DT[v.dtype], _p(v), _p(idx), v.size, _p(out))
    return out


def order_words(keys, mode: str) -> np.ndarray:
    k = _c(keys)
    es = k.dtype.itemsize
    base = k.ctypes.data
    L = lib()
    return np.array([L.msvo_order_word(DT[k.dtype], C.c_void_p(base + i * es), ORD[mode])
                     for i in range(k.size)], dtype=np.uint64)


def sort_perm(keys, mode: str) -> np.ndarray:
    k = _c(keys)
    perm = np.empty(k.size, dtype=np.uint32)
    lib().msvo_sort_perm(DT[k.dtype], ORD[mode], _p(k), k.size, _p(perm))
    return perm


def radix_sort_by_key(keys, values):
    """Returns (sorted keys, permuted values) in total order, stable."""
    k, v = _c(keys), _c(values)
    ok, ov = np.empty_like(k), np.empty_like(v)
    lib().msvo_sort_by_key(DT[k.dtype], DT[v.dtype], ORD["total"], _p(k), _p(v), k.size,
                           _p(ok), _p(ov))
    return ok, ov


def sort_by_key(keys, values, mode: str = "less"):
    k, v = _c(keys), _c(values)
    ok, ov = np.empty_like(k), np.empty_like(v)
    lib().msvo_sort_by_key(DT[k.dtype], DT[v.dtype], ORD[mode], _p(k), _p(v), k.size,
                           _p(ok), _p(ov))
    return ok, ov


def sort(keys, mode: str = "less") -> np.ndarray:
    k = _c(keys)
    return k[sort_perm(k, mode)]


def merge_src(left, right, mode: str = "less") -> np.ndarray:
    a, b = _c(left), _c(right)
    src = np.empty(a.size + b.size, dtype=np.uint32)
    lib().msvo_merge_src(DT[a.dtype], ORD[mode], _p(a), a.size, _p(b), b.size, _p(src))
    return src


def merge(left, right, mode: str = "less") -> np.ndarray:
    a, b = _c(left), _c(right)
    return np.concatenate([a, b])[merge_src(a, b, mode)]


def merge_by_key(lk, lv, rk, rv, mode: str = "less"):
    src = merge_src(lk, rk, mode)
    return (np.concatenate([_c(lk), _c(rk)])[src], np.concatenate([_c(lv), _c(rv)])[src])


SET_KIND = {"union": 0, "intersection": 1, "difference": 2}


def set_op(kind: str, left, right, mode: str = "less") -> np.ndarray:
    a, b = _c(left), _c(right)
    out = np.empty(a.size + b.size, dtype=a.dtype)
    k = lib().msvo_set_op(DT[a.dtype], ORD[mode], SET_KIND[kind], _p(a), a.size, _p(b), b.size,
                          _p(out))
    return out[:k].copy()


def set_op_src(kind: str, left, right, mode: str = "less") -> np.ndarray:
    """Source of each output of set_op: an index into left, or len(left) + an
    index into right (the same walk, for orders given by derived keys)."""
    a, b = _c(left), _c(right)
    src = np.empty(a.size + b.size, dtype=np.uint32)
    k = lib().msvo_set_op_src(DT[a.dtype], ORD[mode], SET_KIND[kind], _p(a), a.size, _p(b), b.size,
                              _p(src))
    return src[:k].copy()


def _offs(offsets) -> np.ndarray:
    return np.ascontiguousarray(offsets, dtype=np.uint32)


def seg_inclusive_scan(x, offsets, op: str) -> np.ndarray:
    x, o = _c(x), _offs(offsets)
    y = np.empty_like(x)
    lib().msvo_seg_inclusive_scan(DT[x.dtype], OPS[op], _p(x), _p(y), _p(o), o.size - 1)
    return y


def seg_exclusive_scan(x, offsets, init, op: str) -> np.ndarray:
    x, o = _c(x), _offs(offsets)
    y = np.empty_like(x)
    i = np.array([init], dtype=x.dtype)
    lib().msvo_seg_exclusive_scan(DT[x.dtype], OPS[op], _p(x), _p(y), _p(o), o.size - 1, _p(i))
    return y


def seg_reduce(x, offsets, init, op: str) -> np.ndarray:
    x, o = _c(x), _offs(offsets)
    out = np.empty(o.size - 1, dtype=x.dtype)
    i = np.array([init], dtype=x.dtype)
    lib().msvo_seg_reduce(DT[x.dtype], OPS[op], _p(x), _p(o), o.size - 1, _p(i), _p(out))
    return out


def seg_sort_perm(keys, offsets, mode: str = "less") -> np.ndarray:
    k, o = _c(keys), _offs(offsets)
    perm = np.empty(k.size, dtype=np.uint32)
    lib().msvo_seg_sort_perm(DT[k.dtype], ORD[mode], _p(k), _p(o), o.size - 1, _p(perm))
    return perm


def inclusive_scan_by_key(keys, values, op: str) -> np.ndarray:
    k, v = _c(keys), _c(values)
    y = np.empty_like(v)
    lib().msvo_inclusive_scan_by_key(DT[k.dtype], DT[v.dtype], OPS[op], _p(k), _p(v), _p(y), v.size)
    return y


def exclusive_scan_by_key(keys, values, init, op: str) -> np.ndarray:
    k, v = _c(keys), _c(values)
    y = np.empty_like(v)
    i = np.array([init], dtype=v.dtype)
    lib().msvo_exclusive_scan_by_key(DT[k.dtype], DT[v.dtype], OPS[op], _p(k), _p(v), _p(y), v.size,
                                     _p(i))
    return y


def reduce_by_key(keys, values, init, op: str):
    k, v = _c(keys), _c(values)
    ok, ov = np.empty_like(k), np.empty_like(v)
    i = np.array([init], dtype=v.dtype)
    r = lib().msvo_reduce_by_key(DT[k.dtype], DT[v.dtype], OPS[op], _p(k), _p(v), v.size, _p(i),
                                 _p(ok), _p(ov))
    return ok[:r].copy(), ov[:r].copy()


def unique_by_key(keys, values) -> np.ndarray:
    k, v = _c(keys), _c(values)
    ov = np.empty_like(v)
    r = lib().msvo_unique_by_key(DT[k.dtype], DT[v.dtype], _p(k), _p(v), v.size, _p(ov))
    return ov[:r].copy()


def pcg_hash32(x: int) -> int:
    return int(lib().msvo_pcg_hash32(x & 0xFFFFFFFF))


def seed_key(seed: int) -> int:
    return int(lib().msvo_seed_key(seed))


def uniform_u32(lo: int, hi: int, seed: int, n: int, start: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().msvo_uniform_u32(lo, hi, seed, start, n, _p(out))
    return out


def uniform_u64(lo: int, hi: int, seed: int, n: int, start: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().msvo_uniform_u64(lo, hi, seed, start, n, _p(out))
    return out


def partition(x, kind: int, a_bits: int = 0, b_bits: int = 0):
    """(out, boundary): matching elements first, then failing ones."""
    x = _c(x)
    out = np.empty_like(x)
    k = lib().msvo_partition(DT[x.dtype], _p(x), x.size, kind, a_bits, b_bits, _p(out))
    return out, int(k)


def scatter_reduce(values, indices, init, op: str, output) -> np.ndarray:
    v = _c(values)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    out = np.array(output, dtype=v.dtype, copy=True)
    i = np.array([init], dtype=v.dtype)
    r = lib().msvo_scatter_reduce(DT[v.dtype], OPS[op], _p(v), _p(idx), v.size, _p(i), _p(out),
                                  out.size)
    assert r == 0
    return out


def _seg_compact(x, offsets, form: int, kind: int = 0, a_bits: int = 0, b_bits: int = 0,
                 take: int = 0):
    x, o = _c(x), _offs(offsets)
    out = np.empty_like(x)
    oo = np.empty(o.size, dtype=np.uint32)
    k = lib().msvo_seg_compact(DT[x.dtype], form, _p(x), _p(o), o.size - 1, kind, a_bits, b_bits,
                               take, _p(out), _p(oo))
    return out[:k].copy(), oo


def seg_filter(x, offsets, kind: int, a_bits: int = 0, b_bits: int = 0):
    """(values, offsets) of ForEachSegment(Filter(pred))."""
    return _seg_compact(x, offsets, 0, kind, a_bits, b_bits)


def seg_unique(x, offsets):
    """(values, offsets) of ForEachSegment(Unique(Equal))."""
    return _seg_compact(x, offsets, 1)


def seg_take(x, offsets, k: int):
    """(values, offsets) of ForEachSegment(Take(k))."""
    return _seg_compact(x, offsets, 2, take=k)


def compound_sort_perm(columns) -> np.ndarray:
    """Stable lexicographic row order over up to 3 key columns, leftmost
    primary, each column in integer / IEEE total order (core/radix.rs:499-531)."""
    words = [order_words(col, "total") for col in columns]
    return np.lexsort(tuple(reversed(words))).astype(np.uint32)
