"""GPU parity of msv_inclusive_scan_noalloc, the inclusive scan that allocates
nothing (its scratch is a static per-device array), against msv_inclusive_scan
through the C ABI: bit-exact for every dtype and operator, two calls in a row
with the second on a smaller input, words past the input length untouched,
graph replays, and the host-side rejections (short or overlapping output,
dtype mismatch, inputs whose scratch would exceed the static array)."""
import ctypes as C

import numpy as np
import pytest

import massively_amd as m
from massively_amd import _lib as L
from massively_amd import hiptime

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)
OPS = {"add": 0, "min": 1, "max": 2, "xor": 3}


def rand(dtype, n):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return RNG.standard_normal(n).astype(dt)
    info = np.iinfo(dt)
    return RNG.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)


def plain_scan(ex, d, op):
    h = C.c_void_p()
    assert L.load().msv_inclusive_scan(ex.handle, d.slice()._c(), op, C.byref(h)) == 0, L.last_error().message
    return ex.to_host(m.DeviceVec(ex, h))


def noalloc_scan(ex, d, op, out, offset=0):
    s = out.slice(offset, offset + len(d))._c()
    st = L.load().msv_inclusive_scan_noalloc(ex.handle, d.slice()._c(), op, s)
    assert st == 0, L.last_error().message


def h2d(dptr, arr):
    """Host-to-device copy into existing device memory (hipMemcpy, synchronous)."""
    a = np.ascontiguousarray(arr)
    st = hiptime.hip().hipMemcpy(C.c_void_p(dptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1)
    assert st == 0, st


@pytest.mark.parametrize("dtype,sizes", [
    (np.uint32, (16_777_233, 70_001)),   # large tiles, then small tiles
    (np.int32, (4_194_304, 4097)),
    (np.uint64, (8_388_617, 2049)),
    (np.int64, (300_001, 1)),
    (np.float32, (3_000_001, 4097)),
    (np.float64, (1_048_583, 2048)),
])
def test_noalloc_matches_inclusive_scan(exec_, dtype, sizes):
    ops = ["add", "min", "max"] + ([] if np.dtype(dtype).kind == "f" else ["xor"])
    cap = sizes[0] + 64
    out = exec_.alloc(dtype, cap)  # one output buffer for every call
    for op in ops:
        for n in sizes:  # two calls in a row, the second on a smaller input
            d = exec_.to_device(rand(dtype, n))
            sentinel = rand(dtype, cap)
            exec_.sync()
            h2d(out.device_ptr, sentinel)
            noalloc_scan(exec_, d, OPS[op], out, offset=5)
            got = exec_.to_host(out)
            want = plain_scan(exec_, d, OPS[op])
            assert got[5:5 + n].tobytes() == want.tobytes(), (dtype, op, n)
            assert got[:5].tobytes() == sentinel[:5].tobytes(), (dtype, op, n)
            assert got[5 + n:].tobytes() == sentinel[5 + n:].tobytes(), (dtype, op, n)


def test_noalloc_u32_at_2_pow_28(exec_):
    from massively_amd import random as mr

    n = 1 << 28
    x = exec_.alloc(np.uint32, n)
    mr.uniform_u32(exec_, 0, 1023, 2, x)
    out = exec_.alloc(np.uint32, n)
    noalloc_scan(exec_, x, OPS["add"], out)
    want = plain_scan(exec_, x, OPS["add"])
    got = exec_.to_host(out)
    assert np.array_equal(got, want)
    # Second call, smaller, into the same output.
    y = exec_.to_device(rand(np.uint32, 1 << 20))
    noalloc_scan(exec_, y, OPS["xor"], out)
    assert np.array_equal(exec_.to_host(out)[:1 << 20], plain_scan(exec_, y, OPS["xor"]))


def test_noalloc_graph_replays(exec_):
    n = 1 << 20
    xs = [rand(np.uint32, n) for _ in range(3)]
    d = exec_.to_device(xs[0])
    out = exec_.alloc(np.uint32, n)
    other = exec_.to_device(rand(np.uint32, 5000))
    other_out = exec_.alloc(np.uint32, 5000)
    exec_.sync()
    with exec_.capture() as g:
        noalloc_scan(exec_, d, OPS["add"], out)
    graph = g.graph
    for x in xs:
        exec_.sync()
        h2d(d.device_ptr, x)
        graph.launch()
        exec_.sync()
        assert np.array_equal(exec_.to_host(out), plain_scan(exec_, d, OPS["add"]))
        # Eager calls between the replays use the same static array.
        noalloc_scan(exec_, other, OPS["max"], other_out)
        assert np.array_equal(exec_.to_host(other_out), plain_scan(exec_, other, OPS["max"]))
    graph.close()


def test_noalloc_rejections(exec_):
    lib = L.load()
    x = exec_.to_device(rand(np.uint32, 1000))
    short = exec_.alloc(np.uint32, 999)
    assert lib.msv_inclusive_scan_noalloc(exec_.handle, x.slice()._c(), 0, short.slice()._c()) == L.ERR_OUTPUT_TOO_SHORT
    e = L.last_error()
    assert (e.left, e.right) == (1000, 999)
    f = exec_.alloc(np.float32, 1000)
    assert lib.msv_inclusive_scan_noalloc(exec_.handle, x.slice()._c(), 0, f.slice()._c()) == L.ERR_INVALID_ARGUMENT
    both = exec_.alloc(np.uint32, 1500)
    assert lib.msv_inclusive_scan_noalloc(exec_.handle, both.slice(0, 1000)._c(), 0,
                                          both.slice(500, 1500)._c()) == L.ERR_INVALID_ARGUMENT
    assert b"overlap" in L.last_error().message
    fx, fo = exec_.to_device(rand(np.float32, 10)), exec_.alloc(np.float32, 10)
    assert lib.msv_inclusive_scan_noalloc(exec_.handle, fx.slice()._c(), OPS["xor"],
                                          fo.slice()._c()) == L.ERR_INVALID_ARGUMENT
    # f64 scans of more than 2^29 elements need more than the 4 MiB of static scratch.
    n = (1 << 29) + 1
    big = exec_.alloc(np.float64, 2 * n)
    assert lib.msv_inclusive_scan_noalloc(exec_.handle, big.slice(0, n)._c(), 0,
                                          big.slice(n, 2 * n)._c()) == L.ERR_LENGTH_TOO_LARGE
    e = L.last_error()
    assert e.left == n and b"static scratch" in e.message
    del big
    exec_.trim()
