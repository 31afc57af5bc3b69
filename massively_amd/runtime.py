"""Executor / DeviceVec / DeviceSlice (massively/src/core/runtime.rs:21-545).

``Executor`` owns one HIP stream on one device. ``DeviceVec`` owns a
refcounted device allocation; ``clone()`` shares it (runtime.rs:275-285).
``DeviceSlice`` is a zero-copy ``{vec, offset, len}`` view (runtime.rs:208-272).
Vectors produced by length-changing algorithms (copy_where, set operations)
carry a device-resident logical length (core/extent.rs:17-36); reading it
synchronises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .error import check


class Executor:
    """``Executor::new(device)``: one device, one stream, a caching allocator."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = C.c_void_p()
        check(self._lib.msv_executor_create(int(device), C.byref(h)))
        self._h = h
        self.device = int(device)
        # Operators compiled at run time for lazy inputs, by source (lazy.py).
        self._jit_ops = {}

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("executor is closed")
        return self._h

    @property
    def id(self) -> int:
        return int(self._lib.msv_executor_id(self.handle))

    @property
    def stream(self) -> int:
        return int(self._lib.msv_executor_stream(self.handle) or 0)

    def close(self) -> None:
        for h in getattr(self, "_jit_ops", {}).values():
            self._lib.msv_user_op_destroy(h)
        self._jit_ops = {}
        if self._h is not None:
            self._lib.msv_executor_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        check(self._lib.msv_executor_sync(self.handle))

    def trim(self) -> None:
        check(self._lib.msv_executor_trim(self.handle))

    def capture(self) -> "GraphCapture":
        """``with ex.capture() as g: ...`` records the library calls made
        inside the block into a HIP graph; ``g.graph.launch()`` replays them."""
        return GraphCapture(self)

    # Tuning options of the batching layer (results never depend on them).
    OPT_SEG_LAUNCH, OPT_AUTOTUNE, OPT_SEGSORT_CONFIG, OPT_RADIX_ONESWEEP, OPT_SEGSORT_WAVE = 1, 2, 3, 4, 5
    SEG_LAUNCH_PERSISTENT, SEG_LAUNCH_PER_CLASS = 0, 1

    def set_option(self, opt: int, value: int) -> None:
        check(self._lib.msv_executor_set_option(self.handle, opt, value))

    def get_option(self, opt: int) -> int:
        v = C.c_int64()
        check(self._lib.msv_executor_get_option(self.handle, opt, C.byref(v)))
        return v.value


    def alloc(self, dtype, length: int) -> "DeviceVec":
        dt = L.DTYPE_OF_NP[np.dtype(dtype)]
        h = C.c_void_p()
        check(self._lib.msv_alloc(self.handle, dt, _u32(length), C.byref(h)))
        return DeviceVec(self, h)

    def to_device(self, values, dtype=None) -> "DeviceVec":
        arr = np.ascontiguousarray(values, dtype=dtype)
        if arr.dtype not in L.DTYPE_OF_NP:
            raise TypeError(f"unsupported dtype {arr.dtype}")
        arr = arr.reshape(-1)
        h = C.c_void_p()
        ptr = arr.ctypes.data_as(C.c_void_p) if arr.size else None
        check(
            self._lib.msv_to_device(
                self.handle, L.DTYPE_OF_NP[arr.dtype], ptr, _u32(arr.size), C.byref(h)
            )
        )
        return DeviceVec(self, h)

    def to_host(self, x) -> np.ndarray:
        s = as_slice(x)
        v = s.vec
        cap = v.capacity - s.offset if s.length == L.WHOLE else s.length
        out = np.empty(max(cap, 0), dtype=v.dtype)
        n = C.c_uint32()
        ptr = out.ctypes.data_as(C.c_void_p) if out.size else None
        check(self._lib.msv_to_host(self.handle, s._c(), ptr, _u32(out.size), C.byref(n)))
        return out[: n.value].copy() if n.value != out.size else out

    def wrap(self, device_ptr: int, dtype, length: int) -> "DeviceVec":
        """Borrows caller-owned device memory (e.g. a torch tensor's storage)."""
        h = C.c_void_p()
        check(
            self._lib.msv_wrap_device(
                self.handle, L.DTYPE_OF_NP[np.dtype(dtype)], C.c_void_p(device_ptr),
                _u32(length), C.byref(h),
            )
        )
        return DeviceVec(self, h)


def _u32(n: int) -> int:
    n = int(n)
    if n < 0 or n > 0xFFFFFFFF:
        from .error import LengthTooLarge

        raise LengthTooLarge(n)
    return n


class DeviceVec:
    """``DeviceVec<R, T>`` / ``MVec``: a refcounted device allocation."""

    def __init__(self, executor: Executor, handle: C.c_void_p):
        self.executor = executor
        self._h = handle
        self._lib = executor._lib

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def dtype(self) -> np.dtype:
        return L.NP_DTYPES[int(self._lib.msv_vec_dtype(self._h))]

    @property
    def capacity(self) -> int:
        return int(self._lib.msv_vec_capacity(self._h))

    @property
    def device_ptr(self) -> int:
        return int(self._lib.msv_vec_device_ptr(self._h) or 0)

    @property
    def has_device_len(self) -> bool:
        return bool(self._lib.msv_vec_device_len_ptr(self._h))

    def len(self) -> int:
        n = C.c_uint32()
        check(self._lib.msv_vec_len(self.executor.handle, self._h, C.byref(n)))
        return int(n.value)

    def __len__(self) -> int:
        return self.len()

    def clone(self) -> "DeviceVec":
        """Shallow clone sharing the allocation (runtime.rs:275-285)."""
        check(self._lib.msv_vec_retain(self._h))
        return DeviceVec(self.executor, self._h)

    def slice(self, start: int = 0, end: int | None = None) -> "DeviceSlice":
        if end is None:
            return DeviceSlice(self, start, L.WHOLE)
        if end < start:
            raise ValueError(f"slice end {end} < start {start}")
        return DeviceSlice(self, start, end - start)

    def __del__(self):
        try:
            if self._h:
                self._lib.msv_vec_release(self._h)
                self._h = None
        except Exception:
            pass


class DeviceSlice:
    """``DeviceSlice`` / ``DeviceSliceMut``: a view ``{vec, offset, len}``."""

    __slots__ = ("vec", "offset", "length")

    def __init__(self, vec: DeviceVec, offset: int, length: int):
        self.vec = vec
        self.offset = int(offset)
        self.length = int(length)

    @property
    def dtype(self) -> np.dtype:
        return self.vec.dtype

    @property
    def executor(self) -> Executor:
        return self.vec.executor

    def _c(self) -> L.MsvSlice:
        c = L.MsvSlice(self.vec.handle, self.offset, self.length)
        # The struct keeps the vector alive while it is a call's argument (a
        # temporary vector, e.g. a written-out lazy input, has no other owner).
        c._keep = self.vec
        return c

    def slice(self, start: int = 0, end: int | None = None) -> "DeviceSlice":
        base_len = self.vec.len() - self.offset if self.length == L.WHOLE else self.length
        end = base_len if end is None else end
        if not (0 <= start <= end <= base_len):
            raise ValueError(f"bad sub-slice {start}..{end} of length {base_len}")
        return DeviceSlice(self.vec, self.offset + start, end - start)


def as_slice(x) -> DeviceSlice:
    if isinstance(x, DeviceSlice):
        return x
    if isinstance(x, DeviceVec):
        return x.slice()
    raise TypeError(f"expected DeviceVec or DeviceSlice, got {type(x).__name__}")


NULL_SLICE = L.MsvSlice(None, 0, 0)


def segsort_configs():
    """The seg::Sort tile-shape candidates: [(block, items_per_thread, digit_bits)]."""
    lib = L.lib()
    out = []
    for k in range(lib.msv_segsort_config_count()):
        b, i, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib.msv_segsort_config_info(k, C.byref(b), C.byref(i), C.byref(d)))
        out.append((b.value, i.value, d.value))
    return out


class Graph:
    """A captured sequence of library calls (``msv_graph``)."""

    def __init__(self, executor: Executor, handle: C.c_void_p):
        self.executor = executor
        self._h = handle
        self._lib = executor._lib

    def launch(self) -> None:
        check(self._lib.msv_graph_launch(self._h))

    def close(self) -> None:
        if self._h is not None:
            check(self._lib.msv_graph_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GraphCapture:
    def __init__(self, executor: Executor):
        self.executor = executor
        self.graph: Graph | None = None

    def __enter__(self) -> "GraphCapture":
        check(self.executor._lib.msv_executor_begin_capture(self.executor.handle))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = C.c_void_p()
        st = self.executor._lib.msv_executor_end_capture(self.executor.handle, C.byref(h))
        if exc_type is None:
            check(st)
            self.graph = Graph(self.executor, h)
        elif st == 0:
            Graph(self.executor, h).close()
        return False


def segsort_tuned_config(key_bytes: int, value_bytes: int, launch: int, n: int, nseg: int,
                         windows: bool = False) -> int:
    """Index of the seg::Sort candidate the autotuner chose on the current device
    for this shape (-1: not tuned); ``windows`` for the window-packing form."""
    f = L.lib().msv_segsort_tuned_config_windows if windows else L.lib().msv_segsort_tuned_config
    return int(f(key_bytes, value_bytes, launch, n, nseg))


def batch_tuned_grid(family: int, jobs: int, mean_len: int) -> int:
    """Grid candidate the autotuner chose on the current device for batched
    job-kernel family ``family`` (0 = scan / reduce / copy_where, 1..6 =
    sort_by_key by key and value width) in a batch of ``jobs`` jobs of mean
    length ``mean_len`` (-1: not tuned). Candidate c launches the family's
    resident blocks per CU divided by 2**c."""
    return int(L.lib().msv_batch_tuned_grid(family, jobs, mean_len))


def tuning_save(path: str) -> None:
    """Writes the autotuner's table (every device's entries) to ``path``."""
    check(L.lib().msv_tuning_save(str(path).encode()))


def tuning_load(path: str) -> None:
    """Merges the entries of a file written by :func:`tuning_save`."""
    check(L.lib().msv_tuning_load(str(path).encode()))


def tuning_clear() -> None:
    check(L.lib().msv_tuning_clear())


def tuning_count() -> int:
    return int(L.lib().msv_tuning_count())
