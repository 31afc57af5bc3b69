"""Batched jobs and the serving loop (no counterpart in the reference).

``run_jobs(exec, jobs)`` enqueues many independent scan / reduce / copy_where /
sort_by_key jobs over caller-owned device vectors in one call (small jobs share
one persistent launch per job family). ``Server`` is the request loop of
``msv_server_*``: requests are admitted in arrival order by an HBM budget,
their inputs are created on the device, admitted requests are batched
continuously over several streams, and outputs stay device-resident until
fetched. Two forms the batched mode is measured against:
``Server(mode="serial")`` runs one request at a time through the device-wide
kernels, and ``Server(mode="per_job")`` runs one request per launch sequence
through the device-wide kernels on every stream (concurrency without
batching).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .error import check
from .op import BinaryPredicateOp, PredicateOp, ReductionOp

INCLUSIVE_SCAN, EXCLUSIVE_SCAN, REDUCE, COPY_WHERE, SORT_BY_KEY = 1, 2, 3, 4, 5
KINDS = {"inclusive_scan": INCLUSIVE_SCAN, "exclusive_scan": EXCLUSIVE_SCAN, "reduce": REDUCE,
         "copy_where": COPY_WHERE, "sort_by_key": SORT_BY_KEY}
ORDER_TOTAL, ORDER_LESS, ORDER_GREATER = 0, 1, 2


def _order(cmp) -> int:
    if cmp is None:
        return ORDER_TOTAL
    if isinstance(cmp, BinaryPredicateOp):
        return ORDER_LESS if cmp.code == 0 else ORDER_GREATER
    return int(cmp)


def _bits(v, dtype) -> int:
    a = np.array([v], dtype=np.dtype(dtype))
    return int(a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)[0])


@dataclass
class Job:
    """One job of ``run_jobs`` over device vectors (outputs preallocated).

    kind: "inclusive_scan" | "exclusive_scan" | "reduce" | "copy_where" | "sort_by_key".
    For sorts `inp` are the keys, `inp2` the values (or None), `out2` the
    sorted keys and `out` the sorted values; for copy_where `out2` is a u32
    vector of length 1 that receives the count; for reduce `out` has length 1.
    """

    kind: str
    inp: object
    out: object = None
    op: object = None            # ReductionOp for scans / reduce, Less / Greater / None for sorts
    init: object = 0
    pred: PredicateOp | None = None
    invert: bool = False
    inp2: object = None
    out2: object = None

    def c_struct(self) -> L.MsvJob:
        j = L.MsvJob()
        j.kind = KINDS[self.kind]
        dt = self.inp.dtype
        j.dtype = L.DTYPE_OF_NP[np.dtype(dt)]
        j.n = len(self.inp)
        if j.kind == SORT_BY_KEY:
            j.op = _order(self.op)
            j.value_bytes = 0 if self.inp2 is None else np.dtype(self.inp2.dtype).itemsize
        else:
            j.op = self.op.code if isinstance(self.op, ReductionOp) else int(self.op or 0)
        j.invert = 1 if self.invert else 0
        if self.pred is not None:
            j.pred = self.pred.c_struct(dt)
        j.init = _bits(self.init, dt)
        j.in_ = self.inp.device_ptr
        j.in2 = None if self.inp2 is None else self.inp2.device_ptr
        j.out = None if self.out is None else self.out.device_ptr
        j.out2 = None if self.out2 is None else self.out2.device_ptr
        return j


def run_jobs(exec, jobs) -> None:
    arr = (L.MsvJob * max(1, len(jobs)))(*[j.c_struct() for j in jobs])
    check(L.lib().msv_run_jobs(exec.handle, arr, len(jobs)))


@dataclass
class Request:
    kind: str
    dtype: object
    n: int
    seed: int = 0
    op: object = None
    init: object = 0
    pred: PredicateOp | None = None
    invert: bool = False
    with_values: bool = True      # sorts: u32 payload = input position
    host_input: np.ndarray | None = field(default=None, repr=False)

    def c_struct(self) -> L.MsvRequest:
        r = L.MsvRequest()
        r.kind = KINDS[self.kind]
        dt = np.dtype(self.dtype)
        r.dtype = L.DTYPE_OF_NP[dt]
        r.n = int(self.n)
        if r.kind == SORT_BY_KEY:
            r.op = _order(self.op)
            r.value_bytes = 4 if self.with_values else 0
        else:
            r.op = self.op.code if isinstance(self.op, ReductionOp) else int(self.op or 0)
        r.invert = 1 if self.invert else 0
        if self.pred is not None:
            r.pred = self.pred.c_struct(dt)
        r.init = _bits(self.init, dt)
        r.seed = int(self.seed)
        r.host_input = None
        return r


class Server:
    """``msv_server``: continuous batching ("batched"), one request at a time
    through the device-wide kernels ("serial"), or one request per launch
    sequence on every stream ("per_job")."""

    def __init__(self, device: int = 0, streams: int = 4, max_batch: int = 256,
                 mode: str = "batched", hbm_budget: int = 64 << 30):
        self._lib = L.lib()
        modes = {"batched": 0, "serial": 1, "per_job": 2}
        if mode not in modes:
            raise ValueError(f"mode must be one of {sorted(modes)}")
        cfg = L.MsvServerConfig(device, streams, max_batch, modes[mode], hbm_budget)
        h = C.c_void_p()
        check(self._lib.msv_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._reqs: dict[int, Request] = {}

    def close(self) -> None:
        if self._h is not None:
            check(self._lib.msv_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req: Request) -> int:
        r = req.c_struct()
        keep = None
        if req.host_input is not None:
            keep = np.ascontiguousarray(req.host_input, dtype=np.dtype(req.dtype))
            assert keep.size == req.n
            r.host_input = keep.ctypes.data_as(C.c_void_p)
        t = C.c_uint64()
        check(self._lib.msv_server_submit(self._h, C.byref(r), C.byref(t)))
        self._reqs[t.value] = req
        return t.value

    def wait(self, ticket: int) -> dict:
        info = L.MsvRequestInfo()
        check(self._lib.msv_server_wait(self._h, ticket, C.byref(info)))
        if info.state == 3:
            check(info.status)
        return dict(state=info.state, submit_us=info.submit_us, admit_us=info.admit_us,
                    done_us=info.done_us, out_len=info.out_len, bytes=info.bytes)

    def result(self, ticket: int, out_len: int | None = None):
        """Host copies of a finished request's outputs: the elements (scan,
        copy_where), the value (reduce), or (sorted keys, sorted values)."""
        req = self._reqs[ticket]
        if out_len is None:
            out_len = self.wait(ticket)["out_len"]
        dt = np.dtype(req.dtype)
        main = np.empty(out_len, dtype=dt)
        vals = None
        if req.kind == "sort_by_key" and req.with_values:
            vals = np.empty(req.n, dtype=np.uint32)
        check(self._lib.msv_server_result(
            self._h, ticket, main.ctypes.data_as(C.c_void_p), main.nbytes,
            None if vals is None else vals.ctypes.data_as(C.c_void_p), 0 if vals is None else vals.nbytes))
        if req.kind == "reduce":
            return main[0]
        return (main, vals) if req.kind == "sort_by_key" else main

    def release(self, ticket: int) -> None:
        check(self._lib.msv_server_release(self._h, ticket))
        self._reqs.pop(ticket, None)

    def drain(self) -> None:
        check(self._lib.msv_server_drain(self._h))

    def now_us(self) -> float:
        return float(self._lib.msv_server_now_us(self._h))
