"""``massively::seg``: per-segment algorithms over CSR offsets.

``ForEachSegment(Alg).run(exec, SegmentIterator(values, offsets))`` as in
massively/src/seg/mod.rs:315-330. Offsets are validated into a
:class:`Segmentation` (seg/segmentation.rs:187-264) before any kernel runs.

Served algorithms: ``InclusiveScan``, ``ExclusiveScan``, ``Reduce``, ``Sort``
and the compacting ``Filter``, ``Unique`` and ``Take`` (seg/mod.rs:597-880),
which return the kept values with a new :class:`Segmentation`. ``Sort``
accepts either a key column or a ``(keys, values)`` pair, in which case rows
are ordered by key and the values move with them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .error import check
from .op import BinaryPredicateOp, PredicateOp, ReductionOp, UserEqual, UserLess
from .runtime import DeviceSlice, DeviceVec, Executor, NULL_SLICE, as_slice

_lib = L.lib


class Segmentation:
    """A validated ordered partition (``seg::Segmentation``)."""

    def __init__(self, exec: Executor, handle: C.c_void_p):
        self.executor = exec
        self._h = handle

    @staticmethod
    def from_offsets(exec: Executor, offsets) -> "Segmentation":
        h = C.c_void_p()
        check(_lib().msv_segmentation_from_offsets(exec.handle, as_slice(offsets)._c(),
                                                   C.byref(h)))
        return Segmentation(exec, h)

    @staticmethod
    def from_lengths(exec: Executor, lengths) -> "Segmentation":
        h = C.c_void_p()
        check(_lib().msv_segmentation_from_lengths(exec.handle, as_slice(lengths)._c(),
                                                   C.byref(h)))
        return Segmentation(exec, h)

    @property
    def handle(self):
        return self._h

    def offsets(self) -> DeviceVec:
        h = C.c_void_p()
        check(_lib().msv_segmentation_offsets(self._h, C.byref(h)))
        return DeviceVec(self.executor, h)

    def segment_count(self) -> int:
        return int(_lib().msv_segmentation_segment_count(self._h))

    def value_count(self) -> int:
        return int(_lib().msv_segmentation_value_count(self._h))

    def segments(self, values) -> "SegmentIterator":
        return SegmentIterator(values, self)

    def __del__(self):
        try:
            if self._h:
                _lib().msv_segmentation_destroy(self._h)
                self._h = None
        except Exception:
            pass


@dataclass
class SegmentIterator:
    """A flat value stream and the offsets delimiting its segments (seg/mod.rs:48-74)."""

    values: object
    offsets: object

    def into_parts(self):
        return self.values, self.offsets

    def segmentation(self, exec: Executor) -> Segmentation:
        if isinstance(self.offsets, Segmentation):
            return self.offsets
        return Segmentation.from_offsets(exec, self.offsets)


@dataclass(frozen=True)
class InclusiveScan:
    op: ReductionOp


@dataclass(frozen=True)
class ExclusiveScan:
    op: ReductionOp
    init: object


@dataclass(frozen=True)
class Reduce:
    op: ReductionOp
    init: object


@dataclass(frozen=True)
class Sort:
    less: BinaryPredicateOp


@dataclass(frozen=True)
class Filter:
    pred: PredicateOp


@dataclass(frozen=True)
class Unique:
    """Unique(equal): the built-in numeric equality (None or op.Equal) or a
    UserEqual."""

    equal: object = None


@dataclass(frozen=True)
class Take:
    count: int


class ForEachSegment:
    def __init__(self, algorithm):
        self.algorithm = algorithm

    def run(self, exec: Executor, segments: SegmentIterator):
        alg = self.algorithm
        seg = segments.segmentation(exec)
        values = segments.values
        if isinstance(alg, Sort):
            if isinstance(values, tuple):
                keys, payload = as_slice(values[0]), as_slice(values[1])
                pc = payload._c()
            else:
                keys, payload, pc = as_slice(values), None, NULL_SLICE
            hk, hv = C.c_void_p(), C.c_void_p()
            if isinstance(alg.less, UserLess):
                check(_lib().msv_seg_sort_user(exec.handle, keys._c(), pc, seg.handle,
                                               alg.less.handle(exec, keys.dtype), C.byref(hk),
                                               C.byref(hv) if payload is not None else None))
            else:
                check(_lib().msv_seg_sort(exec.handle, keys._c(), pc, seg.handle, alg.less.code,
                                          C.byref(hk), C.byref(hv) if payload is not None else None))
            k = DeviceVec(exec, hk)
            out = (k, DeviceVec(exec, hv)) if payload is not None else k
            return SegmentIterator(out, seg)
        s = as_slice(values)
        h = C.c_void_p()
        if isinstance(alg, InclusiveScan):
            check(_lib().msv_seg_inclusive_scan(exec.handle, s._c(), seg.handle, alg.op.code,
                                                C.byref(h)))
            return SegmentIterator(DeviceVec(exec, h), seg)
        if isinstance(alg, ExclusiveScan):
            init = np.array([alg.init], dtype=s.dtype)
            check(_lib().msv_seg_exclusive_scan(exec.handle, s._c(), seg.handle,
                                                init.ctypes.data_as(C.c_void_p), alg.op.code,
                                                C.byref(h)))
            return SegmentIterator(DeviceVec(exec, h), seg)
        if isinstance(alg, Reduce):
            init = np.array([alg.init], dtype=s.dtype)
            check(_lib().msv_seg_reduce(exec.handle, s._c(), seg.handle,
                                        init.ctypes.data_as(C.c_void_p), alg.op.code,
                                        C.byref(h)))
            return DeviceVec(exec, h)
        if isinstance(alg, (Filter, Unique, Take)):
            hs = C.c_void_p()
            if isinstance(alg, Filter):
                p = alg.pred.c_struct(s.dtype)
                check(_lib().msv_seg_filter(exec.handle, s._c(), seg.handle, C.byref(p),
                                            C.byref(h), C.byref(hs)))
            elif isinstance(alg, Unique) and isinstance(alg.equal, UserEqual):
                check(_lib().msv_seg_unique_user(exec.handle, s._c(), seg.handle,
                                                 alg.equal.handle(exec, s.dtype), C.byref(h),
                                                 C.byref(hs)))
            elif isinstance(alg, Unique):
                check(_lib().msv_seg_unique(exec.handle, s._c(), seg.handle, C.byref(h),
                                            C.byref(hs)))
            else:
                check(_lib().msv_seg_take(exec.handle, s._c(), seg.handle, int(alg.count),
                                          C.byref(h), C.byref(hs)))
            return SegmentIterator(DeviceVec(exec, h), Segmentation(exec, hs))
        raise TypeError(f"unsupported segment algorithm {alg!r}")
