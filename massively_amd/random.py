"""Seeded synthetic inputs generated on the device.

Bit-identical to the reference's counter-based PCG streams
(massively/src/api/util/random.rs:12-22,331-396): element i of
``uniform_u32(min, max, seed)`` is ``min + pcg_hash32(i ^ seed_key(seed)) %
(max - min + 1)`` (the raw hash when the span is the whole range).
"""
from __future__ import annotations

from . import _lib as L
from .error import check
from .runtime import Executor, as_slice


def uniform_u32(exec: Executor, lo: int, hi: int, seed: int, out) -> None:
    check(L.lib().msv_random_uniform_u32(exec.handle, lo, hi, seed, as_slice(out)._c()))


def uniform_u64(exec: Executor, lo: int, hi: int, seed: int, out) -> None:
    check(L.lib().msv_random_uniform_u64(exec.handle, lo, hi, seed, as_slice(out)._c()))


def iota_u32(exec: Executor, out, start: int = 0) -> None:
    check(L.lib().msv_iota_u32(exec.handle, start, as_slice(out)._c()))
