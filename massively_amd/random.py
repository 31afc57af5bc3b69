"""Seeded synthetic inputs generated on the device.

Bit-identical to the reference's counter-based PCG streams
(massively/src/api/util/random.rs:12-22,331-396): element i of
``uniform_u32(min, max, seed)`` is ``min + pcg_hash32(i ^ seed_key(seed)) %
(max - min + 1)`` (the raw hash when the span is the whole range).

``normal_f32`` is N(0, 1) by the Box-Muller transform in f64 on the host over
the full-range ``uniform_u64(seed)`` stream, rounded to f32. The reference's
own normal_f32 (random.rs:398-405) evaluates f32 ``ln``/``cos`` on the device,
whose last bits depend on the GPU's math library; computing in f64 and
rounding once fixes every bit from the integer stream alone.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .error import check
from .runtime import Executor, as_slice


def uniform_u32(exec: Executor, lo: int, hi: int, seed: int, out) -> None:
    check(L.lib().msv_random_uniform_u32(exec.handle, lo, hi, seed, as_slice(out)._c()))


def uniform_u64(exec: Executor, lo: int, hi: int, seed: int, out) -> None:
    check(L.lib().msv_random_uniform_u64(exec.handle, lo, hi, seed, as_slice(out)._c()))


def iota_u32(exec: Executor, out, start: int = 0) -> None:
    check(L.lib().msv_iota_u32(exec.handle, start, as_slice(out)._c()))


def normal_from_u64(bits: np.ndarray) -> np.ndarray:
    """Box-Muller over 64-bit words: u1 = (lo + 1) / 2^32 in (0, 1], u2 = hi /
    2^32 in [0, 1); z = sqrt(-2 ln u1) cos(2 pi u2) in f64, rounded to f32."""
    bits = np.asarray(bits, dtype=np.uint64)
    lo = (bits & np.uint64(0xFFFFFFFF)).astype(np.float64)
    hi = (bits >> np.uint64(32)).astype(np.float64)
    u1 = (lo + 1.0) * 2.0**-32
    u2 = hi * 2.0**-32
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).astype(np.float32)


def normal_f32_host(exec: Executor, seed: int, n: int, chunk: int = 1 << 24) -> np.ndarray:
    """n values N(0, 1) from the seed (see the module docstring), on the host."""
    out = np.empty(n, dtype=np.float32)
    bits = exec.alloc(np.uint64, n) if n else None
    if n:
        uniform_u64(exec, 0, 2**64 - 1, seed, bits)
        for lo in range(0, n, chunk):
            hi = min(n, lo + chunk)
            out[lo:hi] = normal_from_u64(exec.to_host(bits.slice(lo, hi)))
    return out


def normal_f32(exec: Executor, seed: int, n: int):
    """A new device vector of n f32 values N(0, 1) from the seed."""
    return exec.to_device(normal_f32_host(exec, seed, n))
