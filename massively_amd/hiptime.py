"""HIP event timing on an explicit stream (the executor's own stream).

torch.cuda.Event only records on torch's current stream; the engine launches
on the stream its executor owns, so timing goes through hipEvent* directly.
"""
from __future__ import annotations

import ctypes as C

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
    return _hip


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        if hip().hipEventCreate(C.byref(self.h)) != 0:
            raise RuntimeError("hipEventCreate failed")

    def record(self, stream: int) -> None:
        if hip().hipEventRecord(self.h, C.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def synchronize(self) -> None:
        hip().hipEventSynchronize(self.h)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        if hip().hipEventElapsedTime(C.byref(ms), self.h, end.h) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)

    def __del__(self):
        try:
            hip().hipEventDestroy(self.h)
        except Exception:
            pass
