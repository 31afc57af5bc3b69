"""Error type mirroring ``massively::Error`` (massively/src/core/error.rs:5-39).

Each variant is a subclass of :class:`Error`; equality compares the variant and
its fields, as the reference's ``#[derive(Eq, PartialEq)]`` does.
"""
from __future__ import annotations

from . import _lib as L


class Error(Exception):
    fields: tuple = ()

    def __eq__(self, other):
        return type(self) is type(other) and all(
            getattr(self, f) == getattr(other, f) for f in self.fields
        )

    def __hash__(self):
        return hash((type(self).__name__,) + tuple(getattr(self, f) for f in self.fields))


class LengthMismatch(Error):
    fields = ("left", "right")

    def __init__(self, left: int, right: int):
        super().__init__(f"input lengths differ: {left} != {right}")
        self.left, self.right = left, right


class OutputTooShort(Error):
    fields = ("input", "output")

    def __init__(self, input: int, output: int):
        super().__init__(f"output length {output} is shorter than input length {input}")
        self.input, self.output = input, output


class LengthTooLarge(Error):
    fields = ("len",)

    def __init__(self, len: int):
        super().__init__(f"length does not fit in u32: {len}")
        self.len = len


class InvalidSegmentation(Error):
    def __init__(self):
        super().__init__("segmentation is not a valid ordered partition")


class UnboundColumn(Error):
    def __init__(self):
        super().__init__("column is not bound to device storage")


class ForeignExecutor(Error):
    def __init__(self):
        super().__init__("executor does not own this device data")


class Launch(Error):
    fields = ("message",)

    def __init__(self, message: str):
        super().__init__(f"launch failed: {message}")
        self.message = message


class InvalidArgument(Error):
    fields = ("message",)

    def __init__(self, message: str):
        super().__init__(message)
        self.message = message


class Unsupported(Error):
    fields = ("message",)

    def __init__(self, message: str):
        super().__init__(message)
        self.message = message


class OutOfMemory(Error):
    fields = ("message",)

    def __init__(self, message: str):
        super().__init__(message)
        self.message = message


def from_status(code: int) -> Error:
    e = L.last_error()
    msg = e.message.decode(errors="replace")
    if code == L.ERR_LENGTH_MISMATCH:
        return LengthMismatch(int(e.left), int(e.right))
    if code == L.ERR_OUTPUT_TOO_SHORT:
        return OutputTooShort(int(e.left), int(e.right))
    if code == L.ERR_LENGTH_TOO_LARGE:
        return LengthTooLarge(int(e.left))
    if code == L.ERR_INVALID_SEGMENTATION:
        return InvalidSegmentation()
    if code == L.ERR_UNBOUND_COLUMN:
        return UnboundColumn()
    if code == L.ERR_FOREIGN_EXECUTOR:
        return ForeignExecutor()
    if code == L.ERR_LAUNCH:
        return Launch(msg)
    if code == L.ERR_UNSUPPORTED:
        return Unsupported(msg)
    if code == L.ERR_OUT_OF_MEMORY:
        return OutOfMemory(msg)
    return InvalidArgument(msg or f"status {code}")


def check(code: int) -> None:
    if code != L.OK:
        raise from_status(code)
