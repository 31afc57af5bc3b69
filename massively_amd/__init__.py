"""massively_amd: MI355X (gfx950) engine for massively's vector and segment primitives.

Host-side mirror of the reference's public API (``Executor``, ``DeviceVec``,
``DeviceSlice``, ``vector::*``, ``seg::*``, ``Error``) over the C ABI of
``libmsv.so`` (include/msv.h). All computation runs in hand-written HIP
kernels; importing this package does not load the library, but every
operation raises if it is missing.
"""
from . import lazy, op, seg, vector
from .error import (
    Error,
    ForeignExecutor,
    InvalidSegmentation,
    LengthMismatch,
    LengthTooLarge,
    OutputTooShort,
    UnboundColumn,
    Unsupported,
)
from .runtime import (DeviceSlice, DeviceVec, Executor, batch_tuned_grid, segsort_configs,
                      segsort_tuned_config, tuning_clear, tuning_count, tuning_load, tuning_save)

__all__ = [
    "Executor",
    "DeviceVec",
    "DeviceSlice",
    "Error",
    "LengthMismatch",
    "OutputTooShort",
    "LengthTooLarge",
    "InvalidSegmentation",
    "UnboundColumn",
    "ForeignExecutor",
    "Unsupported",
    "batch_tuned_grid",
    "segsort_configs",
    "segsort_tuned_config",
    "tuning_clear",
    "tuning_count",
    "tuning_load",
    "tuning_save",
    "lazy",
    "op",
    "seg",
    "vector",
]
