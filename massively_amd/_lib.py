"""ctypes binding of libmsv.so (the C ABI declared in include/msv.h).

The library is loaded lazily and the product path fails loudly when it is
missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("MSV_LIBRARY", PKG / "libmsv.so"))

# msv_status
OK = 0
ERR_LENGTH_MISMATCH = 1
ERR_OUTPUT_TOO_SHORT = 2
ERR_LENGTH_TOO_LARGE = 3
ERR_INVALID_SEGMENTATION = 4
ERR_UNBOUND_COLUMN = 5
ERR_FOREIGN_EXECUTOR = 6
ERR_LAUNCH = 7
ERR_INVALID_ARGUMENT = 8
ERR_UNSUPPORTED = 9
ERR_OUT_OF_MEMORY = 10

# msv_dtype
U32, I32, F32, U64, I64, F64 = range(6)
NP_DTYPES = {
    U32: np.dtype(np.uint32),
    I32: np.dtype(np.int32),
    F32: np.dtype(np.float32),
    U64: np.dtype(np.uint64),
    I64: np.dtype(np.int64),
    F64: np.dtype(np.float64),
}
DTYPE_OF_NP = {v: k for k, v in NP_DTYPES.items()}

WHOLE = 0xFFFFFFFF


class MsvError(C.Structure):
    _fields_ = [
        ("code", C.c_int32),
        ("left", C.c_uint64),
        ("right", C.c_uint64),
        ("message", C.c_char * 256),
    ]


class MsvSlice(C.Structure):
    _fields_ = [("vec", C.c_void_p), ("offset", C.c_uint32), ("len", C.c_uint32)]


LAZY_SLOTS = 13
LAZY_PARAMS = 16


class MsvLazyArgs(C.Structure):
    _fields_ = [
        ("slots", MsvSlice * LAZY_SLOTS),
        ("nslots", C.c_uint32),
        ("len", C.c_uint32),
        ("params", C.c_uint64 * LAZY_PARAMS),
    ]


class MsvPredicate(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("reserved", C.c_uint32),
        ("a", C.c_uint64),
        ("b", C.c_uint64),
    ]


class MsvJob(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("dtype", C.c_uint32),
        ("op", C.c_uint32),
        ("n", C.c_uint32),
        ("value_bytes", C.c_uint32),
        ("invert", C.c_uint32),
        ("pred", MsvPredicate),
        ("init", C.c_uint64),
        ("in_", C.c_void_p),
        ("in2", C.c_void_p),
        ("out", C.c_void_p),
        ("out2", C.c_void_p),
    ]


class MsvServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("streams", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("mode", C.c_uint32),
        ("hbm_budget", C.c_uint64),
    ]


class MsvRequest(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("dtype", C.c_uint32),
        ("op", C.c_uint32),
        ("n", C.c_uint32),
        ("value_bytes", C.c_uint32),
        ("invert", C.c_uint32),
        ("pred", MsvPredicate),
        ("init", C.c_uint64),
        ("seed", C.c_uint64),
        ("host_input", C.c_void_p),
    ]


class MsvRequestInfo(C.Structure):
    _fields_ = [
        ("state", C.c_uint32),
        ("status", C.c_int32),
        ("submit_us", C.c_double),
        ("admit_us", C.c_double),
        ("done_us", C.c_double),
        ("out_len", C.c_uint32),
        ("reserved", C.c_uint32),
        ("bytes", C.c_uint64),
    ]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_S = MsvSlice
_I = C.c_int32
_U32 = C.c_uint32
_U64 = C.c_uint64

# name -> (restype, argtypes)
PROTOTYPES = {
    "msv_abi_version": (_U32, []),
    "msv_last_error": (None, [C.POINTER(MsvError)]),
    "msv_executor_create": (_I, [_I, _PP]),
    "msv_executor_destroy": (_I, [_P]),
    "msv_executor_sync": (_I, [_P]),
    "msv_executor_id": (_U64, [_P]),
    "msv_executor_stream": (_P, [_P]),
    "msv_executor_trim": (_I, [_P]),
    "msv_executor_set_option": (_I, [_P, _I, C.c_int64]),
    "msv_executor_get_option": (_I, [_P, _I, C.POINTER(C.c_int64)]),
    "msv_inclusive_scan_by_key": (_I, [_P, _S, _S, _I, _PP]),
    "msv_exclusive_scan_by_key": (_I, [_P, _S, _S, _P, _I, _PP]),
    "msv_reduce_by_key": (_I, [_P, _S, _S, _P, _I, _PP, _PP]),
    "msv_unique_by_key": (_I, [_P, _S, _S, _PP]),
    "msv_executor_begin_capture": (_I, [_P]),
    "msv_executor_end_capture": (_I, [_P, _PP]),
    "msv_graph_launch": (_I, [_P]),
    "msv_graph_destroy": (_I, [_P]),
    "msv_run_jobs": (_I, [_P, C.POINTER(MsvJob), _U32]),
    "msv_server_create": (_I, [C.POINTER(MsvServerConfig), _PP]),
    "msv_server_destroy": (_I, [_P]),
    "msv_server_submit": (_I, [_P, C.POINTER(MsvRequest), C.POINTER(_U64)]),
    "msv_server_wait": (_I, [_P, _U64, C.POINTER(MsvRequestInfo)]),
    "msv_server_result": (_I, [_P, _U64, _P, _U64, _P, _U64]),
    "msv_server_release": (_I, [_P, _U64]),
    "msv_server_drain": (_I, [_P]),
    "msv_server_now_us": (C.c_double, [_P]),
    "msv_segsort_config_count": (_U32, []),
    "msv_segsort_tuned_config": (_I, [_U32, _U32, C.c_int64, _U32, _U32]),
    "msv_batch_tuned_grid": (_I, [_U32, _U32, _U32]),
    "msv_segsort_tuned_config_windows": (_I, [_U32, _U32, C.c_int64, _U32, _U32]),
    "msv_tuning_save": (_I, [C.c_char_p]),
    "msv_tuning_load": (_I, [C.c_char_p]),
    "msv_tuning_clear": (_I, []),
    "msv_tuning_count": (_U32, []),
    "msv_segsort_config_info": (_I, [_U32, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32)]),
    "msv_alloc": (_I, [_P, _I, _U32, _PP]),
    "msv_to_device": (_I, [_P, _I, _P, _U32, _PP]),
    "msv_to_host": (_I, [_P, _S, _P, _U32, C.POINTER(_U32)]),
    "msv_wrap_device": (_I, [_P, _I, _P, _U32, _PP]),
    "msv_vec_retain": (_I, [_P]),
    "msv_vec_release": (_I, [_P]),
    "msv_vec_len": (_I, [_P, _P, C.POINTER(_U32)]),
    "msv_vec_capacity": (_U32, [_P]),
    "msv_vec_dtype": (_I, [_P]),
    "msv_vec_device_ptr": (_P, [_P]),
    "msv_vec_device_len_ptr": (_P, [_P]),
    "msv_inclusive_scan": (_I, [_P, _S, _I, _PP]),
    "msv_inclusive_scan_noalloc": (_I, [_P, _S, _I, _S]),
    "msv_exclusive_scan": (_I, [_P, _S, _P, _I, _PP]),
    "msv_reduce": (_I, [_P, _S, _P, _I, _P]),
    "msv_copy_where": (_I, [_P, _S, _S, _PP]),
    "msv_remove_where": (_I, [_P, _S, _S, _PP]),
    "msv_copy_where_pred": (_I, [_P, _S, _S, C.POINTER(MsvPredicate), _I, _PP]),
    "msv_gather": (_I, [_P, _S, _S, _PP]),
    "msv_scatter": (_I, [_P, _S, _S, _S]),
    "msv_scatter_reduce": (_I, [_P, _S, _S, _P, _I, _S]),
    "msv_partition": (_I, [_P, _S, _P, _PP, C.POINTER(C.c_uint32)]),
    "msv_user_op_create": (_I, [_P, C.c_int32, _I, C.c_char_p, C.c_char_p, _PP]),
    "msv_user_op_destroy": (_I, [_P]),
    "msv_reduce_user": (_I, [_P, _S, _P, _P, _P]),
    "msv_inclusive_scan_user": (_I, [_P, _S, _P, _PP]),
    "msv_exclusive_scan_user": (_I, [_P, _S, _P, _P, _PP]),
    "msv_copy_where_user": (_I, [_P, _S, _P, _PP]),
    "msv_lazy_op_create": (_I, [_P, C.c_int32, _I, C.c_char_p, C.c_int32, C.c_char_p, C.c_char_p, _PP]),
    "msv_lazy_op_check": (_I, [C.c_int32, _I, C.c_char_p, C.c_int32, C.c_char_p, C.c_char_p]),
    "msv_reduce_lazy": (_I, [_P, C.POINTER(MsvLazyArgs), _P, _P, _P]),
    "msv_inclusive_scan_lazy": (_I, [_P, C.POINTER(MsvLazyArgs), _P, _PP]),
    "msv_exclusive_scan_lazy": (_I, [_P, C.POINTER(MsvLazyArgs), _P, _P, _PP]),
    "msv_copy_where_lazy": (_I, [_P, C.POINTER(MsvLazyArgs), _P, C.POINTER(MsvPredicate), _PP]),
    "msv_map_lazy": (_I, [_P, C.POINTER(MsvLazyArgs), _P, _PP]),
    "msv_radix_sort_by_key_columns": (_I, [_P, C.POINTER(MsvSlice), C.c_uint32,
                                           C.POINTER(MsvSlice), C.c_uint32, _PP]),
    "msv_radix_sort_by_key": (_I, [_P, _S, _S, _PP, _PP]),
    "msv_gather_columns": (_I, [_P, C.POINTER(MsvSlice), C.c_uint32, _S, _PP]),
    "msv_scatter_columns": (_I, [_P, C.POINTER(MsvSlice), C.c_uint32, _S, C.POINTER(MsvSlice)]),
    "msv_copy_where_columns": (_I, [_P, C.POINTER(MsvSlice), C.c_uint32, _S,
                                    C.POINTER(MsvPredicate), C.c_int32, _PP]),
    "msv_sort_by_key": (_I, [_P, _S, _S, _I, _PP, _PP]),
    "msv_sort": (_I, [_P, _S, _I, _PP]),
    "msv_sort_user": (_I, [_P, _S, _P, _PP]),
    "msv_sort_by_key_user": (_I, [_P, _S, _S, _P, _PP, _PP]),
    "msv_merge_user": (_I, [_P, _S, _S, _P, _PP]),
    "msv_merge_by_key_user": (_I, [_P, _S, _S, _S, _S, _P, _PP, _PP]),
    "msv_set_op_user": (_I, [_P, _I, _S, _S, _P, _PP]),
    "msv_seg_sort_user": (_I, [_P, _S, _S, _P, _P, _PP, _PP]),
    "msv_inclusive_scan_by_key_user": (_I, [_P, _S, _S, _P, _I, _PP]),
    "msv_exclusive_scan_by_key_user": (_I, [_P, _S, _S, _P, _P, _I, _PP]),
    "msv_reduce_by_key_user": (_I, [_P, _S, _S, _P, _P, _I, _PP, _PP]),
    "msv_unique_by_key_user": (_I, [_P, _S, _S, _P, _PP]),
    "msv_seg_unique_user": (_I, [_P, _S, _P, _P, _PP, _PP]),
    "msv_merge": (_I, [_P, _S, _S, _I, _PP]),
    "msv_merge_by_key": (_I, [_P, _S, _S, _S, _S, _I, _PP, _PP]),
    "msv_set_op": (_I, [_P, _I, _S, _S, _I, _PP]),
    "msv_segmentation_from_offsets": (_I, [_P, _S, _PP]),
    "msv_segmentation_from_lengths": (_I, [_P, _S, _PP]),
    "msv_segmentation_destroy": (_I, [_P]),
    "msv_segmentation_offsets": (_I, [_P, _PP]),
    "msv_segmentation_segment_count": (_U32, [_P]),
    "msv_segmentation_value_count": (_U32, [_P]),
    "msv_seg_inclusive_scan": (_I, [_P, _S, _P, _I, _PP]),
    "msv_seg_exclusive_scan": (_I, [_P, _S, _P, _P, _I, _PP]),
    "msv_seg_reduce": (_I, [_P, _S, _P, _P, _I, _PP]),
    "msv_seg_sort": (_I, [_P, _S, _S, _P, _I, _PP, _PP]),
    "msv_seg_filter": (_I, [_P, _S, _P, _P, _PP, _PP]),
    "msv_seg_unique": (_I, [_P, _S, _P, _PP, _PP]),
    "msv_seg_take": (_I, [_P, _S, _P, C.c_uint32, _PP, _PP]),
    "msv_random_uniform_u32": (_I, [_P, _U32, _U32, _U64, _S]),
    "msv_random_uniform_u64": (_I, [_P, _U64, _U64, _U64, _S]),
    "msv_iota_u32": (_I, [_P, _U32, _S]),
}

_lock = threading.Lock()
_lib = None


def load(path: Path | str | None = None):
    """Loads libmsv.so and binds every prototype; raises if it is missing."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"libmsv.so not found at {p}; build it with `python -m massively_amd.build` "
                "(the HIP extension is required, there is no CPU fallback)"
            )
        lib = C.CDLL(str(p))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def lib():
    return _lib if _lib is not None else load()


def last_error() -> MsvError:
    e = MsvError()
    lib().msv_last_error(C.byref(e))
    return e
