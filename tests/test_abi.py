"""CPU checks of the C-ABI boundary: libmsv.so loads, exports every function
include/msv.h declares, the Python mirror binds each of them, and calls made
without a GPU fail with a status code (no crash). No kernel runs here."""
import ctypes as C
import re
from pathlib import Path

import pytest

import massively_amd as m
from massively_amd import _lib as L
from massively_amd.error import InvalidArgument, from_status

ROOT = Path(__file__).resolve().parent.parent
HEADER = (ROOT / "include" / "msv.h").read_text()


def declared_functions():
    body = re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)
    return sorted(set(re.findall(r"\b(msv_[a-z0-9_]+)\s*\(", body)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("msv_executor_create", "msv_inclusive_scan", "msv_exclusive_scan",
                     "msv_reduce", "msv_radix_sort_by_key", "msv_sort_by_key", "msv_sort",
                     "msv_merge", "msv_merge_by_key", "msv_set_op", "msv_copy_where",
                     "msv_remove_where", "msv_gather", "msv_scatter", "msv_seg_inclusive_scan",
                     "msv_seg_exclusive_scan", "msv_seg_reduce", "msv_seg_sort",
                     "msv_segmentation_from_offsets", "msv_segmentation_from_lengths"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    assert set(declared_functions()) <= set(L.PROTOTYPES), set(declared_functions()) - set(L.PROTOTYPES)


def test_abi_version():
    assert L.load().msv_abi_version() == 1


def test_no_gpu_fails_cleanly():
    try:
        ex = m.Executor(0)
    except m.Error as e:
        assert "device" in str(e).lower() or isinstance(e, m.error.Launch) or isinstance(e, m.error.InvalidArgument)
        return
    ex.close()
    pytest.skip("a GPU is present")


def test_null_handles_are_rejected():
    lib = L.load()
    h = C.c_void_p()
    assert lib.msv_executor_create(0, None) == L.ERR_INVALID_ARGUMENT
    s = L.MsvSlice(None, 0, 0)
    assert lib.msv_inclusive_scan(None, s, 0, C.byref(h)) == L.ERR_INVALID_ARGUMENT
    assert lib.msv_executor_sync(None) == L.ERR_INVALID_ARGUMENT
    e = L.last_error()
    assert e.code == L.ERR_INVALID_ARGUMENT and b"executor" in e.message


def test_error_mapping_mirrors_massively_error():
    # massively/src/core/error.rs:5-13 variants and their Display texts.
    assert m.LengthMismatch(2, 1) == m.LengthMismatch(2, 1)
    assert m.LengthMismatch(2, 1) != m.LengthMismatch(1, 2)
    assert str(m.LengthMismatch(2, 1)) == "input lengths differ: 2 != 1"
    assert str(m.OutputTooShort(5, 3)) == "output length 3 is shorter than input length 5"
    assert str(m.InvalidSegmentation()) == "segmentation is not a valid ordered partition"
    assert str(m.ForeignExecutor()) == "executor does not own this device data"
    assert isinstance(from_status(L.ERR_FOREIGN_EXECUTOR), m.ForeignExecutor)
    assert isinstance(from_status(L.ERR_INVALID_SEGMENTATION), m.InvalidSegmentation)


def test_predicate_threshold_bits():
    import numpy as np

    p = m.op.Lt(np.float32(0.25)).c_struct(np.dtype(np.float32))
    assert p.kind == 2 and p.a == 0x3E800000
    p = m.op.Ge(-3).c_struct(np.dtype(np.int64))
    assert p.a == (2**64 - 3)
    p = m.op.BitsEq(0x30, 0x10).c_struct(np.dtype(np.uint32))
    assert (p.kind, p.a, p.b) == (7, 0x30, 0x10)


def test_host_mirror_surface():
    """The Python mirror keeps the reference API surface (no GPU needed)."""
    import massively_amd as m

    for a in ("alloc", "to_device", "to_host", "wrap", "sync", "trim", "set_option", "get_option"):
        assert hasattr(m.Executor, a), a
    cfgs = m.segsort_configs()
    assert len(cfgs) >= 2 and all(b % 64 == 0 and i > 0 and d in (4, 8) for b, i, d in cfgs)


def test_tuning_file_round_trip(tmp_path):
    """The autotuner's table saves to and merges from a text file (entries of
    any device signature), rejects a malformed line, and clears."""
    m.tuning_clear()
    f = tmp_path / "tune.txt"
    f.write_text("# comment\ngfx950:sramecc+:xnack-/256 segsort 1048849 3\n"
                 "gfx942:sramecc+:xnack-/304 segsort 1048849 1\n")
    m.tuning_load(f)
    assert m.tuning_count() == 2
    g = tmp_path / "tune2.txt"
    m.tuning_save(g)
    lines = sorted(x for x in g.read_text().splitlines() if not x.startswith("#"))
    assert lines == ["gfx942:sramecc+:xnack-/304 segsort 1048849 1",
                     "gfx950:sramecc+:xnack-/256 segsort 1048849 3"]
    m.tuning_clear()
    assert m.tuning_count() == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("not an entry\n")
    with pytest.raises(InvalidArgument):
        m.tuning_load(bad)
    with pytest.raises(InvalidArgument):
        m.tuning_load(tmp_path / "missing.txt")
    m.tuning_clear()
