"""Writes tests/golden/vectors.json: known-answer vectors of the reference.

Every entry records the reference file:line its inputs and expected outputs
come from (paths relative to the massively repository). Literal vectors are
transcribed from the reference's tests and doc examples; the proptest
regression inputs are parsed from the reference's committed
``*.proptest-regressions`` data files (inputs only: their expected outputs are
defined by the reference as the sequential oracle's result, which the tests
compute with numpy library routines, not with the oracle under test).

Run in the build container (it reads the reference checkout once):
    python tests/golden/make_fixtures.py [/path/to/reference]
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent

U32_MAX = 0xFFFFFFFF
U64_MAX = 0xFFFFFFFFFFFFFFFF
I32_MIN, I32_MAX = -(2**31), 2**31 - 1
I64_MIN, I64_MAX = -(2**63), 2**63 - 1


def literal_vectors() -> list[dict]:
    v = []
    v.append(dict(name="inclusive_scan_add_doc", src="massively/src/api/algorithm/scan.rs:76-95",
                  op="inclusive_scan", dtype="u32", reduce="add", input=[1, 2, 3, 4],
                  expected=[1, 3, 6, 10]))
    v.append(dict(name="exclusive_scan_add_doc", src="massively/src/api/algorithm/scan.rs:196-215",
                  op="exclusive_scan", dtype="u32", reduce="add", init=0, input=[1, 2, 3, 4],
                  expected=[0, 1, 3, 6]))
    v.append(dict(name="exclusive_scan_init_once", src="massively/src/core/scan.rs:1467-1483",
                  op="exclusive_scan_into", dtype="u32", reduce="add", init=10,
                  input=[1, 2, 3, 4], output_before=[99] * 6, output_range=[1, 5],
                  expected=[99, 10, 11, 13, 16, 99]))
    v.append(dict(name="inclusive_scan_70001_ones", src="massively/src/core/scan.rs:1304-1319",
                  op="inclusive_scan", dtype="u32", reduce="add", fill=1, len=70001,
                  check_indices=[0, 255, 256, 65535, 70000],
                  expected_at=[1, 256, 257, 65536, 70001], expected_last=70001))
    v.append(dict(name="exclusive_scan_70001_ones_init10", src="massively/src/core/scan.rs:1486-1501",
                  op="exclusive_scan", dtype="u32", reduce="add", init=10, fill=1, len=70001,
                  check_indices=[0, 255, 256, 65535, 70000],
                  expected_at=[10, 265, 266, 65545, 70010]))
    v.append(dict(name="reduce_add_doc", src="massively/src/api/algorithm/reduce.rs:9-30",
                  op="reduce", dtype="u32", reduce="add", init=0, input=[1, 2, 3, 4], expected=10))
    v.append(dict(name="copy_where_doc", src="massively/src/api/algorithm/selection.rs:131-145",
                  op="copy_where", dtype="u32", input=[10, 20, 30, 40], stencil=[1, 0, 1, 0],
                  expected=[10, 30]))
    v.append(dict(name="remove_where_doc", src="massively/src/api/algorithm/selection.rs:188-202",
                  op="remove_where", dtype="u32", input=[10, 20, 30, 40], stencil=[1, 0, 1, 0],
                  expected=[20, 40]))
    v.append(dict(name="gather_doc", src="massively/src/api/algorithm/indexed.rs:157-171",
                  op="gather", dtype="u32", values=[10, 20, 30], indices=[2, 0, 1],
                  expected=[30, 10, 20]))
    v.append(dict(name="scatter_doc", src="massively/src/api/algorithm/scatter.rs:225-247",
                  op="scatter", dtype="u32", values=[10, 20, 30], indices=[2, 0, 1],
                  output_before=[0, 0, 0], expected=[20, 30, 10]))
    v.append(dict(name="merge_stable", src="massively/src/core/merge.rs:1262-1309",
                  op="merge_by_key", dtype="u32", left_keys=[1, 2, 2, 5], right_keys=[2, 3, 4],
                  left_values=[10, 20, 21, 50], right_values=[200, 300, 400],
                  expected_keys=[1, 2, 2, 2, 3, 4, 5],
                  expected_values=[10, 20, 21, 200, 300, 400, 50]))
    v.append(dict(name="set_multiplicity", src="massively/src/core/set.rs:418-471",
                  op="set_ops", dtype="u32", left=[1, 2, 2, 2, 4], right=[2, 2, 3, 4, 4],
                  expected_union=[1, 2, 2, 2, 3, 4, 4], expected_intersection=[2, 2, 4],
                  expected_difference=[1, 2]))
    # Radix extremes: values are iota; expected = stable order of the keys.
    for dt, keys in (("u32", [U32_MAX, 0, 1, 1 << 31, 0, U32_MAX]),
                     ("u64", [U64_MAX, 0, 1, 1 << 63, 0, U64_MAX]),
                     ("i32", [I32_MAX, I32_MIN, -1, 0, I32_MIN, I32_MAX]),
                     ("i64", [I64_MAX, I64_MIN, -1, 0, I64_MIN, I64_MAX])):
        v.append(dict(name=f"radix_extremes_{dt}", src="massively/tests/vector/api/radix.rs:26-49",
                      op="radix_sort_by_key", dtype=dt, keys=keys, values=list(range(6)),
                      expected=[1, 4, 2, 3, 0, 5]))
    # Float total order (f32::total_cmp): -NaN < -inf < -1.5 < -0 (2), -0 (8) < +0 < 1.5 < inf < +NaN.
    v.append(dict(name="radix_total_order_f32", src="massively/tests/vector/api/radix.rs:74-115",
                  op="radix_sort_by_key", dtype="f32", key_bits=[
                      0x7FC00001, 0x7F800000, 0x80000000, 0x00000000, 0xFF800000, 0xFFC00001,
                      0x3FC00000, 0xBFC00000, 0x80000000],
                  values=list(range(9)), expected=[5, 4, 7, 2, 8, 3, 6, 1, 0]))
    v.append(dict(name="radix_total_order_f64", src="massively/tests/vector/api/radix.rs:74-115",
                  op="radix_sort_by_key", dtype="f64", key_bits=[
                      0x7FF8000000000001, 0x7FF0000000000000, 0x8000000000000000, 0,
                      0xFFF0000000000000, 0xFFF8000000000001, 0x3FF8000000000000,
                      0xBFF8000000000000, 0x8000000000000000],
                  values=list(range(9)), expected=[5, 4, 7, 2, 8, 3, 6, 1, 0]))
    v.append(dict(name="radix_doc_i32", src="massively/src/api/algorithm/by_key.rs:259-271",
                  op="radix_sort_by_key", dtype="i32", keys=[3, -1, 2, -1],
                  values=[30, 10, 20, 11], expected=[10, 11, 20, 30]))
    v.append(dict(name="radix_stable_513", src="massively/tests/vector/api/radix.rs:262-275",
                  op="radix_sort_by_key", dtype="u32",
                  keys=[(i * 37) % 11 for i in range(513)], values=list(range(513)),
                  expected="stable_argsort"))
    v.append(dict(name="radix_empty_and_mismatch", src="massively/tests/vector/api/radix.rs:244-259",
                  op="radix_errors", dtype="u32", empty=True, mismatch_keys=[1, 2],
                  mismatch_values=[10], expected_error=["LengthMismatch", 2, 1]))
    v.append(dict(name="seg_owned_results", src="massively/tests/seg/api.rs:286-323",
                  op="seg_reduce", dtype="u32", reduce="add", init=0, values=[1, 1, 2, 3, 3],
                  offsets=[0, 3, 5], expected=[4, 6]))
    v.append(dict(name="seg_exclusive_from_reduce", src="massively/tests/seg/api.rs:326-348",
                  op="seg_exclusive_scan", dtype="u32", reduce="add", init=0, values=[1, 2, 4],
                  offsets=[0, 2, 3], expected=[0, 1, 0], expected_reduce=[3, 4]))
    # by-key family: doc examples and the owned-API test.
    v.append(dict(name="inclusive_scan_by_key_doc", src="massively/src/api/algorithm/by_key.rs:366-397",
                  op="inclusive_scan_by_key", dtype="u32", reduce="add", keys=[1, 1, 2, 2],
                  values=[10, 20, 30, 40], expected=[10, 30, 30, 70]))
    v.append(dict(name="exclusive_scan_by_key_doc", src="massively/src/api/algorithm/by_key.rs:470-495",
                  op="exclusive_scan_by_key", dtype="u32", reduce="add", init=0, keys=[1, 1, 2, 2],
                  values=[10, 20, 30, 40], expected=[0, 10, 0, 30]))
    v.append(dict(name="reduce_by_key_doc", src="massively/src/api/algorithm/by_key.rs:586-617",
                  op="reduce_by_key", dtype="u32", reduce="add", init=0, keys=[1, 1, 2, 2],
                  values=[10, 20, 30, 40], expected_keys=[1, 2], expected=[30, 70]))
    v.append(dict(name="unique_by_key_doc", src="massively/src/api/algorithm/by_key.rs:730-757",
                  op="unique_by_key", dtype="u32", keys=[1, 1, 2, 2, 3], values=[10, 11, 20, 21, 30],
                  expected=[10, 20, 30]))
    v.append(dict(name="by_key_owned", src="massively/tests/vector/api/owned.rs:186-228",
                  op="by_key_owned", dtype="u32", reduce="add", init=0, keys=[1, 1, 2, 3],
                  values=[10, 11, 20, 30], expected_inclusive=[10, 21, 20, 30],
                  expected_exclusive=[0, 10, 0, 0], expected_reduce_keys=[1, 2, 3],
                  expected_reduce=[21, 20, 30], expected_unique=[10, 20, 30]))
    first = [1, -1, 1, -1, 1, -1, 1, -1, 1]
    second = [2, 2, 1, 2, 1, 1, 1, 1, 1]
    third = [0, 0, 2, 0, 1, 2, 1, 1, 1]
    order = sorted(range(9), key=lambda i: (first[i], second[i], third[i]))
    v.append(dict(name="radix_compound_keys_seven_values", src="massively/tests/vector/api/radix.rs:118-170",
                  op="radix_columns", key_dtypes=["i32", "u32", "u32"], keys=[first, second, third],
                  values=[[c * 100 + i for i in range(9)] for c in range(7)],
                  expected=[[c * 100 + i for i in order] for c in range(7)]))
    rows = [[0, 0, 0], [0, 0, 1], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [0, 1, 1]]
    order = sorted(range(7), key=lambda i: rows[i])
    v.append(dict(name="radix_three_keys_twelve_values", src="massively/tests/vector/api/radix.rs:172-242",
                  op="radix_columns", key_dtypes=["u32", "u32", "u32"],
                  keys=[[r[c] for r in rows] for c in range(3)],
                  values=[[c * 100 + i for i in range(7)] for c in range(12)],
                  expected=[[c * 100 + i for i in order] for c in range(12)]))
    v.append(dict(name="partition_even_doc", src="massively/src/api/algorithm/selection.rs:245-271",
                  op="partition", dtype="u32", input=[1, 2, 3, 4], pred="even",
                  expected=[2, 4, 1, 3], expected_boundary=2))
    v.append(dict(name="scatter_reduce_doc", src="massively/src/api/algorithm/scatter.rs:337-376",
                  op="scatter_reduce", dtype="u32", reduce="add", init=0, values=[2, 3, 5, 7],
                  indices=[1, 0, 1, 1], output_before=[10, 20, 30], expected=[13, 34, 30]))
    v.append(dict(name="scatter_reduce_owned", src="massively/tests/vector/api/owned.rs:283-295",
                  op="scatter_reduce", dtype="u32", reduce="add", init=0, values=[1, 2],
                  indices=[0, 0], output_before=[10, 20], expected=[13, 20]))
    v.append(dict(name="seg_unique_owned", src="massively/tests/seg/api.rs:285-313",
                  op="seg_unique", dtype="u32", input=[1, 1, 2, 3, 3], offsets=[0, 3, 5],
                  expected=[1, 2, 3], expected_offsets=[0, 2, 3]))
    v.append(dict(name="seg_take_empty_segments", src="massively/tests/seg/api.rs:766-796",
                  op="seg_take", dtype="u32", input=[2, 4, 6, 1, 1, 3, 5, 7, 8, 8],
                  offsets=[0, 0, 3, 6, 7, 10], take=2, expected=[2, 4, 1, 1, 5, 7, 8],
                  expected_offsets=[0, 0, 2, 4, 5, 7], expected_take0=[],
                  expected_offsets_take0=[0, 0, 0, 0, 0, 0]))
    # lazy inputs, each written out with vector::map(.., op::Identity)
    L = "massively/src/api/lazy.rs"
    v.append(dict(name="lazy_stride_doc", src=f"{L}:700-703", op="lazy", tree="stride",
                  start=2, step=3, take=4, expected=[2, 5, 8, 11]))
    v.append(dict(name="lazy_map_double_doc", src=f"{L}:771-776", op="lazy", tree="map",
                  input=[1, 2, 3], unary="x * 2u", expected=[2, 4, 6]))
    v.append(dict(name="lazy_permute_doc", src=f"{L}:790-796", op="lazy", tree="permute",
                  values=[10, 20, 30], indices=[2, 0], expected=[30, 10]))
    v.append(dict(name="lazy_reverse_doc", src=f"{L}:813-817", op="lazy", tree="reverse",
                  input=[10, 20, 30], expected=[30, 20, 10]))
    v.append(dict(name="lazy_repeat_each_doc", src=f"{L}:834-842", op="lazy", tree="repeat_each",
                  input=[10, 20], repeats=3, expected=[10, 10, 10, 20, 20, 20]))
    v.append(dict(name="lazy_tile_doc", src=f"{L}:859-863", op="lazy", tree="tile",
                  input=[10, 20], repeats=3, expected=[10, 20, 10, 20, 10, 20]))
    v.append(dict(name="lazy_adjacent_difference_doc", src=f"{L}:880-898", op="lazy",
                  tree="adjacent_difference", input=[3, 8, 10], reduction="b - a",
                  expected=[3, 5, 2]))
    v.append(dict(name="lazy_identity_doc", src=f"{L}:914-918", op="lazy", tree="identity",
                  input=[1, 2, 3], expected=[1, 2, 3]))
    v.append(dict(name="lazy_counting_slice", src="massively/tests/vector/api/runtime.rs:46-63",
                  op="lazy", tree="counting_slice", start=10, take=5, slice=[1, 4],
                  expected=[11, 12, 13]))
    v.append(dict(name="lazy_constant_gather", src="massively/tests/vector/api/runtime.rs:77-87",
                  op="lazy", tree="constant_gather", value=7, take=1, indices_value=0,
                  indices_take=3, expected_gather=[7, 7, 7], expected_map=[7]))
    v.append(dict(name="lazy_with_table_doc", src=f"{L}:724-744", op="lazy", tree="with_table",
                  contexts=[2, 0, 1], table=[10, 20, 30], unary="table.at(x)", expected=[30, 10, 20]))
    T = "massively/tests/vector/api/lazy.rs"
    v.append(dict(name="lazy_reverse_sliced", src=f"{T}:295-299", op="lazy", tree="reverse_slices",
                  input=[10, 20, 30, 40, 50], slices=[[1, 4], [1, 2]], expected=[30]))
    v.append(dict(name="lazy_repeat_each_sliced", src=f"{T}:311-325", op="lazy",
                  tree="repeat_each_slices", input=[10, 20, 30], repeats=3, slices=[[2, 8], [1, 5]],
                  expected=[20, 20, 20, 30], input_second=[1, 2, 3], expected_second=[2, 2, 2, 3]))
    v.append(dict(name="lazy_tile_sliced", src=f"{T}:327-341", op="lazy", tree="tile_slices",
                  input=[10, 20, 30], repeats=3, slices=[[2, 8], [1, 5]], expected=[10, 20, 30, 10],
                  input_second=[1, 2, 3], expected_second=[1, 2, 3, 1]))
    v.append(dict(name="zip3_map_sliced", src="massively/tests/vector/api/iter.rs:178-189", op="lazy",
                  tree="zip3_map_slices", columns=[[1, 2, 3, 4], [10, 20, 30, 40], [100, 200, 300, 400]],
                  slices=[[1, 4], [1, 2]], expected=[333]))
    v.append(dict(name="vector_map_doc", src="massively/src/api/algorithm/transform.rs:72-90",
                  op="map", dtype="u32", input=[1, 2, 3], unary="x + 1u", expected=[2, 3, 4]))
    v.append(dict(name="adjacent_difference_doc", src="massively/src/api/algorithm/scan.rs:132-155",
                  op="adjacent_difference", dtype="u32", input=[1, 2, 3, 4], expected=[1, 3, 5, 7]))
    v.append(dict(name="segmentation_from_lengths", src="massively/src/seg/segmentation.rs:105-128",
                  op="segmentation", lengths=[1, 2, 3], expected_offsets=[0, 1, 3, 6],
                  expected_segment_ids=[0, 1, 1, 2, 2, 2]))
    return v


def regression_inputs(ref: Path) -> list[dict]:
    out = []
    base = ref / "verification/oracle/tests/vector"
    for fname, kind in (("by_key.proptest-regressions", "pairs"),
                        ("arity_value.proptest-regressions", "seed")):
        path = base / fname
        if not path.exists():
            continue
        for line_no, line in enumerate(path.read_text().splitlines(), start=1):
            m = re.search(r"shrinks to (\w+) = (\[.*\])\s*$", line)
            if not m:
                continue
            text = m.group(2).replace("(", "[").replace(")", "]")
            data = json.loads(text)
            out.append(dict(name=f"{fname.split('.')[0]}_{line_no}",
                            src=f"verification/oracle/tests/vector/{fname}:{line_no}",
                            kind=kind, data=data))
    return out


def main() -> None:
    ref = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
    doc = {
        "about": "known-answer vectors transcribed from the massively reference tests/docs",
        "vectors": literal_vectors(),
        "regressions": regression_inputs(ref),
    }
    (HERE / "vectors.json").write_text(json.dumps(doc, indent=1) + "\n")
    print(f"wrote {len(doc['vectors'])} vectors, {len(doc['regressions'])} regression inputs")


if __name__ == "__main__":
    main()
