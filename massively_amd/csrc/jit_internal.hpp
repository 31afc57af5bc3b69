// Shared by the run-time compiled operators (jit.hip: reductions, predicates,
// lazy inputs; jit_order.hip: comparators): the operator object and the
// hiprtc compile / load / launch helpers.
#pragma once

#include <hip/hiprtc.h>

#include <string>
#include <vector>

#include "msv_internal.hpp"

struct msv_user_op {
  int32_t kind;
  msv_dtype dtype;
  uint64_t owner_id;
  int device;
  bool lazy = false;         // compiled with a read source
  bool builtin = false;      // built-in reduction code / built-in predicates
  hipModule_t module = nullptr;
  // reduction over a pointer input: reduce (512 x 16), tile partials (256 x 4),
  // exclusive / inclusive prefixed scans
  hipFunction_t f_reduce = nullptr, f_partials = nullptr, f_scan_ex = nullptr, f_scan_in = nullptr;
  // the same kernels reading the lazy source
  hipFunction_t f_reduce_lz = nullptr, f_partials_lz = nullptr, f_scan_ex_lz = nullptr,
                f_scan_in_lz = nullptr;
  // built-in integer operator over a lazy input: the one-pass look-back scan
  // [inclusive, exclusive] x [short tiles, long tiles]
  hipFunction_t f_lb[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  uint64_t identity = 0;  // bits of the identity, for host-side padding values
  // predicate: compaction kernels for short and long inputs
  hipFunction_t f_sel_small = nullptr, f_sel_large = nullptr;
  // map: writes the lazy input out
  hipFunction_t f_map = nullptr;
  // comparator (MSV_USER_LESS): the comparison sort's block sort, pass
  // partition and merge pass, merge partition and tile, set partition and
  // set tiles; [i] of a kernel that moves values is for 0 / 4 / 8 value bytes,
  // of a set tile for union / intersection / difference.
  hipFunction_t f_block_sort[3] = {nullptr, nullptr, nullptr};
  hipFunction_t f_pass_part = nullptr;
  hipFunction_t f_pass[3] = {nullptr, nullptr, nullptr};
  hipFunction_t f_merge_part = nullptr, f_set_part = nullptr;
  hipFunction_t f_merge[3] = {nullptr, nullptr, nullptr};
  hipFunction_t f_set[3] = {nullptr, nullptr, nullptr};
  // equality (MSV_USER_EQUAL): flags of the run heads of a key column.
  hipFunction_t f_heads = nullptr;
};

namespace msv {

const char* jit_c_type(msv_dtype dt);
std::string jit_fmt(const char* f, ...);
// Compiles `src` for gfx950 and gives the code object and the symbol of each
// name expression. Needs no device.
msv_status jit_compile(const std::string& src, const std::vector<std::string>& exprs,
                       std::vector<char>& code, std::vector<std::string>& lowered);
// Compiles `src`, loads the code object on op's device and looks up each
// name expression.
msv_status jit_build(msv_user_op* op, const std::string& src,
                     const std::vector<std::string>& exprs, std::vector<hipFunction_t>& fns);
msv_status jit_launch(hipFunction_t f, uint32_t grid, uint32_t block, hipStream_t s, void** args);
// Checks that `op` is an operator of `kind` for dtype `dt` owned by `ex`.
msv_status jit_check_op(msv_executor* ex, const msv_user_op* op, int32_t kind, msv_dtype dt,
                        bool lazy);

// jit_order.hip: source and kernels of a comparator operator.
msv_status cmp_gen_source(int32_t kind, msv_dtype dtype, const char* expr, std::string& src,
                          std::vector<std::string>& exprs);
msv_status cmp_op_create(msv_executor* ex, int32_t kind, msv_dtype dtype, const char* expr,
                         msv_user_op** out);

}  // namespace msv
