// The by-key family for gfx950: vector::{inclusive_scan_by_key,
// exclusive_scan_by_key, reduce_by_key, unique_by_key} with the built-in
// Equal (numeric == on the key dtype: -0.0 == +0.0, a NaN key equals nothing).
//
// Reference: massively/src/api/algorithm/by_key.rs:359-811 and
// core/by_key.rs:30-213; the oracle's loops are verification/oracle/src/
// vector.rs:92-175 (scans restart where adjacent keys differ; reduce_by_key
// emits the key of each run's first element and init (+) the run's values)
// and :578-594 (unique_by_key keeps each run's first value).
//
// A run head is an element whose key differs from its predecessor's. Every
// kernel here finds the heads inside its own tile from the keys (each thread
// compares its keys with the previous lane's last key), so the keys are read
// once and no head array is written:
// * the scans are the segmented tile scan with those heads;
// * reduce_by_key is the same scan, whose tiles also count their heads (a
//   second decoupled look-back on the counts gives each run its output
//   slot): every run head writes its key and every run end (an element
//   followed by a head, or the last element) writes init (+) its segmented
//   inclusive value, in one pass;
// * unique_by_key is the one-pass compaction (copy_where) with "is a run
//   head" as the flag.
#include "device_common.hpp"
#include "msv_internal.hpp"

namespace msv {

static __global__ void zero_len_kernel(uint32_t* a, uint32_t* b) {
  if (a) *a = 0;
  if (b) *b = 0;
}

}  // namespace msv

using namespace msv;

static bool op_ok_for(msv_dtype dt, int op) {
  if (op < MSV_OP_ADD || op > MSV_OP_XOR) return false;
  return !(op == MSV_OP_XOR && dtype_is_float(dt));
}

static msv_status bykey_inputs(msv_executor* ex, msv_slice keys, msv_slice values, View* k,
                               View* v) {
  MSV_TRY(resolve(ex, keys, k));
  MSV_TRY(resolve(ex, values, v));
  if (k->len != v->len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, k->len, v->len, "input lengths differ: %u != %u", k->len,
              v->len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  return MSV_OK;
}

static msv_status scan_by_key_api(msv_executor* ex, msv_slice keys, msv_slice values,
                                  const void* init, msv_op op, bool exclusive, msv_vec** out) {
  clear_error();
  View k, v;
  MSV_TRY(bykey_inputs(ex, keys, values, &k, &v));
  if (!op_ok_for(v.dtype, op) || !out || (exclusive && !init)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad operator, init or output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, v.dtype, v.len, false, out));
  Ctx c(ex);
  msv_status s = runs_scan_run(c, k.dtype, k.ptr, v.dtype, op,
                               exclusive ? RUNS_EXCLUSIVE : RUNS_INCLUSIVE, v.ptr, v.len,
                               (*out)->ptr, init, nullptr, nullptr, nullptr);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

extern "C" {

msv_status msv_inclusive_scan_by_key(msv_executor* ex, msv_slice keys, msv_slice values, msv_op op,
                                     msv_vec** out) {
  return scan_by_key_api(ex, keys, values, nullptr, op, false, out);
}

msv_status msv_exclusive_scan_by_key(msv_executor* ex, msv_slice keys, msv_slice values,
                                     const void* init, msv_op op, msv_vec** out) {
  return scan_by_key_api(ex, keys, values, init, op, true, out);
}

msv_status msv_reduce_by_key(msv_executor* ex, msv_slice keys, msv_slice values, const void* init,
                             msv_op op, msv_vec** out_keys, msv_vec** out_values) {
  clear_error();
  View k, v;
  MSV_TRY(bykey_inputs(ex, keys, values, &k, &v));
  if (!op_ok_for(v.dtype, op) || !init || !out_keys || !out_values) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad operator, init or output handles");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, k.dtype, k.len, true, out_keys));
  msv_status s = new_vec(ex, v.dtype, v.len, true, out_values);
  if (s != MSV_OK) {
    msv_vec_release(*out_keys);
    *out_keys = nullptr;
    return s;
  }
  Ctx c(ex);
  const uint32_t n = k.len;
  if (n == 0) {
    hipLaunchKernelGGL(zero_len_kernel, dim3(1), dim3(1), 0, c.stream(), (*out_keys)->d_len,
                       (*out_values)->d_len);
    s = hipGetLastError() == hipSuccess ? MSV_OK : MSV_ERR_LAUNCH;
  } else {
    s = runs_scan_run(c, k.dtype, k.ptr, v.dtype, op, RUNS_REDUCE, v.ptr, n, (*out_values)->ptr,
                      init, (*out_keys)->ptr, (*out_keys)->d_len, (*out_values)->d_len);
  }
  if (s != MSV_OK) {
    msv_vec_release(*out_keys);
    msv_vec_release(*out_values);
    *out_keys = *out_values = nullptr;
  }
  return s;
}

msv_status msv_unique_by_key(msv_executor* ex, msv_slice keys, msv_slice values,
                             msv_vec** out_values) {
  clear_error();
  View k, v;
  MSV_TRY(bykey_inputs(ex, keys, values, &k, &v));
  if (!out_values) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, v.dtype, v.len, true, out_values));
  Ctx c(ex);
  msv_status s;
  if (k.len == 0) {
    hipLaunchKernelGGL(zero_len_kernel, dim3(1), dim3(1), 0, c.stream(), (*out_values)->d_len,
                       (uint32_t*)nullptr);
    s = hipGetLastError() == hipSuccess ? MSV_OK : MSV_ERR_LAUNCH;
  } else {
    s = unique_runs_run(c, k.dtype, k.ptr, dtype_size(v.dtype), v.ptr, k.len, (*out_values)->ptr,
                        (*out_values)->d_len);
  }
  if (s != MSV_OK) {
    msv_vec_release(*out_values);
    *out_values = nullptr;
  }
  return s;
}

}  // extern "C"
