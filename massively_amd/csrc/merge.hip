// Merge-path merge, merge_by_key and ordered multiset operations for gfx950:
// host dispatch of the kernels in merge_kernels.hpp with the built-in
// orderings (Less / Greater over order words).
//
// Reference: massively/src/api/algorithm/merge.rs:109-282, core/merge.rs
// (stable merge, ties take the left element), api/algorithm/set.rs:44-195 and
// core/set.rs:18-46 (multiset rules).
#include "merge_kernels.hpp"
#include "msv_internal.hpp"

namespace msv {

__global__ void set_len_zero_kernel(uint32_t* p) { *p = 0; }

// ---- host ---------------------------------------------------------------------
template <typename K, int MODE, bool SET>
static msv_status partition(Ctx& c, const K* A, uint32_t na, const K* B, uint32_t nb,
                            uint32_t ntiles, uint32_t tile, uint32_t** sa, uint32_t** sb) {
  *sa = (uint32_t*)c.scratch(sizeof(uint32_t) * (ntiles + 1));
  *sb = (uint32_t*)c.scratch(sizeof(uint32_t) * (ntiles + 1));
  if (!*sa || !*sb) return MSV_ERR_OUT_OF_MEMORY;
  hipLaunchKernelGGL((merge_partition_kernel<K, OrdLess<K, MODE>, SET>), dim3((ntiles + 1 + 255) / 256),
                     dim3(256), 0, c.stream(), A, na, B, nb, ntiles, tile, *sa, *sb);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename K, int MODE, int VS>
static msv_status merge_impl(Ctx& c, const void* A, uint32_t na, const void* B, uint32_t nb,
                             const void* Av, const void* Bv, void* out_k, void* out_v) {
  const uint64_t total = (uint64_t)na + nb;
  if (total == 0) return MSV_OK;
  const uint32_t ntiles = (uint32_t)((total + kMergeTile - 1) / kMergeTile);
  uint32_t *sa, *sb;
  MSV_TRY((partition<K, MODE, false>(c, (const K*)A, na, (const K*)B, nb, ntiles, kMergeTile, &sa,
                                     &sb)));
  hipLaunchKernelGGL((merge_tile_kernel<K, OrdLess<K, MODE>, VS>), dim3(ntiles), dim3(kMergeBlock), 0,
                     c.stream(), (const K*)A, na, (const K*)B, nb, Av, Bv, sa, sb, (K*)out_k,
                     out_v);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename K, int MODE>
static msv_status set_impl(Ctx& c, int kind, const void* A, uint32_t na, const void* B,
                           uint32_t nb, void* out, uint32_t* out_len) {
  const uint64_t total = (uint64_t)na + nb;
  if (total == 0) {
    hipLaunchKernelGGL(set_len_zero_kernel, dim3(1), dim3(1), 0, c.stream(), out_len);
    MSV_HIP(hipGetLastError());
    return MSV_OK;
  }
  const uint32_t ntiles = (uint32_t)((total + kSetSpan - 1) / kSetSpan);
  uint32_t *sa, *sb;
  MSV_TRY((partition<K, MODE, true>(c, (const K*)A, na, (const K*)B, nb, ntiles, kSetSpan, &sa,
                                    &sb)));
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
#define MSV_SET_LAUNCH(KIND)                                                                    \
  hipLaunchKernelGGL((set_tile_kernel<K, OrdLess<K, MODE>, KIND>), dim3(ntiles), dim3(kSetBlock), 0,      \
                     c.stream(), (const K*)A, na, (const K*)B, nb, sa, sb, (K*)out, out_len,    \
                     counter, status, gen, ntiles)
  if (kind == SET_UNION) MSV_SET_LAUNCH(SET_UNION);
  else if (kind == SET_INTERSECTION) MSV_SET_LAUNCH(SET_INTERSECTION);
  else MSV_SET_LAUNCH(SET_DIFFERENCE);
#undef MSV_SET_LAUNCH
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename K>
static msv_status merge_by_mode(Ctx& c, int mode, size_t vs, const void* A, uint32_t na,
                                const void* B, uint32_t nb, const void* Av, const void* Bv,
                                void* ok, void* ov) {
#define MSV_MERGE_VS(MODE)                                                               \
  if (vs == 0) return merge_impl<K, MODE, 0>(c, A, na, B, nb, Av, Bv, ok, ov);           \
  if (vs == 4) return merge_impl<K, MODE, 4>(c, A, na, B, nb, Av, Bv, ok, ov);           \
  return merge_impl<K, MODE, 8>(c, A, na, B, nb, Av, Bv, ok, ov);
  if (mode == ORD_GREATER) {
    MSV_MERGE_VS(ORD_GREATER)
  }
  MSV_MERGE_VS(ORD_LESS)
#undef MSV_MERGE_VS
}

template <typename K>
static msv_status set_by_mode(Ctx& c, int mode, int kind, const void* A, uint32_t na,
                              const void* B, uint32_t nb, void* out, uint32_t* out_len) {
  if (mode == ORD_GREATER) return set_impl<K, ORD_GREATER>(c, kind, A, na, B, nb, out, out_len);
  return set_impl<K, ORD_LESS>(c, kind, A, na, B, nb, out, out_len);
}

#define MSV_KEY_SWITCH(FN, ...)                                   \
  switch (kdt) {                                                  \
    case MSV_U32: return FN<uint32_t>(__VA_ARGS__);               \
    case MSV_I32: return FN<int32_t>(__VA_ARGS__);                \
    case MSV_F32: return FN<float>(__VA_ARGS__);                  \
    case MSV_U64: return FN<uint64_t>(__VA_ARGS__);               \
    case MSV_I64: return FN<int64_t>(__VA_ARGS__);                \
    case MSV_F64: return FN<double>(__VA_ARGS__);                 \
  }

msv_status merge_run(Ctx& c, msv_dtype kdt, int mode, size_t vs, const void* A, uint32_t na,
                     const void* B, uint32_t nb, const void* Av, const void* Bv, void* ok,
                     void* ov) {
  MSV_KEY_SWITCH(merge_by_mode, c, mode, vs, A, na, B, nb, Av, Bv, ok, ov)
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad key dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status set_run(Ctx& c, msv_dtype kdt, int mode, int kind, const void* A, uint32_t na,
                   const void* B, uint32_t nb, void* out, uint32_t* out_len) {
  MSV_KEY_SWITCH(set_by_mode, c, mode, kind, A, na, B, nb, out, out_len)
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad key dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

}  // namespace msv

using namespace msv;

static msv_status cmp_mode(msv_cmp cmp, int* mode) {
  if (cmp == MSV_CMP_LESS) *mode = ORD_LESS;
  else if (cmp == MSV_CMP_GREATER) *mode = ORD_GREATER;
  else {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown comparator %d", (int)cmp);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return MSV_OK;
}

static msv_status check_total(uint32_t a, uint32_t b) {
  if ((uint64_t)a + b > 0xffffffffull) {
    set_error(MSV_ERR_LENGTH_TOO_LARGE, (uint64_t)a + b, 0, "length does not fit in u32: %llu",
              (unsigned long long)((uint64_t)a + b));
    return MSV_ERR_LENGTH_TOO_LARGE;
  }
  return MSV_OK;
}

extern "C" {

msv_status msv_merge(msv_executor* ex, msv_slice left, msv_slice right, msv_cmp cmp,
                     msv_vec** out) {
  clear_error();
  int mode;
  MSV_TRY(cmp_mode(cmp, &mode));
  View a, b;
  MSV_TRY(resolve(ex, left, &a));
  MSV_TRY(resolve(ex, right, &b));
  if (a.dtype != b.dtype) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "left and right dtypes differ");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(check_total(a.len, b.len));
  MSV_TRY(new_vec(ex, a.dtype, a.len + b.len, false, out));
  Ctx c(ex);
  msv_status s = merge_run(c, a.dtype, mode, 0, a.ptr, a.len, b.ptr, b.len, nullptr, nullptr,
                           (*out)->ptr, nullptr);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

msv_status msv_merge_by_key(msv_executor* ex, msv_slice left_keys, msv_slice left_values,
                            msv_slice right_keys, msv_slice right_values, msv_cmp cmp,
                            msv_vec** out_keys, msv_vec** out_values) {
  clear_error();
  int mode;
  MSV_TRY(cmp_mode(cmp, &mode));
  View lk, lv, rk, rv;
  MSV_TRY(resolve(ex, left_keys, &lk));
  MSV_TRY(resolve(ex, left_values, &lv));
  MSV_TRY(resolve(ex, right_keys, &rk));
  MSV_TRY(resolve(ex, right_values, &rv));
  if (lk.len != lv.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, lk.len, lv.len, "input lengths differ: %u != %u", lk.len, lv.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  if (rk.len != rv.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, rk.len, rv.len, "input lengths differ: %u != %u", rk.len, rv.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  if (lk.dtype != rk.dtype || lv.dtype != rv.dtype || !out_keys || !out_values) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "dtype mismatch or null output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(check_total(lk.len, rk.len));
  MSV_TRY(new_vec(ex, lk.dtype, lk.len + rk.len, false, out_keys));
  msv_status s = new_vec(ex, lv.dtype, lv.len + rv.len, false, out_values);
  if (s == MSV_OK) {
    Ctx c(ex);
    s = merge_run(c, lk.dtype, mode, dtype_size(lv.dtype), lk.ptr, lk.len, rk.ptr, rk.len, lv.ptr,
                  rv.ptr, (*out_keys)->ptr, (*out_values)->ptr);
  }
  if (s != MSV_OK) {
    msv_vec_release(*out_keys);
    *out_keys = nullptr;
    if (*out_values) msv_vec_release(*out_values);
    *out_values = nullptr;
  }
  return s;
}

msv_status msv_set_op(msv_executor* ex, msv_set_kind kind, msv_slice left, msv_slice right,
                      msv_cmp cmp, msv_vec** out) {
  clear_error();
  int mode;
  MSV_TRY(cmp_mode(cmp, &mode));
  if (kind < MSV_SET_UNION || kind > MSV_SET_DIFFERENCE) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown set operation %d", (int)kind);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  View a, b;
  MSV_TRY(resolve(ex, left, &a));
  MSV_TRY(resolve(ex, right, &b));
  if (a.dtype != b.dtype) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "left and right dtypes differ");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(check_total(a.len, b.len));
  const uint32_t cap = kind == MSV_SET_UNION ? a.len + b.len : a.len;
  MSV_TRY(new_vec(ex, a.dtype, cap, true, out));
  Ctx c(ex);
  msv_status s = set_run(c, a.dtype, mode, (int)kind, a.ptr, a.len, b.ptr, b.len, (*out)->ptr,
                         (*out)->d_len);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

}  // extern "C"
