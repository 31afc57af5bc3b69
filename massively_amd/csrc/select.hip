// vector::copy_where / remove_where (stencil and built-in predicate forms) for
// gfx950, in one pass.
//
// Reference: massively/src/api/algorithm/selection.rs:131-260 and
// core/selection.rs:151-411 (scan of flags, count, selected-index scatter,
// gather: four launches and ~30 B/elem). Here one kernel reads each value and
// its flag once: a 512-thread block owns 4096 or 8192 elements laid out in
// 4-element groups, row-striped (group (r, t) starts at (r*512 + t)*4), counts the
// selected elements per (row, thread), turns the counts into in-tile ranks
// with DPP wave scans, resolves the tile's output offset by decoupled
// lookback on counts, stages the selected elements in LDS in input order and
// writes them out contiguously. The block that owns the last tile writes the
// total count into the output's device-resident length (no host sync).
#include "device_common.hpp"
#include "msv_internal.hpp"
#include "predicate.hpp"
#include "select_kernels.hpp"

namespace msv {

// Number of elements passing the predicate, added to *total (zeroed before).
template <typename S>
__global__ __launch_bounds__(256) void pred_count_kernel(const S* __restrict__ in, uint32_t n,
                                                         Pred<S> pred, bool aligned,
                                                         uint32_t* __restrict__ total) {
  const size_t groups = ((size_t)n + 3) / 4;
  uint32_t c = 0;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups;
       g += (size_t)gridDim.x * blockDim.x) {
    S v[4];
    load4<S, true>(in, g * 4, n, aligned && g * 4 + 4 <= n, v);
#pragma unroll
    for (int k = 0; k < 4; ++k) c += (g * 4 + k < n && pred(v[k])) ? 1u : 0u;
  }
  c = wave_reduce<uint32_t, OP_ADD>(c);
  if ((threadIdx.x & 63u) == 0 && c) atomicAdd(total, c);
}

__global__ void store_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

template <typename T, typename S, int FORM>
static msv_status select_launch(Ctx& c, const void* in, const void* st, const msv_predicate* p,
                                bool invert, uint32_t n, void* out, uint32_t* out_len) {
  if (n == 0) {
    hipLaunchKernelGGL(store_u32_kernel, dim3(1), dim3(1), 0, c.stream(), out_len, 0u);
    MSV_HIP(hipGetLastError());
    return MSV_OK;
  }
  const Pred<S> pred = make_pred<S>(p);
  const bool large = n >= kSelLargeMin;
  constexpr int BL = sel_block_large<T, FORM>(), RL = sel_rows_large_of<T, FORM>();
  const uint32_t tile = large ? (uint32_t)BL * RL * 4 : (uint32_t)kSelBlock * kSelRowsSmall * 4;
  const uint32_t ntiles = div_up(n, tile);
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
  const bool aligned = (((uintptr_t)in | (uintptr_t)st) & 15u) == 0;
  if (large)
    hipLaunchKernelGGL((select_kernel<T, S, FORM, RL, false, Pred<S>, const T* __restrict__, BL>),
                       dim3(ntiles), dim3(BL), 0, c.stream(), (const T*)in, (const S*)st, pred, invert,
                       n, aligned, (T*)out, out_len, status, gen, ntiles, (const uint32_t*)nullptr);
  else
    hipLaunchKernelGGL((select_kernel<T, S, FORM, kSelRowsSmall>), dim3(ntiles), dim3(kSelBlock), 0,
                       c.stream(), (const T*)in, (const S*)st, pred, invert, n, aligned, (T*)out,
                       out_len, status, gen, ntiles, (const uint32_t*)nullptr);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename S>
static msv_status select_with_stencil_type(Ctx& c, size_t vsize, bool same, const void* in,
                                           const void* st, const msv_predicate* p, bool invert,
                                           uint32_t n, void* out, uint32_t* out_len) {
  if (same) {
    if constexpr (sizeof(S) == 4)
      return select_launch<uint32_t, S, 2>(c, in, st, p, invert, n, out, out_len);
    else
      return select_launch<uint64_t, S, 2>(c, in, st, p, invert, n, out, out_len);
  }
  if (vsize == 4) return select_launch<uint32_t, S, 1>(c, in, st, p, invert, n, out, out_len);
  return select_launch<uint64_t, S, 1>(c, in, st, p, invert, n, out, out_len);
}

msv_status select_pred_run(Ctx& c, msv_dtype dt, const void* in, const msv_predicate* p,
                           bool invert, uint32_t n, void* out, uint32_t* out_len) {
  const size_t vs = dtype_size(dt);
  switch (dt) {
    case MSV_U32: return select_with_stencil_type<uint32_t>(c, vs, true, in, in, p, invert, n, out, out_len);
    case MSV_I32: return select_with_stencil_type<int32_t>(c, vs, true, in, in, p, invert, n, out, out_len);
    case MSV_F32: return select_with_stencil_type<float>(c, vs, true, in, in, p, invert, n, out, out_len);
    case MSV_U64: return select_with_stencil_type<uint64_t>(c, vs, true, in, in, p, invert, n, out, out_len);
    case MSV_I64: return select_with_stencil_type<int64_t>(c, vs, true, in, in, p, invert, n, out, out_len);
    case MSV_F64: return select_with_stencil_type<double>(c, vs, true, in, in, p, invert, n, out, out_len);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

// vector::partition: the passing count first, then one pass that writes the
// passing elements from 0 and the failing ones from that count on.
template <typename S>
static msv_status partition_launch(Ctx& c, const void* in, uint32_t n, const msv_predicate* p,
                                   void* out, uint32_t* d_total) {
  using U = typename Bits<S>::U;
  hipLaunchKernelGGL(store_u32_kernel, dim3(1), dim3(1), 0, c.stream(), d_total, 0u);
  MSV_HIP(hipGetLastError());
  if (n == 0) return MSV_OK;
  const Pred<S> pred = make_pred<S>(p);
  const bool aligned = ((uintptr_t)in & 15u) == 0;
  const uint32_t groups = (n + 3) / 4;
  const uint32_t blocks = min((groups + 255) / 256, 4096u);
  hipLaunchKernelGGL(pred_count_kernel<S>, dim3(blocks), dim3(256), 0, c.stream(), (const S*)in, n,
                     pred, aligned, d_total);
  MSV_HIP(hipGetLastError());
  const bool large = n >= kSelLargeMin;
  const uint32_t tile = (uint32_t)kSelBlock * (large ? sel_rows_large<U>() : kSelRowsSmall) * 4;
  const uint32_t ntiles = div_up(n, tile);
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
  if (large)
    hipLaunchKernelGGL((select_kernel<U, S, 2, sel_rows_large<U>(), true>), dim3(ntiles), dim3(kSelBlock),
                       0, c.stream(), (const U*)in, (const S*)in, pred, false, n, aligned, (U*)out,
                       (uint32_t*)nullptr, status, gen, ntiles, (const uint32_t*)d_total);
  else
    hipLaunchKernelGGL((select_kernel<U, S, 2, kSelRowsSmall, true>), dim3(ntiles), dim3(kSelBlock),
                       0, c.stream(), (const U*)in, (const S*)in, pred, false, n, aligned, (U*)out,
                       (uint32_t*)nullptr, status, gen, ntiles, (const uint32_t*)d_total);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename S>
static msv_status unique_with_key_type(Ctx& c, size_t vsize, const void* keys, const void* vals,
                                       uint32_t n, void* out, uint32_t* out_len) {
  if (vsize == 4) return select_launch<uint32_t, S, 3>(c, vals, keys, nullptr, false, n, out, out_len);
  return select_launch<uint64_t, S, 3>(c, vals, keys, nullptr, false, n, out, out_len);
}

msv_status unique_runs_run(Ctx& c, msv_dtype kdt, const void* keys, size_t vsize,
                           const void* vals, uint32_t n, void* out, uint32_t* out_len) {
  switch (kdt) {
    case MSV_U32: return unique_with_key_type<uint32_t>(c, vsize, keys, vals, n, out, out_len);
    case MSV_I32: return unique_with_key_type<int32_t>(c, vsize, keys, vals, n, out, out_len);
    case MSV_F32: return unique_with_key_type<float>(c, vsize, keys, vals, n, out, out_len);
    case MSV_U64: return unique_with_key_type<uint64_t>(c, vsize, keys, vals, n, out, out_len);
    case MSV_I64: return unique_with_key_type<int64_t>(c, vsize, keys, vals, n, out, out_len);
    case MSV_F64: return unique_with_key_type<double>(c, vsize, keys, vals, n, out, out_len);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

}  // namespace msv

using namespace msv;

static msv_status select_api(msv_executor* ex, msv_slice input, msv_slice stencil,
                             const msv_predicate* pred, bool invert, msv_vec** out) {
  clear_error();
  View in, st;
  MSV_TRY(resolve(ex, input, &in));
  MSV_TRY(resolve(ex, stencil, &st));
  if (in.len != st.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, in.len, st.len, "input lengths differ: %u != %u", in.len,
              st.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  if (!pred && st.dtype != MSV_U32) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "stencil must be u32 flags (MFlag)");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (pred && pred->kind > PRED_BITS_EQ) {
    set_error(MSV_ERR_INVALID_ARGUMENT, pred->kind, 0, "unknown predicate kind %u", pred->kind);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, in.dtype, in.len, true, out));
  Ctx c(ex);
  const size_t vs = dtype_size(in.dtype);
  msv_status s;
  if (!pred) {
    if (vs == 4)
      s = select_launch<uint32_t, uint32_t, 0>(c, in.ptr, st.ptr, nullptr, invert, in.len,
                                              (*out)->ptr, (*out)->d_len);
    else
      s = select_launch<uint64_t, uint32_t, 0>(c, in.ptr, st.ptr, nullptr, invert, in.len,
                                              (*out)->ptr, (*out)->d_len);
  } else {
    const bool same = in.ptr == st.ptr && in.dtype == st.dtype;
    switch (st.dtype) {
      case MSV_U32: s = select_with_stencil_type<uint32_t>(c, vs, same, in.ptr, st.ptr, pred, invert, in.len, (*out)->ptr, (*out)->d_len); break;
      case MSV_I32: s = select_with_stencil_type<int32_t>(c, vs, same, in.ptr, st.ptr, pred, invert, in.len, (*out)->ptr, (*out)->d_len); break;
      case MSV_F32: s = select_with_stencil_type<float>(c, vs, same, in.ptr, st.ptr, pred, invert, in.len, (*out)->ptr, (*out)->d_len); break;
      case MSV_U64: s = select_with_stencil_type<uint64_t>(c, vs, same, in.ptr, st.ptr, pred, invert, in.len, (*out)->ptr, (*out)->d_len); break;
      case MSV_I64: s = select_with_stencil_type<int64_t>(c, vs, same, in.ptr, st.ptr, pred, invert, in.len, (*out)->ptr, (*out)->d_len); break;
      default: s = select_with_stencil_type<double>(c, vs, same, in.ptr, st.ptr, pred, invert, in.len, (*out)->ptr, (*out)->d_len); break;
    }
  }
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

extern "C" {

msv_status msv_copy_where(msv_executor* ex, msv_slice input, msv_slice stencil, msv_vec** out) {
  return select_api(ex, input, stencil, nullptr, false, out);
}

msv_status msv_remove_where(msv_executor* ex, msv_slice input, msv_slice stencil, msv_vec** out) {
  return select_api(ex, input, stencil, nullptr, true, out);
}

msv_status msv_copy_where_pred(msv_executor* ex, msv_slice input, msv_slice stencil,
                               const msv_predicate* pred, int32_t invert, msv_vec** out) {
  if (!pred) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null predicate");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return select_api(ex, input, stencil, pred, invert != 0, out);
}

}  // extern "C"

extern "C" msv_status msv_partition(msv_executor* ex, msv_slice input, const msv_predicate* pred,
                                    msv_vec** out, uint32_t* boundary) {
  clear_error();
  MSV_TRY(no_capture(ex, "vector::partition"));
  View in;
  MSV_TRY(resolve(ex, input, &in));
  if (!pred || !out || !boundary) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null predicate or output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (pred->kind > PRED_BITS_EQ) {
    set_error(MSV_ERR_INVALID_ARGUMENT, pred->kind, 0, "unknown predicate kind %u", pred->kind);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, in.dtype, in.len, false, out));
  Ctx c(ex);
  uint32_t* d_total = (uint32_t*)c.scratch(sizeof(uint32_t));
  msv_status s = d_total ? MSV_OK : MSV_ERR_OUT_OF_MEMORY;
  if (s == MSV_OK) {
    switch (in.dtype) {
      case MSV_U32: s = partition_launch<uint32_t>(c, in.ptr, in.len, pred, (*out)->ptr, d_total); break;
      case MSV_I32: s = partition_launch<int32_t>(c, in.ptr, in.len, pred, (*out)->ptr, d_total); break;
      case MSV_F32: s = partition_launch<float>(c, in.ptr, in.len, pred, (*out)->ptr, d_total); break;
      case MSV_U64: s = partition_launch<uint64_t>(c, in.ptr, in.len, pred, (*out)->ptr, d_total); break;
      case MSV_I64: s = partition_launch<int64_t>(c, in.ptr, in.len, pred, (*out)->ptr, d_total); break;
      default: s = partition_launch<double>(c, in.ptr, in.len, pred, (*out)->ptr, d_total); break;
    }
  }
  if (s == MSV_OK) {
    hipError_t e = hipMemcpyAsync(boundary, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost, ex->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ex->stream);
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "partition failed: %s", hipGetErrorString(e));
      s = MSV_ERR_LAUNCH;
    }
  }
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}
