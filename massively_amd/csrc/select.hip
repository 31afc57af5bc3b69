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

namespace msv {

constexpr int kSelBlock = 512;
// Rows of 4-element groups per thread (512-thread blocks, which ran the
// predicate forms 7-9% faster than 256-thread blocks with the same tiles): 2
// (4096-element tiles) for short inputs, 4 (8192) from 2^22 elements on, so
// that fewer tiles take part in the look-back.
constexpr int kSelRowsSmall = 2;
constexpr int kSelRowsLarge = 4;
constexpr uint32_t kSelLargeMin = 1u << 22;

// FORM 0: u32 stencil array (flag = stencil != 0).
// FORM 1: predicate on a separate stencil column of type S.
// FORM 2: predicate on the input itself (S == T, one read).
// FORM 3: run heads of a key column of type S (unique_by_key): an element is
//         kept when it is the first or its key differs from the previous key.
// PART (partition): failing elements are written too, after all passing
// ones: `part_total` holds the number of passing elements, and a failing
// element goes to part_total + (failing elements before it).
template <typename T, typename S, int FORM, int ROWS, bool PART = false>
__global__ __launch_bounds__(kSelBlock) void select_kernel(const T* __restrict__ in,
                                                           const S* __restrict__ stencil,
                                                           Pred<S> pred, bool invert,
                                                           uint32_t n, bool aligned,
                                                           T* __restrict__ out,
                                                           uint32_t* __restrict__ out_len,
                                                           uint64_t* status, uint32_t gen,
                                                           uint32_t ntiles,
                                                           const uint32_t* __restrict__ part_total) {
  constexpr int NW = kSelBlock / kWave;
  constexpr uint32_t kSelTile = (uint32_t)kSelBlock * ROWS * 4;
  __shared__ T s_stage[kSelTile];
  __shared__ uint32_t s_wave[ROWS][NW];
  __shared__ uint32_t s_prefix;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * kSelTile;
  const bool full = aligned && base + kSelTile <= n;

  T x[ROWS][4];
  uint32_t f[ROWS][4];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const size_t i = base + ((size_t)r * kSelBlock + tid) * 4;
    load4<T>(in, i, n, full, x[r]);
    if constexpr (FORM == 0) {
      S s[4];
      load4<S>(stencil, i, n, full, s);
#pragma unroll
      for (int k = 0; k < 4; ++k) f[r][k] = ((s[k] != 0) != invert) && (i + k < n);
    } else if constexpr (FORM == 1) {
      S s[4];
      load4<S>(stencil, i, n, full, s);
#pragma unroll
      for (int k = 0; k < 4; ++k) f[r][k] = (pred(s[k]) != invert) && (i + k < n);
    } else if constexpr (FORM == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k) f[r][k] = (pred(bit_cast<S>(x[r][k])) != invert) && (i + k < n);
    } else {
      S s[4];
      load4<S>(stencil, i, n, full, s);
      // The key before the group: the previous lane's last key (lane 0
      // reads it; consecutive lanes hold consecutive groups).
      const S before = shift_up1<S>(s[3], (lane == 0 && i > 0 && i <= n) ? stencil[i - 1] : S());
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const S pk = k == 0 ? before : s[k - 1];
        f[r][k] = (i + k < n) && (i + k == 0 || !key_equal(pk, s[k]));
      }
    }
  }

  uint32_t ws[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const uint32_t c = f[r][0] + f[r][1] + f[r][2] + f[r][3];
    ws[r] = wave_inclusive_scan<uint32_t, OP_ADD>(c);
    if (lane == 63) s_wave[r][wid] = ws[r];
  }
  __syncthreads();
  // Exclusive count before (row r, wave wid) and the tile total: every wave
  // scans the ROWS x NW partials in (row, wave) order, one per lane, and takes
  // its own with readlane (wave-uniform values, no per-thread copies).
  static_assert(ROWS * NW <= kWave, "one partial per lane");
  constexpr int NP = ROWS * NW;
  uint32_t q = lane < (uint32_t)NP ? s_wave[lane / NW][lane % NW] : 0u;
  q = wave_inclusive_scan<uint32_t, OP_ADD>(q);
  const uint32_t agg = readlane(q, NP - 1);
  const uint32_t xq = shift_up1<uint32_t>(q, 0u);
  const int w0 = __builtin_amdgcn_readfirstlane((int)wid);
  uint32_t pre[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) pre[r] = readlane(xq, r * NW + w0);

  if (wid == 0) {
    TileStatus<uint32_t> st{status, gen};
    uint32_t excl = 0;
    if (tile == 0) {
      if (lane == 0) st.publish(0, kFlagInc, 0, agg);
    } else {
      if (lane == 0) st.publish(tile, kFlagAgg, 0, agg);
      // Self-help: count the selected elements of tile q from the input.
      auto count_of = [&](uint32_t q) -> uint32_t {
        const size_t b = (size_t)q * kSelTile;
        const size_t e = min((size_t)n, b + kSelTile);
        uint32_t c = 0;
        for (size_t i = b + lane; i < e; i += kWave) {
          bool f;
          if constexpr (FORM == 0) f = stencil[i] != 0;
          else if constexpr (FORM == 1) f = pred(stencil[i]);
          else if constexpr (FORM == 2) f = pred(bit_cast<S>(in[i]));
          else f = i == 0 || !key_equal(stencil[i - 1], stencil[i]);
          c += (f != invert) ? 1u : 0u;
        }
        return wave_reduce<uint32_t, OP_ADD>(c);
      };
      excl = wave_lookback_sh<uint32_t, OP_ADD>(st, tile, count_of);
      if (lane == 0) st.publish(tile, kFlagInc, 0, excl + agg);
    }
    if (lane == 0) {
      s_prefix = excl;
      if (out_len && tile == ntiles - 1) *out_len = excl + agg;
    }
  }

  // Stage selected elements at their in-tile rank (input order).
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    uint32_t pos = pre[r] + shift_up1<uint32_t>(ws[r], 0u);
    const uint32_t j0 = ((uint32_t)r * kSelBlock + tid) * 4;  // in-tile index of the group
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (f[r][k]) s_stage[pos] = x[r][k];
      // A failing element's rank among the tile's failing ones: the elements
      // before it minus the passing ones.
      if (PART && !f[r][k] && base + j0 + k < n) s_stage[agg + (j0 + k - pos)] = x[r][k];
      pos += f[r][k];
    }
  }
  __syncthreads();
  const size_t o = s_prefix;
  for (uint32_t i = tid; i < agg; i += kSelBlock) out[o + i] = s_stage[i];
  if constexpr (PART) {
    const uint32_t cnt = (uint32_t)min((size_t)kSelTile, (size_t)n - base);
    const size_t fo = (size_t)*part_total + (base - o);  // failing elements before the tile
    for (uint32_t i = agg + tid; i < cnt; i += kSelBlock)
      if (fo + (i - agg) < n) out[fo + (i - agg)] = s_stage[i];  // always true when counts agree
  }
}

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
    load4<S>(in, g * 4, n, aligned && g * 4 + 4 <= n, v);
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
  const uint32_t tile = (uint32_t)kSelBlock * (large ? kSelRowsLarge : kSelRowsSmall) * 4;
  const uint32_t ntiles = (n + tile - 1) / tile;
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
  const bool aligned = (((uintptr_t)in | (uintptr_t)st) & 15u) == 0;
  if (large)
    hipLaunchKernelGGL((select_kernel<T, S, FORM, kSelRowsLarge>), dim3(ntiles), dim3(kSelBlock), 0,
                       c.stream(), (const T*)in, (const S*)st, pred, invert, n, aligned, (T*)out,
                       out_len, status, gen, ntiles, (const uint32_t*)nullptr);
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
  const uint32_t tile = (uint32_t)kSelBlock * (large ? kSelRowsLarge : kSelRowsSmall) * 4;
  const uint32_t ntiles = (n + tile - 1) / tile;
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
  if (large)
    hipLaunchKernelGGL((select_kernel<U, S, 2, kSelRowsLarge, true>), dim3(ntiles), dim3(kSelBlock),
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
