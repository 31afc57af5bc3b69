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
// A run head is an element whose key differs from its predecessor's. One
// kernel turns the keys into a head bitmask (a wave ballot per 64 keys, so
// 1 bit per element). The scans are then the segmented scan with heads from
// the mask. reduce_by_key and unique_by_key are one compaction pass over the
// mask: a tile counts its heads and its run ends (an end is an element whose
// successor is a head, or the last element), a decoupled look-back (blockIdx
// tiles, self-help from the mask) turns both counts into output positions,
// and each thread copies its heads' keys / values and its ends' values there.
// reduce_by_key's values are the segmented inclusive scan at the run ends,
// combined with init afterwards.
#include "device_common.hpp"
#include "msv_internal.hpp"

namespace msv {

constexpr int kByKeyBlock = 256;
constexpr int kByKeyIpt = 16;  // elements per thread in the compaction (half a mask word)
constexpr uint32_t kByKeyTile = kByKeyBlock * kByKeyIpt;

template <typename K>
__device__ inline bool key_equal(K a, K b) {
  return a == b;  // numeric for float keys, bitwise for integer keys
}

// mask[w] bit b = element 32 w + b starts a run.
template <typename K>
__global__ __launch_bounds__(256) void bykey_heads_kernel(const K* __restrict__ keys, uint32_t n,
                                                          uint32_t* __restrict__ mask) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool head = false;
  if (i < n) head = i == 0 || !key_equal(keys[i - 1], keys[i]);
  const uint64_t b = ballot(head);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nw = (n + 31) / 32;
  const uint32_t w = (i - lane) >> 5;  // first word of this wave's 64 elements
  if (lane == 0 && w < nw) mask[w] = (uint32_t)b;
  if (lane == 32 && w + 1 < nw) mask[w + 1] = (uint32_t)(b >> 32);
}

template <int S>
struct Word;
template <>
struct Word<4> { using T = uint32_t; };
template <>
struct Word<8> { using T = uint64_t; };

// Head and end bits of elements [e0, e0 + 16) (e0 a multiple of 16).
__device__ inline void head_end_bits(const uint32_t* __restrict__ mask, uint32_t n, uint32_t e0,
                                     uint32_t& hb, uint32_t& eb) {
  hb = eb = 0;
  if (e0 >= n) return;
  const uint32_t w = e0 >> 5, sh = e0 & 31u;
  const uint32_t nw = (n + 31) / 32;
  const uint64_t two = (uint64_t)mask[w] | ((w + 1 < nw) ? (uint64_t)mask[w + 1] << 32 : 0ull);
  const uint32_t cnt = min(16u, n - e0);
  const uint32_t valid = cnt == 32 ? ~0u : ((1u << cnt) - 1u);
  hb = (uint32_t)(two >> sh) & valid & 0xffffu;
  // The end of element j is the head of element j + 1; the last element
  // always ends a run.
  uint32_t next = (uint32_t)(two >> (sh + 1)) & 0xffffu;
  if (e0 + cnt == n) next |= 1u << (cnt - 1);
  eb = next & valid;
}

// Compacts A (AS-byte elements, 0 = none) at run heads and B (BS-byte
// elements, 0 = none) at run ends. Both outputs hold one element per run;
// the last tile stores the run count in *len_a / *len_b (device lengths).
template <int AS, int BS>
__global__ __launch_bounds__(kByKeyBlock) void bykey_compact_kernel(
    const uint32_t* __restrict__ mask, uint32_t n, const void* __restrict__ src_a,
    void* __restrict__ out_a, const void* __restrict__ src_b, void* __restrict__ out_b,
    uint32_t* len_a, uint32_t* len_b, uint64_t* status, uint32_t gen, uint32_t ntiles) {
  constexpr int NW = kByKeyBlock / kWave;
  __shared__ uint64_t s_wave[NW];
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const uint32_t tile = blockIdx.x;
  const uint32_t e0 = tile * kByKeyTile + tid * kByKeyIpt;
  uint32_t hb, eb;
  head_end_bits(mask, n, e0, hb, eb);
  // Counts packed as (ends << 32) | heads; both sums stay below 2^32.
  const uint64_t mine = ((uint64_t)__builtin_popcount(eb) << 32) | (uint32_t)__builtin_popcount(hb);
  const uint64_t incl = wave_inclusive_scan<uint64_t, OP_ADD>(mine);
  if (lane == 63) s_wave[wid] = incl;
  __syncthreads();
  uint64_t woff = 0, agg = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    if (w < (int)wid) woff += s_wave[w];
    agg += s_wave[w];
  }
  if (wid == 0) {
    TileStatus<uint64_t> st{status, gen};
    uint64_t excl = 0;
    if (tile == 0) {
      if (lane == 0) st.publish(0, kFlagInc, 0, agg);
    } else {
      if (lane == 0) st.publish(tile, kFlagAgg, 0, agg);
      // Self-help: the head and end counts of tile q, from the mask.
      auto count_of = [&](uint32_t q) -> uint64_t {
        uint64_t c = 0;
        for (uint32_t t = lane; t < (uint32_t)kByKeyBlock; t += kWave) {
          uint32_t h, e;
          head_end_bits(mask, n, q * kByKeyTile + t * kByKeyIpt, h, e);
          c += ((uint64_t)__builtin_popcount(e) << 32) | (uint32_t)__builtin_popcount(h);
        }
        return wave_reduce<uint64_t, OP_ADD>(c);
      };
      excl = wave_lookback_sh<uint64_t, OP_ADD>(st, tile, count_of);
      if (lane == 0) st.publish(tile, kFlagInc, 0, excl + agg);
    }
    if (lane == 0) {
      s_prefix = excl;
      if (tile == ntiles - 1) {
        const uint64_t tot = excl + agg;
        if (len_a) *len_a = (uint32_t)tot;
        if (len_b) *len_b = (uint32_t)(tot >> 32);
      }
    }
  }
  __syncthreads();
  const uint64_t pre = s_prefix + woff + incl - mine;
  if constexpr (AS > 0) {
    using W = typename Word<AS>::T;
    uint32_t pa = (uint32_t)pre;
    for (uint32_t m = hb; m; m &= m - 1) {
      const uint32_t j = (uint32_t)__builtin_ctz(m);
      ((W*)out_a)[pa++] = ((const W*)src_a)[e0 + j];
    }
  }
  if constexpr (BS > 0) {
    using W = typename Word<BS>::T;
    uint32_t pb = (uint32_t)(pre >> 32);
    for (uint32_t m = eb; m; m &= m - 1) {
      const uint32_t j = (uint32_t)__builtin_ctz(m);
      ((W*)out_b)[pb++] = ((const W*)src_b)[e0 + j];
    }
  }
}

// out[i] = init (+) out[i] for the first *len elements.
template <typename T, int OP>
__global__ void bykey_init_kernel(T* __restrict__ out, uint32_t cap, const uint32_t* __restrict__ len,
                                  T init) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap && i < *len) out[i] = Combine<T, OP>::apply(init, out[i]);
}

// ---- host ----------------------------------------------------------------------

static msv_status build_mask(Ctx& c, msv_dtype kdt, const void* keys, uint32_t n, uint32_t** mask) {
  *mask = (uint32_t*)c.scratch(sizeof(uint32_t) * ((n + 31) / 32 + 2));
  if (!*mask) return MSV_ERR_OUT_OF_MEMORY;
  if (n == 0) return MSV_OK;
  const dim3 grid((n + 255) / 256), block(256);
  switch (kdt) {
    case MSV_U32:
    case MSV_I32:
      hipLaunchKernelGGL(bykey_heads_kernel<uint32_t>, grid, block, 0, c.stream(), (const uint32_t*)keys, n, *mask);
      break;
    case MSV_F32:
      hipLaunchKernelGGL(bykey_heads_kernel<float>, grid, block, 0, c.stream(), (const float*)keys, n, *mask);
      break;
    case MSV_U64:
    case MSV_I64:
      hipLaunchKernelGGL(bykey_heads_kernel<uint64_t>, grid, block, 0, c.stream(), (const uint64_t*)keys, n, *mask);
      break;
    case MSV_F64:
      hipLaunchKernelGGL(bykey_heads_kernel<double>, grid, block, 0, c.stream(), (const double*)keys, n, *mask);
      break;
  }
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <int AS, int BS>
static msv_status compact(Ctx& c, const uint32_t* mask, uint32_t n, const void* a, void* oa,
                          const void* b, void* ob, uint32_t* la, uint32_t* lb) {
  const uint32_t ntiles = (n + kByKeyTile - 1) / kByKeyTile;
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback((size_t)ntiles * TileStatus<uint64_t>::kWords, &counter, &status, &gen));
  hipLaunchKernelGGL((bykey_compact_kernel<AS, BS>), dim3(ntiles), dim3(kByKeyBlock), 0, c.stream(),
                     mask, n, a, oa, b, ob, la, lb, status, gen, ntiles);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

static __global__ void zero_len_kernel(uint32_t* a, uint32_t* b) {
  if (a) *a = 0;
  if (b) *b = 0;
}

template <typename T>
static msv_status init_apply(Ctx& c, msv_op op, void* out, uint32_t cap, const uint32_t* len,
                             const void* init_host) {
  T init;
  memcpy(&init, init_host, sizeof(T));
  const dim3 grid((cap + 255) / 256), block(256);
  switch (op) {
    case MSV_OP_ADD: hipLaunchKernelGGL((bykey_init_kernel<T, OP_ADD>), grid, block, 0, c.stream(), (T*)out, cap, len, init); break;
    case MSV_OP_MIN: hipLaunchKernelGGL((bykey_init_kernel<T, OP_MIN>), grid, block, 0, c.stream(), (T*)out, cap, len, init); break;
    case MSV_OP_MAX: hipLaunchKernelGGL((bykey_init_kernel<T, OP_MAX>), grid, block, 0, c.stream(), (T*)out, cap, len, init); break;
    default:
      if constexpr (!IsFloat<T>::value)
        hipLaunchKernelGGL((bykey_init_kernel<T, OP_XOR>), grid, block, 0, c.stream(), (T*)out, cap, len, init);
      break;
  }
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

static msv_status init_apply_run(Ctx& c, msv_dtype dt, msv_op op, void* out, uint32_t cap,
                                 const uint32_t* len, const void* init) {
  switch (dt) {
    case MSV_U32: return init_apply<uint32_t>(c, op, out, cap, len, init);
    case MSV_I32: return init_apply<int32_t>(c, op, out, cap, len, init);
    case MSV_F32: return init_apply<float>(c, op, out, cap, len, init);
    case MSV_U64: return init_apply<uint64_t>(c, op, out, cap, len, init);
    case MSV_I64: return init_apply<int64_t>(c, op, out, cap, len, init);
    case MSV_F64: return init_apply<double>(c, op, out, cap, len, init);
  }
  return MSV_ERR_INVALID_ARGUMENT;
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
  uint32_t* mask = nullptr;
  msv_status s = build_mask(c, k.dtype, k.ptr, k.len, &mask);
  if (s == MSV_OK && v.len)
    s = mask_scan_run(c, v.dtype, op, exclusive, v.ptr, v.len, mask, (*out)->ptr, init);
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
    uint32_t* mask = nullptr;
    void* y = c.scratch((size_t)n * dtype_size(v.dtype));
    s = y ? build_mask(c, k.dtype, k.ptr, n, &mask) : MSV_ERR_OUT_OF_MEMORY;
    if (s == MSV_OK) s = mask_scan_run(c, v.dtype, op, false, v.ptr, n, mask, y, nullptr);
    const size_t ks = dtype_size(k.dtype), vs = dtype_size(v.dtype);
    if (s == MSV_OK) {
      void *ok = (*out_keys)->ptr, *ov = (*out_values)->ptr;
      uint32_t *lk = (*out_keys)->d_len, *lv = (*out_values)->d_len;
      if (ks == 4 && vs == 4) s = compact<4, 4>(c, mask, n, k.ptr, ok, y, ov, lk, lv);
      else if (ks == 4) s = compact<4, 8>(c, mask, n, k.ptr, ok, y, ov, lk, lv);
      else if (vs == 4) s = compact<8, 4>(c, mask, n, k.ptr, ok, y, ov, lk, lv);
      else s = compact<8, 8>(c, mask, n, k.ptr, ok, y, ov, lk, lv);
    }
    if (s == MSV_OK)
      s = init_apply_run(c, v.dtype, op, (*out_values)->ptr, n, (*out_values)->d_len, init);
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
    uint32_t* mask = nullptr;
    s = build_mask(c, k.dtype, k.ptr, k.len, &mask);
    if (s == MSV_OK) {
      if (dtype_size(v.dtype) == 4)
        s = compact<4, 0>(c, mask, k.len, v.ptr, (*out_values)->ptr, nullptr, nullptr,
                          (*out_values)->d_len, nullptr);
      else
        s = compact<8, 0>(c, mask, k.len, v.ptr, (*out_values)->ptr, nullptr, nullptr,
                          (*out_values)->d_len, nullptr);
    }
  }
  if (s != MSV_OK) {
    msv_vec_release(*out_values);
    *out_values = nullptr;
  }
  return s;
}

}  // extern "C"
