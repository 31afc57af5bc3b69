// Onesweep LSD radix sort for gfx950: vector::radix_sort_by_key, sort_by_key
// and sort with the built-in comparators.
//
// Reference: massively/src/core/radix.rs:369-546 (4-bit digits, 4 launches
// per pass, every pass reading keys through a permutation, a 1-thread bucket
// scan, then a final gather of the values). Order semantics follow
// core/radix.rs:13-118: unsigned order, signed via the flipped sign bit, and
// IEEE total order for floats. For the comparator forms (`Less`/`Greater`)
// -0.0 and +0.0 compare equal, so they are ordered as equal keys (stably).
//
// Pipeline (keys K of 4 or 8 bytes, optional 4/8-byte values):
//  1. histogram: one read of the keys builds the 256-bin histograms of every
//     8-bit digit at once (LDS histograms, wave-uniform digits added once);
//  2. plan (one block): exclusive scan of each histogram into digit starts;
//     a digit position whose histogram has a single bin of count n is
//     skipped; the ping-pong source/destination of every remaining pass is
//     written into a device-side table, so no host round trip is needed;
//  3. one onesweep kernel per digit: block b owns tile b of 4096 keys (each
//     wave a contiguous 1024-key run, lane l holding elements j*64+l), ranks
//     keys stably inside each wave by matching 8-bit digits with ballots
//     (wave-level multi-split), combines the wave counts into tile digit
//     offsets, resolves each digit's global offset by decoupled lookback over
//     (tile, digit) status words (recounting a predecessor's digits from its
//     keys if its status does not appear), reorders keys (then values)
//     through LDS into digit-sorted runs and writes the runs out.
#include "device_common.hpp"
#include "msv_internal.hpp"

namespace msv {

constexpr int kRadixBlock = 256;
constexpr int kRadixIpt = 16;
constexpr int kRadixWaves = kRadixBlock / kWave;
constexpr uint32_t kRadixTile = kRadixBlock * kRadixIpt;
constexpr int kRadixBins = 256;
constexpr int kHistBlock = 256;

struct RadixPass {
  uint32_t active;
  uint32_t shift;
  const void* src_k;
  const void* src_v;
  void* dst_k;
  void* dst_v;
};

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// 1. All-digit histogram.
template <typename K, int MODE>
__global__ __launch_bounds__(kHistBlock) void radix_hist_kernel(const K* __restrict__ keys,
                                                                uint32_t n,
                                                                uint32_t* __restrict__ g_hist) {
  constexpr int P = sizeof(K);
  __shared__ uint32_t s_hist[P][kRadixBins];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < P * kRadixBins; i += kHistBlock) s_hist[i / kRadixBins][i % kRadixBins] = 0;
  __syncthreads();
  const uint32_t lane = __lane_id();
  const size_t stride = (size_t)gridDim.x * kHistBlock;
  for (size_t i = (size_t)blockIdx.x * kHistBlock + tid; i < n; i += stride) {
    const auto e = order_bits<K, MODE>(keys[i]);
    const uint64_t active = ballot(true);
    const uint32_t first = (uint32_t)__builtin_ctzll(active);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const uint32_t d = (uint32_t)(e >> (8 * p)) & 0xffu;
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
      if (ballot(d == d0) == active) {
        if (lane == first) atomicAdd(&s_hist[p][d0], (uint32_t)__builtin_popcountll(active));
      } else {
        atomicAdd(&s_hist[p][d], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < P * kRadixBins; i += kHistBlock) {
    const uint32_t c = s_hist[i / kRadixBins][i % kRadixBins];
    if (c) atomicAdd(&g_hist[i], c);
  }
}

// 2. Digit starts and the pass table.
template <int P>
__global__ __launch_bounds__(kRadixBins) void radix_plan_kernel(
    const uint32_t* __restrict__ g_hist, uint32_t* __restrict__ g_start, uint32_t n,
    RadixPass* __restrict__ table, const void* in_k, const void* in_v, void* out_k, void* out_v,
    void* tmp_k, void* tmp_v) {
  __shared__ uint32_t s_wave[kRadixBins / kWave];
  __shared__ uint32_t s_trivial[P];
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t / 64u;
  for (int p = 0; p < P; ++p) {
    const uint32_t c = g_hist[p * kRadixBins + t];
    const int triv = __syncthreads_or(c == n);
    const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(c);
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < w; ++k) off += s_wave[k];
    g_start[p * kRadixBins + t] = off + incl - c;
    if (t == 0) s_trivial[p] = triv ? 1u : 0u;
    __syncthreads();
  }
  if (t == 0) {
    int active = 0;
    for (int p = 0; p < P; ++p) active += s_trivial[p] ? 0 : 1;
    // All digits constant (all keys equal): one pass still moves the data.
    const bool force_last = active == 0;
    if (force_last) active = 1;
    int j = 0;
    const void* src_k = in_k;
    const void* src_v = in_v;
    for (int p = 0; p < P; ++p) {
      RadixPass rp;
      rp.shift = 8u * (uint32_t)p;
      rp.active = (!s_trivial[p] || (force_last && p == P - 1)) ? 1u : 0u;
      rp.src_k = src_k;
      rp.src_v = src_v;
      rp.dst_k = nullptr;
      rp.dst_v = nullptr;
      if (rp.active) {
        const bool to_out = ((active - 1 - j) % 2) == 0;
        rp.dst_k = to_out ? out_k : tmp_k;
        rp.dst_v = to_out ? out_v : tmp_v;
        src_k = rp.dst_k;
        src_v = rp.dst_v;
        ++j;
      }
      table[p] = rp;
    }
  }
}

template <int VS>
struct ValueType { using T = uint32_t; };
template <>
struct ValueType<8> { using T = uint64_t; };

// 3. One onesweep pass.
template <typename K, int MODE, int VS>
__global__ __launch_bounds__(kRadixBlock) void radix_pass_kernel(
    const RadixPass* __restrict__ table, int pass, const uint32_t* __restrict__ g_start, uint32_t n,
    uint64_t* status, uint32_t gen, uint32_t ntiles) {
  using U = typename Bits<K>::U;
  using VT = typename ValueType<VS>::T;
  constexpr int SLOT = (VS > (int)sizeof(K)) ? VS : (int)sizeof(K);
  __shared__ __align__(16) unsigned char s_stage_raw[kRadixTile * SLOT];
  __shared__ uint32_t s_hist[kRadixWaves][kRadixBins];  // per-wave counts, then wave offsets
  __shared__ uint32_t s_tile_start[kRadixBins];         // in-tile start of each digit
  __shared__ uint32_t s_goff[kRadixBins];               // global position - in-tile position
  __shared__ uint32_t s_wsum[kRadixBins / kWave];
  __shared__ uint32_t s_help[kRadixBins];               // self-help histogram

  const RadixPass rp = table[pass];
  if (!rp.active) return;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  for (uint32_t i = tid; i < kRadixWaves * kRadixBins; i += kRadixBlock) (&s_hist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * kRadixTile + (size_t)wid * (kWave * kRadixIpt);
  const K* __restrict__ src_k = (const K*)rp.src_k;
  const uint32_t shift = rp.shift;

  K key[kRadixIpt];
  uint32_t dig[kRadixIpt];
  uint32_t rank[kRadixIpt];
#pragma unroll
  for (int j = 0; j < kRadixIpt; ++j) {
    const size_t i = base + (size_t)j * kWave + lane;
    key[j] = i < n ? src_k[i] : (K)0;
    dig[j] = i < n ? ((uint32_t)(order_bits<K, MODE>(key[j]) >> shift) & 0xffu) : 0x100u;
  }

  // Wave-level multi-split: rank each key among the wave's earlier keys with
  // the same digit (input order: j-major, lane-minor).
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kRadixIpt; ++j) {
    const uint32_t d = dig[j];
    const bool valid = d < 0x100u;
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = ballot(bit);
      peers &= bit ? m : ~m;
    }
    uint32_t before = 0;
    if (valid) before = s_hist[wid][d];
    const uint32_t below = (uint32_t)__builtin_popcountll(peers & lt);
    if (valid && below == 0) s_hist[wid][d] = before + (uint32_t)__builtin_popcountll(peers);
    rank[j] = before + below;
  }
  __syncthreads();

  // Per digit: wave offsets and the tile count (thread t <-> digit t).
  const uint32_t d_me = tid;
  uint32_t cnt = 0;
#pragma unroll
  for (int w = 0; w < kRadixWaves; ++w) {
    const uint32_t c = s_hist[w][d_me];
    s_hist[w][d_me] = cnt;
    cnt += c;
  }
  TileStatus<uint32_t> st{status, gen};
  const uint32_t sidx = tile * kRadixBins + d_me;
  if (tile == 0) st.publish(sidx, kFlagInc, 0, cnt);
  else st.publish(sidx, kFlagAgg, 0, cnt);

  // In-tile digit starts (block exclusive scan over the 256 digits).
  const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(cnt);
  if (lane == 63) s_wsum[wid] = incl;
  __syncthreads();
  uint32_t woff = 0;
  for (uint32_t k = 0; k < wid; ++k) woff += s_wsum[k];
  const uint32_t tstart = woff + incl - cnt;
  s_tile_start[d_me] = tstart;

  // Decoupled lookback, one digit per thread, stepping the whole block through
  // the predecessor tiles together: at tile p every digit not yet resolved
  // adds p's count and stops at an inclusive prefix. If p's status does not
  // appear within kSelfHelpPolls polls, the block counts p's digits itself.
  uint32_t excl = 0;
  if (tile > 0) {
    bool done = false;
    uint32_t p = tile - 1;
    uint32_t polls = 0;
    for (;;) {
      uint32_t f = 0, v = 0, head;
      if (!done) f = st.try_read(p * kRadixBins + d_me, head, v);
      const bool waiting = !done && f == 0;
      if (__syncthreads_or(waiting)) {
        if (++polls < kSelfHelpPolls) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        s_help[d_me] = 0;
        __syncthreads();
        for (int j = 0; j < kRadixIpt; ++j) {
          const size_t i = (size_t)p * kRadixTile + (size_t)j * kRadixBlock + tid;
          if (i < n) atomicAdd(&s_help[(uint32_t)(order_bits<K, MODE>(src_k[i]) >> shift) & 0xffu], 1u);
        }
        __syncthreads();
        if (waiting) {
          f = kFlagAgg;
          v = s_help[d_me];
        }
      }
      polls = 0;
      if (!done) {
        excl += v;
        if (f == kFlagInc || p == 0) done = true;
      }
      if (!__syncthreads_or(!done)) break;
      --p;
    }
    st.publish(sidx, kFlagInc, 0, excl + cnt);
  }
  s_goff[d_me] = g_start[(shift / 8u) * kRadixBins + d_me] + excl - tstart;
  __syncthreads();

  // Reorder keys into digit-sorted runs in LDS, then write the runs out.
  K* s_keys = reinterpret_cast<K*>(s_stage_raw);
  uint32_t pos[kRadixIpt];
#pragma unroll
  for (int j = 0; j < kRadixIpt; ++j) {
    if (dig[j] < 0x100u) {
      pos[j] = s_tile_start[dig[j]] + s_hist[wid][dig[j]] + rank[j];
      s_keys[pos[j]] = key[j];
    }
  }
  __syncthreads();
  const uint32_t tile_n = (uint32_t)min((size_t)kRadixTile, (size_t)n - (size_t)tile * kRadixTile);
  K* __restrict__ dst_k = (K*)rp.dst_k;
  uint32_t gpos[kRadixIpt];
#pragma unroll
  for (int q = 0; q < kRadixIpt; ++q) {
    const uint32_t i = tid + (uint32_t)q * kRadixBlock;
    if (i < tile_n) {
      const K k = s_keys[i];
      const uint32_t d = (uint32_t)(order_bits<K, MODE>(k) >> shift) & 0xffu;
      gpos[q] = s_goff[d] + i;
      dst_k[gpos[q]] = k;
    }
  }
  if constexpr (VS > 0) {
    const VT* __restrict__ src_v = (const VT*)rp.src_v;
    VT* __restrict__ dst_v = (VT*)rp.dst_v;
    VT val[kRadixIpt];
#pragma unroll
    for (int j = 0; j < kRadixIpt; ++j) {
      const size_t i = base + (size_t)j * kWave + lane;
      if (i < n) val[j] = src_v[i];
    }
    __syncthreads();
    VT* s_vals = reinterpret_cast<VT*>(s_stage_raw);
#pragma unroll
    for (int j = 0; j < kRadixIpt; ++j)
      if (dig[j] < 0x100u) s_vals[pos[j]] = val[j];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRadixIpt; ++q) {
      const uint32_t i = tid + (uint32_t)q * kRadixBlock;
      if (i < tile_n) dst_v[gpos[q]] = s_vals[i];
    }
  }
}

// ---- host -------------------------------------------------------------------
template <typename K, int MODE, int VS>
static msv_status radix_sort_impl(Ctx& c, const void* keys, const void* vals, uint32_t n,
                                  void* out_k, void* out_v) {
  constexpr int P = sizeof(K);
  if (n == 0) return MSV_OK;
  const size_t vsz = VS;
  void* tmp_k = c.scratch((size_t)n * sizeof(K));
  void* tmp_v = VS ? c.scratch((size_t)n * vsz) : nullptr;
  uint32_t* hist = (uint32_t*)c.scratch(sizeof(uint32_t) * P * kRadixBins * 2);
  RadixPass* table = (RadixPass*)c.scratch(sizeof(RadixPass) * P);
  if (!tmp_k || (VS && !tmp_v) || !hist || !table) return MSV_ERR_OUT_OF_MEMORY;
  uint32_t* start = hist + P * kRadixBins;

  hipLaunchKernelGGL(fill_u32_kernel, dim3(P), dim3(kRadixBins), 0, c.stream(), hist,
                     (uint32_t)(P * kRadixBins), 0u);
  MSV_HIP(hipGetLastError());
  const uint32_t hblocks =
      (uint32_t)std::min<size_t>((n + kHistBlock * 16 - 1) / (kHistBlock * 16), (size_t)c.num_cus() * 8);
  hipLaunchKernelGGL((radix_hist_kernel<K, MODE>), dim3(std::max(1u, hblocks)), dim3(kHistBlock), 0,
                     c.stream(), (const K*)keys, n, hist);
  MSV_HIP(hipGetLastError());
  hipLaunchKernelGGL((radix_plan_kernel<P>), dim3(1), dim3(kRadixBins), 0, c.stream(), hist, start,
                     n, table, keys, vals, out_k, out_v, tmp_k, tmp_v);
  MSV_HIP(hipGetLastError());
  const uint32_t ntiles = (n + kRadixTile - 1) / kRadixTile;
  for (int p = 0; p < P; ++p) {
    uint64_t *counter, *status;
    uint32_t gen;
    MSV_TRY(c.lookback((size_t)ntiles * kRadixBins, &counter, &status, &gen));
    hipLaunchKernelGGL((radix_pass_kernel<K, MODE, VS>), dim3(ntiles), dim3(kRadixBlock), 0,
                       c.stream(), table, p, start, n, status, gen, ntiles);
    MSV_HIP(hipGetLastError());
  }
  return MSV_OK;
}

template <typename K, int MODE>
static msv_status radix_by_vsize(Ctx& c, size_t vs, const void* keys, const void* vals,
                                 uint32_t n, void* out_k, void* out_v) {
  if (vs == 0) return radix_sort_impl<K, MODE, 0>(c, keys, nullptr, n, out_k, nullptr);
  if (vs == 4) return radix_sort_impl<K, MODE, 4>(c, keys, vals, n, out_k, out_v);
  return radix_sort_impl<K, MODE, 8>(c, keys, vals, n, out_k, out_v);
}

template <typename K>
static msv_status radix_by_mode(Ctx& c, int mode, size_t vs, const void* keys, const void* vals,
                                uint32_t n, void* out_k, void* out_v) {
  if (mode == ORD_GREATER) return radix_by_vsize<K, ORD_GREATER>(c, vs, keys, vals, n, out_k, out_v);
  if constexpr (IsFloat<K>::value) {
    if (mode == ORD_LESS) return radix_by_vsize<K, ORD_LESS>(c, vs, keys, vals, n, out_k, out_v);
  }
  return radix_by_vsize<K, ORD_TOTAL>(c, vs, keys, vals, n, out_k, out_v);
}

msv_status radix_sort_run(Ctx& c, msv_dtype kdt, int mode, size_t vsize, const void* keys,
                          const void* vals, uint32_t n, void* out_k, void* out_v) {
  switch (kdt) {
    case MSV_U32: return radix_by_mode<uint32_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_I32: return radix_by_mode<int32_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_F32: return radix_by_mode<float>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_U64: return radix_by_mode<uint64_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_I64: return radix_by_mode<int64_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_F64: return radix_by_mode<double>(c, mode, vsize, keys, vals, n, out_k, out_v);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad key dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

}  // namespace msv

using namespace msv;

// Shared front end of the three sorting entry points.
static msv_status sort_api(msv_executor* ex, msv_slice keys, const msv_slice* values, int mode,
                           msv_vec** out_values, msv_vec** out_keys) {
  clear_error();
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  if (values) {
    MSV_TRY(resolve(ex, *values, &v));
    if (k.len != v.len) {
      set_error(MSV_ERR_LENGTH_MISMATCH, k.len, v.len, "input lengths differ: %u != %u", k.len,
                v.len);
      return MSV_ERR_LENGTH_MISMATCH;
    }
    if (!out_values) {
      set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
      return MSV_ERR_INVALID_ARGUMENT;
    }
  }
  msv_vec* ov = nullptr;
  msv_vec* ok = nullptr;
  if (values) MSV_TRY(new_vec(ex, v.dtype, v.len, false, &ov));
  Ctx c(ex);
  void* kdst = nullptr;
  if (out_keys) {
    msv_status s = new_vec(ex, k.dtype, k.len, false, &ok);
    if (s != MSV_OK) {
      if (ov) msv_vec_release(ov);
      return s;
    }
    kdst = ok->ptr;
  } else {
    kdst = c.scratch((size_t)k.len * dtype_size(k.dtype) + 16);
    if (!kdst) {
      if (ov) msv_vec_release(ov);
      return MSV_ERR_OUT_OF_MEMORY;
    }
  }
  msv_status s = radix_sort_run(c, k.dtype, mode, values ? dtype_size(v.dtype) : 0, k.ptr,
                                values ? v.ptr : nullptr, k.len, kdst, ov ? ov->ptr : nullptr);
  if (s != MSV_OK) {
    if (ov) msv_vec_release(ov);
    if (ok) msv_vec_release(ok);
    return s;
  }
  if (out_values) *out_values = ov;
  if (out_keys) *out_keys = ok;
  return MSV_OK;
}

extern "C" {

msv_status msv_radix_sort_by_key(msv_executor* ex, msv_slice keys, msv_slice values,
                                 msv_vec** out_values, msv_vec** out_keys) {
  return sort_api(ex, keys, &values, ORD_TOTAL, out_values, out_keys);
}

msv_status msv_sort_by_key(msv_executor* ex, msv_slice keys, msv_slice values, msv_cmp cmp,
                           msv_vec** out_values, msv_vec** out_keys) {
  if (cmp != MSV_CMP_LESS && cmp != MSV_CMP_GREATER) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown comparator %d", (int)cmp);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return sort_api(ex, keys, &values, cmp == MSV_CMP_LESS ? ORD_LESS : ORD_GREATER, out_values,
                  out_keys);
}

msv_status msv_sort(msv_executor* ex, msv_slice input, msv_cmp cmp, msv_vec** out) {
  if (cmp != MSV_CMP_LESS && cmp != MSV_CMP_GREATER) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown comparator %d", (int)cmp);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (!out) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return sort_api(ex, input, nullptr, cmp == MSV_CMP_LESS ? ORD_LESS : ORD_GREATER, nullptr, out);
}

}  // extern "C"
