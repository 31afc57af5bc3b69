// LSD radix sort kernels (device code), shared by radix.hip and the
// tuning probes. See radix.hip for the pipeline description.
#pragma once

#include "device_common.hpp"

namespace msv {

constexpr int kRadixIpt = 16;     // keys per thread in a pass tile
constexpr int kLookbackBatch = 4;  // predecessors a onesweep look-back step reads
constexpr int kRadixBins = 256;   // 8-bit digits

// Element moved by the passes: the key (1 or 2 dwords) followed by the value
// (0, 1 or 2 dwords). Between passes elements are kept interleaved (AoS), so
// each pass reorders one array and writes one stream of runs; the first pass
// reads the caller's separate key/value arrays and the last writes them.
template <int EW>
struct alignas(4) Elem {
  uint32_t w[EW];
};

template <typename K, int EW>
__device__ inline K elem_key(const Elem<EW>& e) {
  if constexpr (sizeof(K) == 4) return bit_cast<K>(e.w[0]);
  else return bit_cast<K>(((uint64_t)e.w[1] << 32) | e.w[0]);
}

struct RadixPass {
  uint32_t active;
  uint32_t shift;
  uint32_t in_aos;   // source is an element array (else separate keys/values)
  uint32_t out_aos;  // destination is an element array
  const void* src_k;
  const void* src_v;
  void* dst_k;
  void* dst_v;
};

// Lanes of the wave holding a valid digit equal to ours (the low DBITS bits of
// d). One ballot per digit bit; a lane differs from us where its ballot bit
// differs from our own bit, so the differences are collected with 32-bit
// XOR/OR against our bit spread to 0 / all-ones (signed bit extract).
template <int DBITS>
__device__ inline uint64_t match_digit(uint32_t d, bool valid) {
  uint32_t diff_lo = 0, diff_hi = 0;
#pragma unroll
  for (int b = 0; b < DBITS; ++b) {
    const uint32_t s = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);
    const uint64_t m = ballot(s != 0);
    diff_lo |= (uint32_t)m ^ s;
    diff_hi |= (uint32_t)(m >> 32) ^ s;
  }
  return ballot(valid) & ~(((uint64_t)diff_hi << 32) | diff_lo);
}

static __global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// 1. First upsweep (pass 0's digit counts, see 3a) from the caller's keys,
//    which also records per tile the OR and the AND of the keys' order
//    words: a digit position on which OR ^ AND is zero holds one value for
//    every key, and its pass is skipped. (An all-digit histogram would need
//    one LDS atomic per key per digit; two bitwise reductions per tile do.)
template <typename K, int MODE, int BLOCK, int IPT = kRadixIpt>
__global__ __launch_bounds__(BLOCK) void radix_upsweep0_kernel(const K* __restrict__ keys,
                                                              uint32_t n, uint32_t ntiles,
                                                              uint32_t* __restrict__ counts,
                                                              uint64_t* __restrict__ orand) {
  using U = typename Bits<K>::U;
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  __shared__ uint32_t s_cnt[NW][kRadixBins];
  __shared__ uint64_t s_or[NW], s_and[NW];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  for (uint32_t i = tid; i < NW * kRadixBins; i += BLOCK) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * TILE;
  uint64_t o = 0, a = ~0ull;
  uint32_t dig[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const size_t i = base + (size_t)j * BLOCK + tid;
    const U w = order_bits<K, MODE>(keys[i < n ? i : (size_t)n - 1]);
    const bool ok = i < n;
    o |= ok ? (uint64_t)w : 0ull;
    a &= ok ? (uint64_t)w : ~0ull;
    dig[j] = ok ? ((uint32_t)w & 0xffu) : 0x100u;
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (dig[j] < 0x100u) atomicAdd(&s_cnt[wid][dig[j]], 1u);
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    o |= __shfl_xor(o, m);
    a &= __shfl_xor(a, m);
  }
  if (lane == 0) {
    s_or[wid] = o;
    s_and[wid] = a;
  }
  __syncthreads();
  for (uint32_t d = tid; d < (uint32_t)kRadixBins; d += BLOCK) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) c += s_cnt[w][d];
    counts[(size_t)d * ntiles + blockIdx.x] = c;
  }
  if (tid == 0) {
    uint64_t ro = 0, ra = ~0ull;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      ro |= s_or[w];
      ra &= s_and[w];
    }
    orand[2 * (size_t)blockIdx.x] = ro;
    orand[2 * (size_t)blockIdx.x + 1] = ra;
  }
}

// 2. Pass plan: which digit positions need a pass (a position on which the
//    OR and the AND of all order words agree is skipped) and the source and
//    destination of each remaining pass: caller arrays -> element buffer A0
//    -> A1 -> A0 ... -> caller arrays. `orand` holds (OR, AND) per tile of
//    the first upsweep.
template <int P, bool SOA = false>
__global__ __launch_bounds__(kRadixBins) void radix_plan_kernel(
    const uint64_t* __restrict__ orand, uint32_t ntiles, uint32_t n, RadixPass* __restrict__ table,
    const void* in_k, const void* in_v, void* out_k, void* out_v, void* aos0, void* aos1) {
  __shared__ uint32_t s_trivial[P];
  __shared__ uint64_t s_or[kRadixBins / kWave], s_and[kRadixBins / kWave];
  const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
  uint64_t o = 0, a = ~0ull;
  for (uint32_t q = t; q < ntiles; q += kRadixBins) {
    o |= orand[2 * (size_t)q];
    a &= orand[2 * (size_t)q + 1];
  }
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    o |= __shfl_xor(o, m);
    a &= __shfl_xor(a, m);
  }
  if (lane == 0) {
    s_or[wid] = o;
    s_and[wid] = a;
  }
  __syncthreads();
  if (t == 0) {
    uint64_t ro = 0, ra = ~0ull;
    for (int w = 0; w < kRadixBins / kWave; ++w) {
      ro |= s_or[w];
      ra &= s_and[w];
    }
    const uint64_t varying = ro ^ ra;
    for (int p = 0; p < P; ++p) s_trivial[p] = ((varying >> (8 * p)) & 0xffu) == 0 ? 1u : 0u;
  }
  __syncthreads();
  if (t == 0) {
    int active = 0;
    for (int p = 0; p < P; ++p) active += s_trivial[p] ? 0 : 1;
    // All digits constant (all keys equal): one pass still moves the data.
    const bool force_last = active == 0;
    if (force_last) active = 1;
    int j = 0;
    for (int p = 0; p < P; ++p) {
      RadixPass rp;
      rp.shift = 8u * (uint32_t)p;
      rp.active = (!s_trivial[p] || (force_last && p == P - 1)) ? 1u : 0u;
      rp.in_aos = rp.out_aos = 0;
      rp.src_k = rp.src_v = nullptr;
      rp.dst_k = rp.dst_v = nullptr;
      if (rp.active) {
        // SOA: an intermediate buffer holds the n keys followed by the values.
        rp.in_aos = !SOA && j > 0;
        rp.out_aos = !SOA && j < active - 1;
        if (j == 0) {
          rp.src_k = in_k;
          rp.src_v = in_v;
        } else {
          char* b = (char*)(((j - 1) & 1) ? aos1 : aos0);
          rp.src_k = b;
          rp.src_v = SOA ? b + (size_t)n * P : nullptr;
        }
        if (j == active - 1) {
          rp.dst_k = out_k;
          rp.dst_v = out_v;
        } else {
          char* b = (char*)((j & 1) ? aos1 : aos0);
          rp.dst_k = b;
          rp.dst_v = SOA ? b + (size_t)n * P : nullptr;
        }
        ++j;
      }
      table[p] = rp;
    }
  }
}

template <typename K, int EW>
__device__ inline K load_key(const RadixPass& rp, size_t i) {
  if (rp.in_aos) return elem_key<K, EW>(((const Elem<EW>*)rp.src_k)[i]);
  return ((const K*)rp.src_k)[i];
}

template <typename K, int EW>
__device__ inline Elem<EW> load_elem(const RadixPass& rp, size_t i) {
  if (rp.in_aos) return ((const Elem<EW>*)rp.src_k)[i];
  Elem<EW> e;
  const K k = ((const K*)rp.src_k)[i];
  if constexpr (sizeof(K) == 4) {
    e.w[0] = bit_cast<uint32_t>(k);
  } else {
    const uint64_t u = bit_cast<uint64_t>(k);
    e.w[0] = (uint32_t)u;
    e.w[1] = (uint32_t)(u >> 32);
  }
  constexpr int KW = sizeof(K) / 4;
#pragma unroll
  for (int v = 0; v < EW - KW; ++v) e.w[KW + v] = ((const uint32_t*)rp.src_v)[i * (EW - KW) + v];
  return e;
}

template <typename K, int EW>
__device__ inline void store_elem(const RadixPass& rp, size_t i, const Elem<EW>& e) {
  if (rp.out_aos) {
    ((Elem<EW>*)rp.dst_k)[i] = e;
    return;
  }
  constexpr int KW = sizeof(K) / 4;
  ((K*)rp.dst_k)[i] = elem_key<K, EW>(e);
#pragma unroll
  for (int v = 0; v < EW - KW; ++v) ((uint32_t*)rp.dst_v)[i * (EW - KW) + v] = e.w[KW + v];
}

// The key or the value part of element i of the pass's destination.
template <typename K, int EW>
__device__ inline void store_key_part(const RadixPass& rp, size_t i, K k) {
  if (rp.out_aos) {
    uint32_t* w = ((Elem<EW>*)rp.dst_k)[i].w;
    if constexpr (sizeof(K) == 4) {
      w[0] = bit_cast<uint32_t>(k);
    } else {
      const uint64_t u = bit_cast<uint64_t>(k);
      w[0] = (uint32_t)u;
      w[1] = (uint32_t)(u >> 32);
    }
    return;
  }
  ((K*)rp.dst_k)[i] = k;
}
template <typename K, int EW, typename V>
__device__ inline void store_val_part(const RadixPass& rp, size_t i, V v) {
  constexpr int KW = sizeof(K) / 4;
  if (rp.out_aos) {
    uint32_t* w = ((Elem<EW>*)rp.dst_k)[i].w;
    if constexpr (sizeof(V) == 4) {
      w[KW] = v;
    } else {
      w[KW] = (uint32_t)v;
      w[KW + 1] = (uint32_t)(v >> 32);
    }
    return;
  }
  ((V*)rp.dst_v)[i] = v;
}

// 3. Per-pass reduce-then-scan over tiles of BLOCK * IPT keys.
//  3a. upsweep: digit counts of every tile, digit-major
//      (counts[d * ntiles + tile]), so that one exclusive scan of `counts`
//      yields the global position of each (digit, tile) run.
template <typename K, int MODE, int EW, int BLOCK, int IPT = kRadixIpt>
__global__ __launch_bounds__(BLOCK) void radix_upsweep_kernel(const RadixPass* __restrict__ table,
                                                             int pass, uint32_t n, uint32_t ntiles,
                                                             uint32_t* __restrict__ counts) {
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  constexpr int KW = sizeof(K) / 4;
  static_assert(IPT % 4 == 0, "groups of 4 elements");
  __shared__ uint32_t s_cnt[NW][kRadixBins];
  const RadixPass rp = table[pass];
  if (!rp.active) return;
  const uint32_t tid = threadIdx.x, wid = tid / 64u;
  for (uint32_t i = tid; i < NW * kRadixBins; i += BLOCK) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * TILE;
  const uint32_t shift = rp.shift;
  const bool full = base + TILE <= n;
  // A full tile whose source is 16-byte aligned is read as groups of 4 whole
  // elements (EW dwordx4 loads for an element array, KW for a key array);
  // counting does not depend on the order in which keys are seen.
  const bool vec = full && (((uintptr_t)rp.src_k & 15u) == 0);
  uint32_t dig[IPT];
  if (vec) {
#pragma unroll
    for (int r = 0; r < IPT / 4; ++r) {
      const size_t g = base / 4 + (size_t)r * BLOCK + tid;  // group of elements 4g .. 4g+3
      uint32_t w[4 * (EW > KW ? EW : KW)];
      const int QW = rp.in_aos ? EW : KW;  // dwordx4 loads per group
      const uint4* q = (const uint4*)rp.src_k + g * (size_t)QW;
#pragma unroll
      for (int t = 0; t < (EW > KW ? EW : KW); ++t) {
        if (t < QW) {
          const uint4 v = q[t];
          w[4 * t] = v.x;
          w[4 * t + 1] = v.y;
          w[4 * t + 2] = v.z;
          w[4 * t + 3] = v.w;
        }
      }
      const int stride = rp.in_aos ? EW : KW;  // dwords per element in the group
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        K k;
        if constexpr (KW == 1) {
          k = bit_cast<K>(w[m * stride]);
        } else {
          k = bit_cast<K>(((uint64_t)w[m * stride + 1] << 32) | w[m * stride]);
        }
        dig[4 * r + m] = (uint32_t)(order_bits<K, MODE>(k) >> shift) & 0xffu;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const size_t i = base + (size_t)j * BLOCK + tid;
      dig[j] = i < n ? ((uint32_t)(order_bits<K, MODE>(load_key<K, EW>(rp, i)) >> shift) & 0xffu)
                     : 0x100u;
    }
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (dig[j] < 0x100u) atomicAdd(&s_cnt[wid][dig[j]], 1u);
  __syncthreads();
  for (uint32_t d = tid; d < (uint32_t)kRadixBins; d += BLOCK) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) c += s_cnt[w][d];
    counts[(size_t)d * ntiles + blockIdx.x] = c;
  }
}

//  3b. downsweep: rank the tile's keys stably, reorder the elements through
//      LDS into digit-sorted runs, and write the runs to their offsets
//      (offsets[d * ntiles + tile], the exclusive scan of `counts`).
template <typename K, int MODE, int EW, int BLOCK, int IPT = kRadixIpt>
__global__ __launch_bounds__(BLOCK) void radix_downsweep_kernel(
    const RadixPass* __restrict__ table, int pass, const uint32_t* __restrict__ offsets,
    uint32_t n, uint32_t ntiles) {
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  __shared__ Elem<EW> s_el[TILE];
  __shared__ uint32_t s_hist[NW][kRadixBins];     // per-wave counts, then wave offsets
  __shared__ uint32_t s_tile_start[kRadixBins];   // in-tile start of each digit
  __shared__ uint32_t s_goff[kRadixBins];         // global position - in-tile position
  __shared__ uint32_t s_wsum[kRadixBins / kWave];

  const RadixPass rp = table[pass];
  if (!rp.active) return;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const bool digit_thread = tid < (uint32_t)kRadixBins;
  for (uint32_t i = tid; i < NW * kRadixBins; i += BLOCK) (&s_hist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * TILE + (size_t)wid * (kWave * IPT);
  const uint32_t shift = rp.shift;

  Elem<EW> el[IPT];
  uint32_t dig[IPT];
  uint32_t rank[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const size_t i = base + (size_t)j * kWave + lane;
    if (i < n) {
      el[j] = load_elem<K, EW>(rp, i);
      dig[j] = (uint32_t)(order_bits<K, MODE>(elem_key<K, EW>(el[j])) >> shift) & 0xffu;
    } else {
      dig[j] = 0x100u;
    }
  }
  uint32_t goff = 0;
  if (digit_thread) goff = offsets[(size_t)tid * ntiles + tile];

  // Wave-level multi-split: rank each key among the wave's earlier keys with
  // the same digit (input order: j-major, lane-minor).
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t d = dig[j];
    const bool valid = d < 0x100u;
    const uint64_t peers = match_digit<8>(d, valid);
    uint32_t before = 0;
    if (valid) before = s_hist[wid][d];
    const uint32_t below = (uint32_t)__builtin_popcountll(peers & lt);
    if (valid && below == 0) s_hist[wid][d] = before + (uint32_t)__builtin_popcountll(peers);
    rank[j] = before + below;
  }
  __syncthreads();

  // Per digit (thread t <-> digit t): wave offsets, tile count, tile start.
  uint32_t cnt = 0;
  if (digit_thread) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t c = s_hist[w][tid];
      s_hist[w][tid] = cnt;
      cnt += c;
    }
  }
  uint32_t incl = 0;
  if (digit_thread) {
    incl = wave_inclusive_scan<uint32_t, OP_ADD>(cnt);
    if (lane == 63) s_wsum[wid] = incl;
  }
  __syncthreads();
  if (digit_thread) {
    uint32_t woff = 0;
    for (uint32_t k = 0; k < wid; ++k) woff += s_wsum[k];
    const uint32_t tstart = woff + incl - cnt;
    s_tile_start[tid] = tstart;
    s_goff[tid] = goff - tstart;
  }
  __syncthreads();

  // Reorder elements into digit-sorted runs in LDS, then write the runs out.
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (dig[j] < 0x100u) s_el[s_tile_start[dig[j]] + s_hist[wid][dig[j]] + rank[j]] = el[j];
  __syncthreads();
  const uint32_t tile_n = (uint32_t)min((size_t)TILE, (size_t)n - (size_t)tile * TILE);
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const uint32_t i = tid + (uint32_t)q * BLOCK;
    if (i < tile_n) {
      const Elem<EW> e = s_el[i];
      const uint32_t d = (uint32_t)(order_bits<K, MODE>(elem_key<K, EW>(e)) >> shift) & 0xffu;
      store_elem<K, EW>(rp, (size_t)s_goff[d] + i, e);
    }
  }
}


// ---- onesweep: one histogram pass, then one pass per digit ---------------------
// 5. Global digit histograms of every digit position in one read of the keys:
//    hist[p * 256 + d] = number of keys whose digit p is d. Each thread reads
//    16 keys per step with 16-byte loads (all in flight before the first is
//    counted), each wave counts into its own LDS histogram, then the block adds
//    its counts to the global histogram. On a full step the wave ORs every
//    key's difference from its first key once (a wave reduction), and a digit
//    position that is constant over the step's 1024 keys (constant bytes are
//    common: small keys, a constant top byte) is counted by one add; the other
//    positions take one LDS add per key and no per-key test (per key and
//    position, a uniformity test had cost more instructions than the add).
template <typename K, int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void radix_hist_kernel(const K* __restrict__ keys, uint32_t n,
                                                          uint32_t* __restrict__ hist) {
  using U = typename Bits<K>::U;
  constexpr int P = sizeof(K);
  constexpr int KPT = 16;                   // keys per thread per step
  constexpr int V = 16 / sizeof(K);         // keys per 16-byte load
  constexpr int NW = BLOCK / kWave;
  __shared__ uint32_t s_hw[NW][P * kRadixBins];  // per-wave histograms
  const uint32_t tid = threadIdx.x;
  uint32_t* s_h = s_hw[tid / kWave];
  for (uint32_t i = tid; i < (uint32_t)(NW * P * kRadixBins); i += BLOCK) (&s_hw[0][0])[i] = 0;
  __syncthreads();
  const bool aligned = ((uintptr_t)keys & 15u) == 0;
  const size_t step = (size_t)gridDim.x * BLOCK * KPT;
  for (size_t b0 = (size_t)blockIdx.x * BLOCK * KPT; b0 < n; b0 += step) {
    U w[KPT];
    bool ok[KPT];
    // Row-striped: load r of thread t covers keys b0 + (r * BLOCK + t) * V ...
    if (aligned && b0 + (size_t)BLOCK * KPT <= n) {
#pragma unroll
      for (int r = 0; r < KPT / V; ++r) {
        const uint4 q = *reinterpret_cast<const uint4*>(keys + b0 + ((size_t)r * BLOCK + tid) * V);
        const uint32_t qq[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < V; ++k) {
          K key;
          if constexpr (sizeof(K) == 4) key = bit_cast<K>(qq[k]);
          else key = bit_cast<K>(((uint64_t)qq[2 * k + 1] << 32) | qq[2 * k]);
          w[r * V + k] = order_bits<K, MODE>(key);
        }
      }
      // Bytes in which some key of the wave's step differs from its first key.
      uint32_t lo0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)w[0]);
      uint32_t hi0 = 0, xlo = 0, xhi = 0;
      if constexpr (sizeof(U) == 8)
        hi0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)w[0] >> 32));
#pragma unroll
      for (int j = 0; j < KPT; ++j) {
        xlo |= (uint32_t)w[j] ^ lo0;
        if constexpr (sizeof(U) == 8) xhi |= (uint32_t)((uint64_t)w[j] >> 32) ^ hi0;
      }
#pragma unroll
      for (int k = 1; k < kWave; k <<= 1) {
        xlo |= (uint32_t)__shfl_xor((int)xlo, k);
        if constexpr (sizeof(U) == 8) xhi |= (uint32_t)__shfl_xor((int)xhi, k);
      }
      xlo = (uint32_t)__builtin_amdgcn_readfirstlane((int)xlo);
      xhi = (uint32_t)__builtin_amdgcn_readfirstlane((int)xhi);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const uint32_t sh = 8u * (uint32_t)(p & 3);
        const uint32_t diff = ((p < 4 ? xlo : xhi) >> sh) & 0xffu;
        if (diff == 0) {
          const uint32_t d0 = ((p < 4 ? lo0 : hi0) >> sh) & 0xffu;
          if (__lane_id() == 0) atomicAdd(&s_h[p * kRadixBins + d0], (uint32_t)(kWave * KPT));
        } else {
#pragma unroll
          for (int j = 0; j < KPT; ++j)
            atomicAdd(&s_h[p * kRadixBins + ((uint32_t)(w[j] >> (8 * p)) & 0xffu)], 1u);
        }
      }
      continue;
    } else {
#pragma unroll
      for (int r = 0; r < KPT / V; ++r)
#pragma unroll
        for (int k = 0; k < V; ++k) {
          const size_t i = b0 + ((size_t)r * BLOCK + tid) * V + k;
          ok[r * V + k] = i < n;
          w[r * V + k] = i < n ? order_bits<K, MODE>(keys[i]) : (U)0;
        }
    }
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const uint64_t valid = ballot(ok[j]);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const uint32_t d = (uint32_t)(w[j] >> (8 * p)) & 0xffu;
        const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
        if (ballot(ok[j] && d != d0) == 0) {
          if (__lane_id() == 0 && valid) atomicAdd(&s_h[p * kRadixBins + d0], (uint32_t)__builtin_popcountll(valid));
        } else if (ok[j]) {
          atomicAdd(&s_h[p * kRadixBins + d], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < (uint32_t)(P * kRadixBins); i += BLOCK) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) c += s_hw[w][i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// 6. Onesweep plan (one block of 256 threads): the exclusive scan of every
//    digit position's histogram (the global start of each digit in that
//    pass), which positions need a pass (a position whose histogram puts all
//    n keys in one bin is skipped), the pass table (as the reduce-then-scan
//    plan) and zeroed tile counters.
//    vofs = 0: the intermediate buffers hold interleaved elements; else each
//    holds the keys and, vofs bytes in, the values (the two-phase passes).
template <int P>
__global__ __launch_bounds__(kRadixBins) void radix_onesweep_plan_kernel(
    const uint32_t* __restrict__ hist, uint32_t n, RadixPass* __restrict__ table,
    uint32_t* __restrict__ gbase, uint32_t* __restrict__ tile_counters, const void* in_k,
    const void* in_v, void* out_k, void* out_v, void* aos0, void* aos1, size_t vofs) {
  __shared__ uint32_t s_trivial[P];
  __shared__ uint32_t s_wsum[kRadixBins / kWave];
  const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
  if (t < (uint32_t)P) {
    s_trivial[t] = 0;
    tile_counters[t] = 0;
  }
  __syncthreads();
  for (int p = 0; p < P; ++p) {
    const uint32_t c = hist[p * kRadixBins + t];
    if (c == n) s_trivial[p] = 1;
    const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(c);
    if (lane == 63) s_wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t k = 0; k < wid; ++k) woff += s_wsum[k];
    gbase[p * kRadixBins + t] = woff + incl - c;
    __syncthreads();
  }
  if (t == 0) {
    int active = 0;
    for (int p = 0; p < P; ++p) active += s_trivial[p] ? 0 : 1;
    // All digits constant (all keys equal): one pass still moves the data.
    const bool force_last = active == 0;
    if (force_last) active = 1;
    int j = 0;
    for (int p = 0; p < P; ++p) {
      RadixPass rp;
      rp.shift = 8u * (uint32_t)p;
      rp.active = (!s_trivial[p] || (force_last && p == P - 1)) ? 1u : 0u;
      rp.in_aos = rp.out_aos = 0;
      rp.src_k = rp.src_v = nullptr;
      rp.dst_k = rp.dst_v = nullptr;
      if (rp.active) {
        char* src = (char*)(((j - 1) & 1) ? aos1 : aos0);
        char* dst = (char*)((j & 1) ? aos1 : aos0);
        rp.in_aos = j > 0 && vofs == 0;
        rp.out_aos = j < active - 1 && vofs == 0;
        rp.src_k = j == 0 ? in_k : src;
        rp.src_v = j == 0 ? in_v : (vofs ? src + vofs : nullptr);
        rp.dst_k = j == active - 1 ? out_k : dst;
        rp.dst_v = j == active - 1 ? out_v : (vofs ? dst + vofs : nullptr);
        ++j;
      }
      table[p] = rp;
    }
  }
}

// Whether a onesweep pass over EW-word elements reorders keys and values in
// two phases (and keeps them in separate arrays between passes).
constexpr bool radix_split(int EW) { return EW >= 3; }

// 7. Onesweep pass: a block claims the next tile id from the pass's counter
//    (so every predecessor of a tile is already running), ranks its keys as
//    the downsweep does, publishes its per-digit counts as an aggregate, and
//    one thread per digit walks back over the predecessors' statuses (adding
//    aggregates until it meets an inclusive prefix), publishes its own
//    inclusive prefix and places the digit's run at
//    global start + prefix. A tile publishes its aggregate before it waits on
//    anything, so every walk ends.
//    Elements of 12 or 16 bytes are reordered through LDS in two phases, the
//    keys and then the values (the same LDS), so LDS holds one of them per
//    tile instead of the whole element: 6 waves per SIMD instead of 4 for
//    u64 keys with u32 values (11.33 -> 10.52 ms at 2^27). Narrower elements
//    gain no waves that way (registers bound them) and keep one phase.
template <typename K, int MODE, int EW, int BLOCK, int IPT = kRadixIpt>
__global__ __launch_bounds__(BLOCK) void radix_onesweep_kernel(
    const RadixPass* __restrict__ table, int pass, const uint32_t* __restrict__ gbase,
    uint32_t* __restrict__ tile_counters, uint64_t* __restrict__ status, uint32_t gen,
    uint32_t n) {
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  constexpr int KW = (int)sizeof(K) / 4, VW = EW - KW;
  using V = typename std::conditional<VW == 2, uint64_t, uint32_t>::type;
  constexpr bool kSplit = radix_split(EW);
  constexpr int MW = kSplit ? (KW > VW ? KW : VW) : EW;  // LDS words per element
  __shared__ __align__(16) uint32_t s_part[TILE * MW];
  __shared__ uint32_t s_hist[NW][kRadixBins];     // per-wave counts, then wave offsets
  __shared__ uint32_t s_tile_start[kRadixBins];   // in-tile start of each digit
  __shared__ uint32_t s_goff[kRadixBins];         // global position - in-tile position
  __shared__ uint32_t s_wsum[kRadixBins / kWave];
  __shared__ uint32_t s_cnt[kRadixBins];          // the tile's digit counts (early)
  __shared__ uint32_t s_tile;

  const RadixPass rp = table[pass];
  if (!rp.active) return;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const bool digit_thread = tid < (uint32_t)kRadixBins;
  for (uint32_t i = tid; i < NW * kRadixBins; i += BLOCK) (&s_hist[0][0])[i] = 0;
  if (digit_thread) s_cnt[tid] = 0;
  if (tid == 0) s_tile = atomicAdd(tile_counters + pass, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t base = (size_t)tile * TILE + (size_t)wid * (kWave * IPT);
  const uint32_t shift = rp.shift;

  Elem<EW> el[IPT];
  uint32_t dig[IPT];
  uint32_t rank[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const size_t i = base + (size_t)j * kWave + lane;
    if (i < n) {
      el[j] = load_elem<K, EW>(rp, i);
      dig[j] = (uint32_t)(order_bits<K, MODE>(elem_key<K, EW>(el[j])) >> shift) & 0xffu;
    } else {
      dig[j] = 0x100u;
    }
  }
  uint32_t gstart = 0;
  if (digit_thread) gstart = gbase[pass * kRadixBins + tid];

  // The tile's digit counts first (LDS atomics), published at once as its
  // aggregate: successors walking back meet an aggregate instead of waiting
  // for this tile's ranking.
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (dig[j] < 0x100u) atomicAdd(&s_cnt[dig[j]], 1u);
  __syncthreads();
  uint64_t* my_status = status + (size_t)tile * kRadixBins + tid;
  uint32_t cnt = 0;
  if (digit_thread) {
    cnt = s_cnt[tid];
    st_agent(my_status, pack_status(gen, tile == 0 ? kFlagInc : kFlagAgg, 0, cnt));
  }

  // Wave-level multi-split, as in the downsweep.
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t d = dig[j];
    const bool valid = d < 0x100u;
    const uint64_t peers = match_digit<8>(d, valid);
    uint32_t before = 0;
    if (valid) before = s_hist[wid][d];
    const uint32_t below = (uint32_t)__builtin_popcountll(peers & lt);
    if (valid && below == 0) s_hist[wid][d] = before + (uint32_t)__builtin_popcountll(peers);
    rank[j] = before + below;
  }
  __syncthreads();

  // Per digit (thread t <-> digit t): wave offsets (their sum is cnt).
  if (digit_thread) {
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t c = s_hist[w][tid];
      s_hist[w][tid] = acc;
      acc += c;
    }
  }
  uint32_t incl = 0;
  if (digit_thread) {
    incl = wave_inclusive_scan<uint32_t, OP_ADD>(cnt);
    if (lane == 63) s_wsum[wid] = incl;
  }
  __syncthreads();
  if (digit_thread) {
    uint32_t woff = 0;
    for (uint32_t k = 0; k < wid; ++k) woff += s_wsum[k];
    const uint32_t tstart = woff + incl - cnt;
    s_tile_start[tid] = tstart;
    // Look back: the number of keys with this digit in all earlier tiles,
    // reading kLookbackBatch predecessors per step (loads in flight
    // together); an unpublished predecessor is polled again from there.
    uint32_t prefix = 0;
    if (tile > 0) {
      constexpr int B = kLookbackBatch;
      int64_t q = (int64_t)tile - 1;  // nearest predecessor not yet added
      for (;;) {
        uint64_t w[B];
#pragma unroll
        for (int k = 0; k < B; ++k)
          w[k] = q - k >= 0 ? ld_agent(status + (size_t)(q - k) * kRadixBins + tid) : 0ull;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int k = 0; k < B; ++k) {
          if (done || used < k) continue;
          const uint32_t f = status_flag(w[k], gen);
          if (f == 0) continue;  // not published yet (tile 0 always is)
          prefix += (uint32_t)w[k];
          used = k + 1;
          done = f == kFlagInc;
        }
        if (done) break;
        q -= used;
        if (used == 0) __builtin_amdgcn_s_sleep(1);
      }
      st_agent(my_status, pack_status(gen, kFlagInc, 0, prefix + cnt));
    }
    s_goff[tid] = gstart + prefix - tstart;
  }
  __syncthreads();

  // Reorder into digit-sorted runs in LDS, then write the runs out (split:
  // the keys first, each output position's digit kept for the values, then
  // the values).
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (dig[j] < 0x100u) rank[j] += s_tile_start[dig[j]] + s_hist[wid][dig[j]];
  const uint32_t tile_n = (uint32_t)min((size_t)TILE, (size_t)n - (size_t)tile * TILE);
  if constexpr (!kSplit) {
    Elem<EW>* s_el = reinterpret_cast<Elem<EW>*>(s_part);
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (dig[j] < 0x100u) s_el[rank[j]] = el[j];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t i = tid + (uint32_t)q * BLOCK;
      if (i < tile_n) {
        const Elem<EW> e = s_el[i];
        const uint32_t d = (uint32_t)(order_bits<K, MODE>(elem_key<K, EW>(e)) >> shift) & 0xffu;
        store_elem<K, EW>(rp, (size_t)s_goff[d] + i, e);
      }
    }
    return;
  }
  K* s_key = reinterpret_cast<K*>(s_part);
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (dig[j] < 0x100u) s_key[rank[j]] = elem_key<K, EW>(el[j]);
  __syncthreads();
  uint32_t dq[IPT];
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const uint32_t i = tid + (uint32_t)q * BLOCK;
    dq[q] = 0;
    if (i < tile_n) {
      const K k = s_key[i];
      dq[q] = (uint32_t)(order_bits<K, MODE>(k) >> shift) & 0xffu;
      store_key_part<K, EW>(rp, (size_t)s_goff[dq[q]] + i, k);
    }
  }
  if constexpr (VW > 0) {
    __syncthreads();  // every key read out before the values overwrite them
    V* s_val = reinterpret_cast<V*>(s_part);
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (dig[j] < 0x100u) {
        V v;
        if constexpr (VW == 1) v = el[j].w[KW];
        else v = ((uint64_t)el[j].w[KW + 1] << 32) | el[j].w[KW];
        s_val[rank[j]] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t i = tid + (uint32_t)q * BLOCK;
      if (i < tile_n) store_val_part<K, EW, V>(rp, (size_t)s_goff[dq[q]] + i, s_val[i]);
    }
  }
}

}  // namespace msv
