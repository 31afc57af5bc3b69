// Short segments of seg::Sort (up to 1024 elements) sorted inside one wave,
// in registers, by a bitonic network over (order word, position) pairs.
//
// A segment of length L belongs to size class c = ceil(log2 L) and takes
// 2^c slots (padding slots sort last). Classes 0..6 (L <= 64): a wave sorts
// 64 >> c segments at once, one slot per lane, each segment in its own group
// of 2^c lanes (every compare-exchange partner lies inside the group).
// Classes 7..10 (L <= 1024): a wave sorts one segment with 2^(c-6) slots per
// lane, striped (slot e = r * 64 + lane), so loads and stores are coalesced:
// partners at distance < 64 are other lanes (one shuffle), partners at
// distance >= 64 are other registers of the same lane. Ties are broken by the
// input position, so the order is stable. Keys and values move as they were
// read (so -0.0 / +0.0 keep their bits): after the network, slot e takes the
// key and value of input position i(e), from the lane that read it (one slot
// per lane) or from the wave's LDS staging area.
//
// Work: a classify kernel counts the segments of every class (a histogram
// per block, one global add per class), a fill kernel lists them per class
// as (start, length), the block path's segments (1024 < L <= long limit) and
// the long ones by id; the multi-register classes and classes 6 and 5 run
// one launch per class, classes 0..4 in one, longest class first (with
// 4-byte keys, class 10 by pairs of waves). Wave w takes work
// units w, w + W, w + 2W, ... (a unit is 64 >> c segments of a one-slot
// class or one segment of a larger class) and loads the next unit's
// descriptors before it sorts the current one.
#pragma once

#include "tile_sort.hpp"

namespace msv {

constexpr uint32_t kWaveSortMaxLen = 1024;
constexpr uint32_t kWaveSortClasses = 11;  // 0..10
constexpr uint32_t kSegClassBlock = 11;    // 1024 < L <= long limit: one block per segment
constexpr uint32_t kSegClassLong = 12;     // L > long limit: many blocks per segment
constexpr uint32_t kSegClasses = 13;
constexpr int kWaveSortBlock = 256;
constexpr int kWaveSortWaves = kWaveSortBlock / kWave;
constexpr int kClassifySpt = 8;  // segments per thread in classify / fill

__host__ __device__ inline uint32_t seg_class(uint32_t L, uint32_t long_min) {
  if (L <= 1) return 0;
  if (L >= long_min) return kSegClassLong;
  const uint32_t lg = 32u - (uint32_t)__builtin_clz(L - 1);  // ceil(log2 L)
  return lg <= 10 ? lg : kSegClassBlock;
}
__device__ inline uint32_t wave_sort_class(uint32_t L) {
  return L <= 1 ? 0u : 32u - (uint32_t)__builtin_clz(L - 1);
}
// Segments per work unit of class c.
__host__ __device__ inline uint32_t wave_sort_unit(uint32_t c) { return c <= 6 ? (64u >> c) : 1u; }

// Sort slot: the key's order word and the element's position in its segment;
// for 4-byte keys one 64-bit word (order word << 32 | position), so that a
// comparison is one 64-bit compare. Padding slots are all ones.
template <int KW>
struct SortSlot;
template <>
struct SortSlot<1> {
  uint64_t k;
};
template <>
struct SortSlot<2> {
  uint64_t w;
  uint32_t i;
};

template <int KW>
__device__ __forceinline__ SortSlot<KW> slot_make(KeyWord<KW> w, uint32_t i) {
  SortSlot<KW> s;
  if constexpr (KW == 1) {
    s.k = ((uint64_t)w << 32) | i;
  } else {
    s.w = w;
    s.i = i;
  }
  return s;
}

template <int KW>
__device__ __forceinline__ SortSlot<KW> slot_pad() {
  return slot_make<KW>((KeyWord<KW>)~(KeyWord<KW>)0, 0xffffffffu);
}

template <int KW>
__device__ __forceinline__ uint32_t slot_pos(const SortSlot<KW>& s) {
  if constexpr (KW == 1) return (uint32_t)s.k;
  else return s.i;
}

template <int KW>
__device__ __forceinline__ bool slot_less(const SortSlot<KW>& a, const SortSlot<KW>& b) {
  if constexpr (KW == 1) return a.k < b.k;
  else return a.w < b.w || (a.w == b.w && a.i < b.i);
}

// Value of lane (this lane ^ J), J a power of two below 64: inside a row of
// 16 lanes by DPP (xor 1 and 2 by quad permutes, xor 4 by a half-row mirror
// (lane ^ 7) then a quad reversal (lane ^ 3), xor 8 by a row rotation of 8),
// across rows by the LDS crossbar (swizzle for 16, a lane permute for 32).
// DPP moves are VALU instructions without the LDS round trip a permute
// waits for.
template <uint32_t J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
  if constexpr (J == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
  } else if constexpr (J == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
  } else if constexpr (J == 4) {
    return (uint32_t)__builtin_amdgcn_mov_dpp(
        __builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false);
  } else if constexpr (J == 8) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, false);
  } else if constexpr (J == 16) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x401F);
  } else {
    static_assert(J == 32, "lane distance");
    return (uint32_t)__shfl_xor((int)x, 32);
  }
}

template <int KW, uint32_t J>
__device__ __forceinline__ SortSlot<KW> slot_xor(const SortSlot<KW>& v) {
  SortSlot<KW> o;
  if constexpr (KW == 1) {
    const uint32_t lo = lane_xor<J>((uint32_t)v.k);
    const uint32_t hi = lane_xor<J>((uint32_t)(v.k >> 32));
    o.k = ((uint64_t)hi << 32) | lo;
  } else {
    const uint32_t lo = lane_xor<J>((uint32_t)v.w);
    const uint32_t hi = lane_xor<J>((uint32_t)(v.w >> 32));
    o.w = ((uint64_t)hi << 32) | lo;
    o.i = lane_xor<J>(v.i);
  }
  return o;
}

// Keeps the smaller of (v, partner) when keep_min, else the larger.
template <int KW>
__device__ __forceinline__ void cmp_keep(SortSlot<KW>& v, const SortSlot<KW>& p, bool keep_min) {
  const bool p_less = slot_less<KW>(p, v);
  if (p_less == keep_min) v = p;
}

// Bitonic networks, unrolled at compile time by template recursion: merge
// size K, compare distance J. Slot e = r * 64 + lane; the partner of slot e
// is e ^ J, the smaller goes to the lower slot where e & K == 0 (ascending
// runs) and to the higher slot otherwise. Within a lane group of 2^LG lanes
// the slot index is lane & (2^LG - 1), so every group sorts ascending.
template <int KW, uint32_t K, uint32_t J, int R>
__device__ __forceinline__ void bitonic_stage(SortSlot<KW> (&v)[R], uint32_t e0) {
  if constexpr (J >= 64) {
    constexpr uint32_t jr = J / 64;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if ((r & jr) != 0) continue;
      const bool asc = (((uint32_t)r * 64 + e0) & K) == 0;
      SortSlot<KW>& x = v[r];
      SortSlot<KW>& y = v[r ^ jr];
      const bool swap = asc ? slot_less<KW>(y, x) : slot_less<KW>(x, y);
      if (swap) {
        const SortSlot<KW> t = x;
        x = y;
        y = t;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const SortSlot<KW> p = slot_xor<KW, J>(v[r]);
      const uint32_t e = (uint32_t)r * 64 + e0;
      const bool asc = (e & K) == 0;
      const bool lower = (e & J) == 0;
      cmp_keep<KW>(v[r], p, lower == asc);
    }
  }
}

template <int KW, uint32_t K, uint32_t J, int R>
__device__ __forceinline__ void bitonic_merge(SortSlot<KW> (&v)[R], uint32_t e0) {
  bitonic_stage<KW, K, J, R>(v, e0);
  if constexpr (J > 1) bitonic_merge<KW, K, J / 2, R>(v, e0);
}

template <int KW, uint32_t K, uint32_t N, int R>
__device__ __forceinline__ void bitonic_sort(SortSlot<KW> (&v)[R], uint32_t e0) {
  if constexpr (K <= N) {
    bitonic_merge<KW, K, K / 2, R>(v, e0);
    bitonic_sort<KW, K * 2, N, R>(v, e0);
  }
}

// Lane groups of 2^LG lanes, one slot per lane.
template <int KW, int LG>
__device__ __forceinline__ void bitonic_lanes(SortSlot<KW>& v) {
  SortSlot<KW>(&a)[1] = *reinterpret_cast<SortSlot<KW>(*)[1]>(&v);
  uint32_t e0 = __lane_id();
  asm volatile("" : "+v"(e0));  // per call (see bitonic_wave)
  bitonic_sort<KW, 2, (1u << LG), 1>(a, e0 & ((1u << LG) - 1u));
}

// One wave's 64 * R slots.
template <int KW, int R>
__device__ __forceinline__ void bitonic_wave(SortSlot<KW> (&v)[R]) {
  uint32_t e0 = __lane_id();
  // Per call: hoisted out of the unit loop, the direction masks of all stages
  // would not fit the scalar registers (they spill to VGPR lanes).
  asm volatile("" : "+v"(e0));
  bitonic_sort<KW, 2, 64u * R, R>(v, e0);
}

struct WaveSortArgs {
  const void* keys;
  const void* vals;
  void* out_k;
  void* out_v;
  const uint2* list;       // (start, length) per segment, class by class
  const uint32_t* counts;  // [0..10] segments per class
  OrderMap om;
};

template <int KW>
__device__ __forceinline__ KeyWord<KW> load_raw_key(const void* keys, size_t i) {
  if constexpr (KW == 1) return ((const uint32_t*)keys)[i];
  else return ((const uint64_t*)keys)[i];
}

template <int VW>
struct ValWord {
  using T = typename std::conditional<VW == 2, uint64_t, uint32_t>::type;
};

template <int VW>
__device__ __forceinline__ typename ValWord<VW>::T load_val(const void* vals, size_t i) {
  if constexpr (VW == 2) return ((const uint64_t*)vals)[i];
  else if constexpr (VW == 1) return ((const uint32_t*)vals)[i];
  else return 0u;
}

template <typename T>
__device__ __forceinline__ T shfl_any(T x, int src) {
  if constexpr (sizeof(T) == 4) {
    return (T)__shfl((int)x, src);
  } else {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)x, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)((uint64_t)x >> 32), src);
    return (T)(((uint64_t)hi << 32) | lo);
  }
}

// Counts the non-empty segments of every class: a histogram per block in
// LDS, then one global add per class.
__global__ __launch_bounds__(256) void seg_classify_kernel(const uint32_t* __restrict__ offsets,
                                                           uint32_t nseg, uint32_t long_min,
                                                           uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_c[kSegClasses];
  const uint32_t tid = threadIdx.x;
  if (tid < kSegClasses) s_c[tid] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * 256u * kClassifySpt;
#pragma unroll
  for (int q = 0; q < kClassifySpt; ++q) {
    const uint32_t s = base + (uint32_t)q * 256u + tid;
    if (s < nseg) {
      const uint32_t L = offsets[s + 1] - offsets[s];
      if (L > 0) atomicAdd(&s_c[seg_class(L, long_min)], 1u);
    }
  }
  __syncthreads();
  if (tid < kSegClasses && s_c[tid]) atomicAdd(&counts[tid], s_c[tid]);
}

// Lists every non-empty segment under its class: the wave classes as
// (start, length), back to back in class order; the block path's segments and
// the long ones by id (cursors[11], and long_count). Order inside a class is
// the order in which blocks reserve their ranges.
__global__ __launch_bounds__(256) void seg_fill_kernel(const uint32_t* __restrict__ offsets,
                                                       uint32_t nseg, uint32_t long_min,
                                                       const uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ cursors,
                                                       uint2* __restrict__ list,
                                                       uint32_t* __restrict__ big,
                                                       uint32_t* __restrict__ longs,
                                                       uint32_t* __restrict__ long_count) {
  __shared__ uint32_t s_c[kSegClasses], s_base[kSegClasses], s_start[kWaveSortClasses];
  const uint32_t tid = threadIdx.x;
  if (tid < kSegClasses) s_c[tid] = 0;
  if (tid < kWaveSortClasses) {
    uint32_t st = 0;
    for (uint32_t k = 0; k < tid; ++k) st += counts[k];
    s_start[tid] = st;
  }
  __syncthreads();
  const uint32_t base = blockIdx.x * 256u * kClassifySpt;
  uint32_t cls[kClassifySpt], rk[kClassifySpt], g0[kClassifySpt], len[kClassifySpt];
#pragma unroll
  for (int q = 0; q < kClassifySpt; ++q) {
    const uint32_t s = base + (uint32_t)q * 256u + tid;
    cls[q] = 0xffu;
    if (s < nseg) {
      g0[q] = offsets[s];
      len[q] = offsets[s + 1] - g0[q];
      if (len[q] > 0) {
        cls[q] = seg_class(len[q], long_min);
        rk[q] = atomicAdd(&s_c[cls[q]], 1u);
      }
    }
  }
  __syncthreads();
  if (tid < kSegClasses && s_c[tid])
    s_base[tid] = atomicAdd(tid == kSegClassLong ? long_count : &cursors[tid], s_c[tid]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kClassifySpt; ++q) {
    const uint32_t c = cls[q];
    if (c == 0xffu) continue;
    const uint32_t s = base + (uint32_t)q * 256u + tid;
    const uint32_t pos = s_base[c] + rk[q];
    if (c < kWaveSortClasses) list[s_start[c] + pos] = make_uint2(g0[q], len[q]);
    else if (c == kSegClassBlock) big[pos] = s;
    else longs[pos] = s;
  }
}

// A lane's input element: key and value (position lane & (2^c - 1) of its
// group's segment, or nothing past the segment's end).
template <int KW, int VW>
struct LaneElem {
  KeyWord<KW> k;
  typename ValWord<VW>::T v;
};

template <int KW, int VW>
__device__ __forceinline__ LaneElem<KW, VW> load_lane_elem(const WaveSortArgs& a, uint32_t c,
                                                          uint2 d) {
  const uint32_t e = __lane_id() & ((1u << c) - 1u);
  LaneElem<KW, VW> x{0, 0};
  if (e < d.y) {
    x.k = load_raw_key<KW>(a.keys, (size_t)d.x + e);
    x.v = load_val<VW>(a.vals, (size_t)d.x + e);
  }
  return x;
}

// Classes 0..6: a unit's segments, one per group of 2^c lanes. Every lane
// keeps the key and value it read; slot e then takes them from the lane that
// read input position i(e).
template <int KW, int VW, int LG>
__device__ __forceinline__ void wave_sort_groups(const WaveSortArgs& a, uint2 d,
                                                 const LaneElem<KW, VW>& x) {
  constexpr uint32_t c = LG;
  const uint32_t lane = __lane_id();
  const uint32_t e = lane & ((1u << c) - 1u);
  const size_t g0 = d.x;
  const uint32_t L = d.y;
  SortSlot<KW> v = e < L ? slot_make<KW>(order_word<KeyWord<KW>>(x.k, a.om), e) : slot_pad<KW>();
  bitonic_lanes<KW, LG>(v);
  const int src = (int)((lane & ~((1u << c) - 1u)) + (e < L ? slot_pos<KW>(v) : e));
  const KeyWord<KW> ok = shfl_any(x.k, src);
  if (e < L) ((KeyWord<KW>*)a.out_k)[g0 + e] = ok;
  if constexpr (VW > 0) {
    const typename ValWord<VW>::T ov = shfl_any(x.v, src);
    if (e < L) ((typename ValWord<VW>::T*)a.out_v)[g0 + e] = ov;
  }
}

// LDS staging of one wave: the segment's keys and values by input position.
template <int KW, int VW, int N>
struct WaveStage {
  KeyWord<KW> k[N];
  typename ValWord<VW>::T v[VW > 0 ? N : 1];
};

// A wave's input of one segment of class 6 + log2(R): slot r of a lane is
// position r * 64 + lane.
template <int KW, int VW, int R>
struct WaveElems {
  KeyWord<KW> k[R];
  typename ValWord<VW>::T v[VW > 0 ? R : 1];
};

template <int KW, int VW, int R>
__device__ __forceinline__ void load_wave_elems(const WaveSortArgs& a, uint2 d,
                                                WaveElems<KW, VW, R>& x) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    if (e < d.y) {
      x.k[r] = load_raw_key<KW>(a.keys, (size_t)d.x + e);
      if constexpr (VW > 0) x.v[r] = load_val<VW>(a.vals, (size_t)d.x + e);
    }
  }
}

// Classes 7..10: one segment, R slots per lane, from its input elements.
template <int KW, int VW, int R>
__device__ __forceinline__ void wave_sort_one(const WaveSortArgs& a, WaveStage<KW, VW, 64 * R>& st,
                                              uint2 d, const WaveElems<KW, VW, R>& x) {
  const uint32_t lane = __lane_id();
  const size_t g0 = d.x;
  const uint32_t L = d.y;
  SortSlot<KW> v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    v[r] = slot_pad<KW>();
    if (e < L) {
      st.k[e] = x.k[r];
      if constexpr (VW > 0) st.v[e] = x.v[r];
      v[r] = slot_make<KW>(order_word<KeyWord<KW>>(x.k[r], a.om), e);
    }
  }
  bitonic_wave<KW, R>(v);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    if (e < L) {
      const uint32_t p = slot_pos<KW>(v[r]);
      ((KeyWord<KW>*)a.out_k)[g0 + e] = st.k[p];
      if constexpr (VW > 0) ((typename ValWord<VW>::T*)a.out_v)[g0 + e] = st.v[p];
    }
  }
}

// The same, loading the segment at the unit: values straight to the staging
// area, keys into the slots (for wide units, where holding the next unit's
// input would cost the registers the network needs).
template <int KW, int VW, int R>
__device__ __forceinline__ void wave_sort_one_direct(const WaveSortArgs& a,
                                                     WaveStage<KW, VW, 64 * R>& st, uint2 d) {
  const uint32_t lane = __lane_id();
  const size_t g0 = d.x;
  const uint32_t L = d.y;
  if constexpr (VW > 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t e = (uint32_t)r * 64 + lane;
      if (e < L) st.v[e] = load_val<VW>(a.vals, g0 + e);
    }
  }
  SortSlot<KW> v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    v[r] = slot_pad<KW>();
    if (e < L) {
      const KeyWord<KW> raw = load_raw_key<KW>(a.keys, g0 + e);
      st.k[e] = raw;
      v[r] = slot_make<KW>(order_word<KeyWord<KW>>(raw, a.om), e);
    }
  }
  bitonic_wave<KW, R>(v);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    if (e < L) {
      const uint32_t p = slot_pos<KW>(v[r]);
      ((KeyWord<KW>*)a.out_k)[g0 + e] = st.k[p];
      if constexpr (VW > 0) ((typename ValWord<VW>::T*)a.out_v)[g0 + e] = st.v[p];
    }
  }
}

// Classes CLO..CHI (within 0..6) in one launch, class CHI first. A wave's
// units are pipelined: descriptors two units ahead, input elements one unit
// ahead, so the memory latency of the next unit overlaps the current unit's
// network.
template <int KW, int VW, int CLO, int CHI>
__global__ __launch_bounds__(kWaveSortBlock) void wave_sort_small_kernel(WaveSortArgs a) {
  __shared__ uint32_t units[7], starts[7], s_total;
  if (threadIdx.x == 0) {
    uint32_t total = 0, seg_start = 0;
    for (uint32_t c = 0; c < 7; ++c) {
      const uint32_t n = a.counts[c];
      units[c] = (n + wave_sort_unit(c) - 1) / wave_sort_unit(c);
      starts[c] = seg_start;
      seg_start += n;
      if (c >= (uint32_t)CLO && c <= (uint32_t)CHI) total += units[c];
    }
    s_total = total;
  }
  __syncthreads();
  const uint32_t total = s_total;
  const uint32_t lane = __lane_id();
  const uint32_t W = gridDim.x * kWaveSortWaves;
  // Unit u counts from class CHI down: its class, and this lane's descriptor
  // (length 0 past the unit's segments and past the last unit).
  auto locate = [&](uint32_t u, uint32_t* cls) -> uint2 {
    if (u >= total) {
      *cls = CLO;
      return make_uint2(0u, 0u);
    }
    uint32_t c = CHI;
    if constexpr (CLO < CHI) {
      while (u >= units[c]) {
        u -= units[c];
        --c;
      }
    }
    *cls = c;
    const uint32_t per = wave_sort_unit(c);
    const uint32_t g = lane >> c;
    const uint32_t k = u * per + g;
    return g < per && k < a.counts[c] ? a.list[starts[c] + k] : make_uint2(0u, 0u);
  };
  uint32_t u0 = blockIdx.x * kWaveSortWaves + threadIdx.x / kWave;
  uint32_t c0, c1;
  uint2 d0 = locate(u0, &c0);
  uint2 d1 = locate(u0 + W, &c1);
  LaneElem<KW, VW> x0 = load_lane_elem<KW, VW>(a, c0, d0);
  while (u0 < total) {
    const LaneElem<KW, VW> x1 = load_lane_elem<KW, VW>(a, c1, d1);
    uint32_t c2;
    const uint2 d2 = locate(u0 + 2 * W, &c2);
    if constexpr (CLO == CHI) {
      wave_sort_groups<KW, VW, CHI>(a, d0, x0);
    } else {
      switch (c0) {
        case 0: wave_sort_groups<KW, VW, 0>(a, d0, x0); break;
        case 1: wave_sort_groups<KW, VW, 1>(a, d0, x0); break;
        case 2: wave_sort_groups<KW, VW, 2>(a, d0, x0); break;
        case 3: wave_sort_groups<KW, VW, 3>(a, d0, x0); break;
        case 4: wave_sort_groups<KW, VW, 4>(a, d0, x0); break;
        case 5: wave_sort_groups<KW, VW, 5>(a, d0, x0); break;
        default: wave_sort_groups<KW, VW, 6>(a, d0, x0); break;
      }
    }
    u0 += W;
    c0 = c1;
    d0 = d1;
    x0 = x1;
    c1 = c2;
    d1 = d2;
  }
}

// Class 6 + log2(R) (R = 2, 4, 8, 16), one launch per class, pipelined as
// the one-slot classes (descriptors two units ahead; input elements one unit
// ahead where they take at most 16 registers, else loaded at the unit).
// The waves-per-SIMD floor keeps the compiler's register budget near what one
// segment's slots need (left alone it spends 191 VGPRs at R = 16: 2 waves).
template <int KW, int VW, int R>
__global__ __launch_bounds__(kWaveSortBlock) __attribute__((amdgpu_waves_per_eu(R >= 16 ? 3 : R >= 8 ? 5 : 7)))
void wave_sort_large_kernel(WaveSortArgs a) {
  constexpr uint32_t C = 6 + (R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4);
  constexpr bool kAhead = R * (KW + VW) <= 16;
  __shared__ WaveStage<KW, VW, 64 * R> stage[kWaveSortWaves];
  __shared__ uint32_t s_first;
  if (threadIdx.x == 0) {
    uint32_t first = 0;
    for (uint32_t c = 0; c < C; ++c) first += a.counts[c];
    s_first = first;
  }
  __syncthreads();
  const uint32_t first = s_first, n = a.counts[C];
  // The wave index, descriptors and addresses are wave-uniform (scalar).
  const uint32_t wid = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  WaveStage<KW, VW, 64 * R>& st = stage[wid];
  const uint32_t W = gridDim.x * kWaveSortWaves;
  auto desc = [&](uint32_t k) -> uint2 {
    return k < n ? a.list[first + k] : make_uint2(0u, 0u);
  };
  auto uniform = [](uint2 d) -> uint2 {
    return make_uint2((uint32_t)__builtin_amdgcn_readfirstlane((int)d.x),
                      (uint32_t)__builtin_amdgcn_readfirstlane((int)d.y));
  };
  uint32_t k = blockIdx.x * kWaveSortWaves + wid;
  uint2 d0 = uniform(desc(k));
  uint2 d1 = desc(k + W);
  if constexpr (kAhead) {
    WaveElems<KW, VW, R> x0;
    load_wave_elems<KW, VW, R>(a, d0, x0);
    while (k < n) {
      const uint2 d1u = uniform(d1);
      WaveElems<KW, VW, R> x1;
      load_wave_elems<KW, VW, R>(a, d1u, x1);
      const uint2 d2 = desc(k + 2 * W);
      wave_sort_one<KW, VW, R>(a, st, d0, x0);
      x0 = x1;
      k += W;
      d0 = d1u;
      d1 = d2;
    }
  } else {
    while (k < n) {
      const uint2 d1u = uniform(d1);
      const uint2 d2 = desc(k + 2 * W);
      wave_sort_one_direct<KW, VW, R>(a, st, d0);
      k += W;
      d0 = d1u;
      d1 = d2;
    }
  }
}

// Class 10 (513..1024 elements) by a pair of waves per segment, 8 slots per
// lane each (instead of one wave with 16 slots per lane, which holds 136-168
// VGPRs and runs 3 waves per SIMD): wave h holds slots h * 512 + r * 64 +
// lane. Both halves are sorted in opposite directions in registers (the
// stages up to size 512), the size-1024 merge's first stage (distance 512)
// pairs slot e with slot e ^ 512 of the other wave through LDS, and its
// remaining stages stay in registers. One block is one pair, so its barriers
// never depend on the other waves of a larger block.
constexpr int kPairBlock = 2 * kWave;

template <int KW, int VW>
__global__ __launch_bounds__(kPairBlock) __attribute__((amdgpu_waves_per_eu(5)))
void wave_sort_pair_kernel(WaveSortArgs a) {
  constexpr int R = 8;
  constexpr uint32_t C = 10;
  __shared__ WaveStage<KW, VW, 1024> st;
  __shared__ SortSlot<KW> xch[1024];
  __shared__ uint32_t s_first;
  if (threadIdx.x == 0) {
    uint32_t first = 0;
    for (uint32_t c = 0; c < C; ++c) first += a.counts[c];
    s_first = first;
  }
  __syncthreads();
  const uint32_t first = s_first, n = a.counts[C];
  const uint32_t lane = __lane_id();
  const uint32_t h = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {  // uniform over the block
    const uint2 dd = a.list[first + k];
    const size_t g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)dd.x);
    const uint32_t L = (uint32_t)__builtin_amdgcn_readfirstlane((int)dd.y);
    SortSlot<KW> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t e = h * 512u + (uint32_t)r * 64 + lane;
      v[r] = slot_pad<KW>();
      if (e < L) {
        const KeyWord<KW> raw = load_raw_key<KW>(a.keys, g0 + e);
        st.k[e] = raw;
        if constexpr (VW > 0) st.v[e] = load_val<VW>(a.vals, g0 + e);
        v[r] = slot_make<KW>(order_word<KeyWord<KW>>(raw, a.om), e);
      }
    }
    uint32_t e0 = h * 512u + lane;
    asm volatile("" : "+v"(e0));  // as in bitonic_wave
    bitonic_sort<KW, 2, 512, R>(v, e0);
#pragma unroll
    for (int r = 0; r < R; ++r) xch[h * 512u + (uint32_t)r * 64 + lane] = v[r];
    __syncthreads();  // also publishes the staged input of both halves
#pragma unroll
    for (int r = 0; r < R; ++r)
      cmp_keep<KW>(v[r], xch[(h ^ 1u) * 512u + (uint32_t)r * 64 + lane], h == 0);
    bitonic_merge<KW, 1024, 256, R>(v, e0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t e = h * 512u + (uint32_t)r * 64 + lane;
      if (e < L) {
        const uint32_t p = slot_pos<KW>(v[r]);
        ((KeyWord<KW>*)a.out_k)[g0 + e] = st.k[p];
        if constexpr (VW > 0) ((typename ValWord<VW>::T*)a.out_v)[g0 + e] = st.v[p];
      }
    }
    __syncthreads();  // before the next segment overwrites st and xch
  }
}

}  // namespace msv
