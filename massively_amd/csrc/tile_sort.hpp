// Block-level stable LSD radix sorts (device code), shared by the segmented
// sort and the batched small-job kernel:
//  * tile sort: up to BLOCK * IPT elements held in LDS, sorted by key and then,
//    for a tile that packs several segments, by segment slot, so the result is
//    ordered by (segment, key, input position);
//  * range sort: one block sorts a longer range through global memory, a
//    chunk of BLOCK * IPT elements at a time per digit pass.
// Keys are sorted by their order word (unsigned, sign-flipped, IEEE total
// order or comparator order, optionally inverted); the element keeps the raw
// key so that -0.0 / +0.0 survive the comparator modes.
#pragma once

#include "radix_kernels.hpp"

namespace msv {

// Runtime description of the key order (uniform over a launch).
enum KeyKind : uint32_t { KEY_UNSIGNED = 0, KEY_SIGNED = 1, KEY_FLOAT = 2 };
struct OrderMap {
  uint32_t kind;  // KeyKind
  int mode;       // OrderMode
};

// Branch-free (selects only), so that per-item register arrays around it stay
// in plain registers.
template <typename U>
__device__ inline U order_word(U raw, OrderMap om) {
  constexpr U sign = (U)1 << (sizeof(U) * 8 - 1);
  const bool is_float = om.kind == KEY_FLOAT;
  const bool canon = is_float && om.mode != ORD_TOTAL;  // -0 == +0 under `<`
  const U e = (canon && (raw & (U)~sign) == 0) ? (U)0 : raw;
  using SU = typename std::make_signed<U>::type;
  const U float_flip = (U)((SU)e >> (sizeof(U) * 8 - 1)) | sign;  // negative: all ones
  const U flip = is_float ? float_flip : (om.kind == KEY_SIGNED ? sign : (U)0);
  const U inv = om.mode == ORD_GREATER ? (U)~(U)0 : (U)0;
  return e ^ flip ^ inv;
}

template <int KW>
using KeyWord = typename std::conditional<KW == 1, uint32_t, uint64_t>::type;

template <int KW, int EW>
__device__ inline KeyWord<KW> elem_word(const Elem<EW>& e) {
  if constexpr (KW == 1) return e.w[0];
  else return ((uint64_t)e.w[1] << 32) | e.w[0];
}

// Element i of separate key (KW dwords) and value (EW - KW dwords) arrays.
template <int KW, int EW>
__device__ inline Elem<EW> load_kv(const void* keys, const void* vals, size_t i) {
  Elem<EW> e;
  if constexpr (KW == 1) {
    e.w[0] = ((const uint32_t*)keys)[i];
  } else {
    const uint64_t k = ((const uint64_t*)keys)[i];
    e.w[0] = (uint32_t)k;
    e.w[1] = (uint32_t)(k >> 32);
  }
  constexpr int VW = EW - KW;
  if constexpr (VW == 1) {
    e.w[KW] = ((const uint32_t*)vals)[i];
  } else if constexpr (VW == 2) {
    const uint64_t v = ((const uint64_t*)vals)[i];
    e.w[KW] = (uint32_t)v;
    e.w[KW + 1] = (uint32_t)(v >> 32);
  }
  return e;
}

template <int KW, int EW>
__device__ inline void store_kv(void* keys, void* vals, size_t i, const Elem<EW>& e) {
  if constexpr (KW == 1) ((uint32_t*)keys)[i] = e.w[0];
  else ((uint64_t*)keys)[i] = ((uint64_t)e.w[1] << 32) | e.w[0];
  constexpr int VW = EW - KW;
  if constexpr (VW == 1) ((uint32_t*)vals)[i] = e.w[KW];
  else if constexpr (VW == 2) ((uint64_t*)vals)[i] = ((uint64_t)e.w[KW + 1] << 32) | e.w[KW];
}

// Shared memory of the tile sort. `hist` is double-buffered so that a pass
// clears the next pass's counters while it runs.
template <int EW, int BLOCK, int IPT, int DBITS>
struct alignas(16) TileSmem {
  static constexpr int NW = BLOCK / kWave;
  static constexpr int BINS = 1 << DBITS;
  static constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  Elem<EW> el[TILE];
  uint16_t slot[TILE];
  uint32_t hist[2][NW][BINS];
  uint32_t head[TILE / 32];  // segment-start bitmap
  uint32_t hpre[TILE / 32];  // exclusive popcount prefix of `head`
  uint32_t red[4 * NW];
  uint32_t misc[4];
};

// Digit-major exclusive scan of h[wave][digit] by wave 0: afterwards
// h[w][d] = (elements of digits < d) + (elements of digit d in waves < w).
// Digits are interleaved over lanes (d = k * 64 + lane) so that the LDS
// accesses are conflict-free. Returns nothing; `tot_out` (if non-null) gets
// the per-digit totals and `start_out` the per-digit starts.
template <int NW, int BINS>
__device__ __forceinline__ void digit_scan_wave0(uint32_t (*h)[BINS], uint32_t* start_out, uint32_t* tot_out) {
  const uint32_t lane = threadIdx.x & 63u;
  constexpr int DPL = BINS >= kWave ? BINS / kWave : 1;
  uint32_t carry = 0;
#pragma unroll
  for (int k = 0; k < DPL; ++k) {
    const uint32_t d = (uint32_t)k * kWave + lane;
    uint32_t c = 0;
    if (d < (uint32_t)BINS) {
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const uint32_t x = h[w][d];
        h[w][d] = c;
        c += x;
      }
    }
    const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(c);
    const uint32_t start = carry + incl - c;
    if (d < (uint32_t)BINS) {
#pragma unroll
      for (int w = 0; w < NW; ++w) h[w][d] += start;
      if (start_out) start_out[d] = start;
      if (tot_out) tot_out[d] = c;
    }
    carry += readlane(incl, kWave - 1);
  }
}

// Digit-major exclusive scan of h[wave][digit] by the first BINS threads
// (thread t owns digit t): afterwards h[w][d] = (elements of digits < d) +
// (elements of digit d in waves < w). Spreads the work of digit_scan_wave0
// over BINS / 64 waves at the cost of one more barrier. All threads call it.
template <int NW, int BINS, int BLOCK>
__device__ __forceinline__ void digit_scan_block(uint32_t (*h)[BINS], uint32_t* s_wsum) {
  static_assert(BLOCK >= BINS, "one thread per digit");
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  constexpr uint32_t DW = (BINS + kWave - 1) / kWave;  // waves holding digits
  const bool dthread = tid < (uint32_t)BINS;
  uint32_t c = 0;
  if (dthread) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t x = h[w][tid];
      h[w][tid] = c;
      c += x;
    }
  }
  uint32_t incl = 0;
  if (wid < DW) {
    incl = wave_inclusive_scan<uint32_t, OP_ADD>(c);
    if (lane == 63) s_wsum[wid] = incl;
  }
  __syncthreads();
  if (dthread) {
    uint32_t start = incl - c;
    for (uint32_t k = 0; k < wid; ++k) start += s_wsum[k];
#pragma unroll
    for (int w = 0; w < NW; ++w) h[w][tid] += start;
  }
}

// One stable counting pass over s.el[0..n) in LDS by the digit at `shift` of
// the key's order word (or of the slot when `slot_digit`). Items of wave w are
// [w * 64 * IPT, (w + 1) * 64 * IPT), row-major, so ranking order is input
// order. `buf` selects the counter buffer; the other one is cleared. The
// elements are read only after the ranking, so the digits and ranks are the
// only per-item state live across it (register pressure decides occupancy).
template <int KW, int EW, int BLOCK, int IPT, int DBITS, bool SLOTS>
__device__ __forceinline__ void tile_pass(TileSmem<EW, BLOCK, IPT, DBITS>& s, uint32_t n,
                                          uint32_t shift, bool slot_digit, int buf, OrderMap om) {
  using S = TileSmem<EW, BLOCK, IPT, DBITS>;
  constexpr int NW = S::NW, BINS = S::BINS;
  constexpr uint32_t MASK = (uint32_t)BINS - 1;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const uint32_t wbase = wid * (uint32_t)(kWave * IPT);
  // Per-item arrays are written unconditionally (selects, clamped LDS
  // indices): conditional writes make the compiler copy whole arrays.
  uint32_t dg[IPT], rk[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t p = wbase + (uint32_t)j * kWave + lane;
    const uint32_t q = p < n ? p : 0u;
    uint32_t dv;
    if constexpr (SLOTS) {
      dv = slot_digit ? ((uint32_t)s.slot[q] >> shift) & MASK
                      : (uint32_t)(order_word(elem_word<KW, EW>(s.el[q]), om) >> shift) & MASK;
    } else {
      dv = (uint32_t)(order_word(elem_word<KW, EW>(s.el[q]), om) >> shift) & MASK;
    }
    dg[j] = p < n ? dv : (uint32_t)BINS;
  }
  uint32_t(*h)[BINS] = s.hist[buf];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t d = dg[j];
    const bool valid = d < (uint32_t)BINS;
    uint32_t r = 0;
    // Rows past the tile's end are skipped (wave-uniform); no array is
    // written inside the branch.
    if (wbase + (uint32_t)j * kWave < n) {
      const uint64_t peers = match_digit<DBITS>(d, valid);
      const uint32_t before = h[wid][valid ? d : 0u];
      const uint32_t below = (uint32_t)__builtin_popcountll(peers & lt);
      if (valid && below == 0) h[wid][d] = before + (uint32_t)__builtin_popcountll(peers);
      r = before + below;
    }
    rk[j] = valid ? r : 0u;
  }
  __syncthreads();
  for (uint32_t i = tid; i < (uint32_t)(NW * BINS); i += BLOCK) (&s.hist[buf ^ 1][0][0])[i] = 0;
  digit_scan_block<NW, BINS, BLOCK>(h, s.red);
  __syncthreads();
  // Destination of every item, then the items themselves; all reads of the
  // tile finish before the first write.
  Elem<EW> el[IPT];
  uint16_t sl[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t p = wbase + (uint32_t)j * kWave + lane;
    const uint32_t q = p < n ? p : 0u;
    const bool valid = dg[j] < (uint32_t)BINS;
    rk[j] += h[wid][valid ? dg[j] : 0u];
#pragma unroll
    for (int w = 0; w < EW; ++w) el[j].w[w] = s.el[q].w[w];
    if constexpr (SLOTS) sl[j] = s.slot[q];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if (dg[j] < (uint32_t)BINS) {
#pragma unroll
      for (int w = 0; w < EW; ++w) s.el[rk[j]].w[w] = el[j].w[w];
      if constexpr (SLOTS) s.slot[rk[j]] = sl[j];
    }
  }
  __syncthreads();
}

// Sorts s.el[0..n) by key order word (skipping digit positions on which the
// tile does not vary: `varying` = OR ^ AND of the tile's order words), then by
// slot over `slot_bits` bits. The caller has cleared s.hist[0] and synced.
template <int KW, int EW, int BLOCK, int IPT, int DBITS, bool SLOTS>
__device__ __forceinline__ void tile_sort(TileSmem<EW, BLOCK, IPT, DBITS>& s, uint32_t n, KeyWord<KW> varying,
                          uint32_t slot_bits, OrderMap om) {
  constexpr uint32_t MASK = (1u << DBITS) - 1;
  int buf = 0;
  for (uint32_t shift = 0; shift < 32u * KW; shift += DBITS) {
    if (((varying >> shift) & MASK) == 0) continue;
    tile_pass<KW, EW, BLOCK, IPT, DBITS, SLOTS>(s, n, shift, false, buf, om);
    buf ^= 1;
  }
  if constexpr (SLOTS) {
    for (uint32_t shift = 0; shift < slot_bits; shift += DBITS) {
      tile_pass<KW, EW, BLOCK, IPT, DBITS, SLOTS>(s, n, shift, true, buf, om);
      buf ^= 1;
    }
  }
}

// Block-wide OR and AND of a key word (all threads call; result in all).
template <int KW, int BLOCK, int EW, int IPT, int DBITS>
__device__ __forceinline__ KeyWord<KW> block_varying(TileSmem<EW, BLOCK, IPT, DBITS>& s, KeyWord<KW> o,
                                            KeyWord<KW> a) {
  using U = KeyWord<KW>;
  constexpr int NW = BLOCK / kWave;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    o |= __shfl_xor(o, m);
    a &= __shfl_xor(a, m);
  }
  U* red = (U*)s.red;
  if (lane == 0) {
    red[wid] = o;
    red[NW + wid] = a;
  }
  __syncthreads();
  U ro = 0, ra = (U)~(U)0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    ro |= red[w];
    ra &= red[NW + w];
  }
  __syncthreads();
  return ro ^ ra;
}

// Loads [g0, g0 + n) into the tile, sorts it and stores it back at the same
// positions of the output arrays. With SLOTS, the segments that start at
// offsets[first..end) (all inside the range) are kept in place and each is
// sorted on its own.
template <int KW, int EW, int BLOCK, int IPT, int DBITS, bool SLOTS>
__device__ __forceinline__ void tile_sort_range(TileSmem<EW, BLOCK, IPT, DBITS>& s, const void* keys,
                                const void* vals, void* out_k, void* out_v, uint32_t g0, uint32_t n,
                                const uint32_t* offsets, uint32_t first, uint32_t end, OrderMap om) {
  using S = TileSmem<EW, BLOCK, IPT, DBITS>;
  using U = KeyWord<KW>;
  constexpr int NW = S::NW, BINS = S::BINS;
  constexpr uint32_t HW = S::TILE / 32;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  for (uint32_t i = tid; i < (uint32_t)(NW * BINS); i += BLOCK) (&s.hist[0][0][0])[i] = 0;
  if constexpr (SLOTS)
    for (uint32_t i = tid; i < HW; i += BLOCK) s.head[i] = 0;
  U o = 0, a = (U)~(U)0;
  {
    // All IPT loads in flight before the first LDS store (n <= BLOCK * IPT).
    Elem<EW> ev[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t i = tid + (uint32_t)q * BLOCK;
      if (i < n) ev[q] = load_kv<KW, EW>(keys, vals, (size_t)g0 + i);
    }
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t i = tid + (uint32_t)q * BLOCK;
      if (i < n) {
        s.el[i] = ev[q];
        const U w = order_word(elem_word<KW, EW>(ev[q]), om);
        o |= w;
        a &= w;
      }
    }
  }
  uint32_t slot_bits = 0;
  if constexpr (SLOTS) {
    __syncthreads();
    for (uint32_t sg = first + tid; sg < end; sg += BLOCK) {
      const uint32_t b = offsets[sg];
      if (offsets[sg + 1] > b) {
        const uint32_t p = b - g0;
        atomicOr(&s.head[p >> 5], 1u << (p & 31u));
      }
    }
    __syncthreads();
    if (wid == 0) {
      constexpr uint32_t WPL = (HW + kWave - 1) / kWave;
      uint32_t c[WPL], sum = 0;
#pragma unroll
      for (uint32_t k = 0; k < WPL; ++k) {
        const uint32_t w = lane * WPL + k;
        c[k] = w < HW ? (uint32_t)__builtin_popcount(s.head[w]) : 0u;
        sum += c[k];
      }
      const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(sum);
      uint32_t run = incl - sum;
#pragma unroll
      for (uint32_t k = 0; k < WPL; ++k) {
        const uint32_t w = lane * WPL + k;
        if (w < HW) s.hpre[w] = run;
        run += c[k];
      }
      if (lane == 63) s.misc[0] = incl;
    }
    __syncthreads();
    const uint32_t nheads = s.misc[0];
    for (uint32_t i = tid; i < n; i += BLOCK) {
      const uint32_t w = i >> 5;
      const uint32_t below = (uint32_t)__builtin_popcount(s.head[w] & ((2u << (i & 31u)) - 1u));
      s.slot[i] = (uint16_t)(s.hpre[w] + below - 1u);
    }
    slot_bits = nheads > 1 ? 32u - (uint32_t)__builtin_clz(nheads - 1u) : 0u;
  }
  const U varying = block_varying<KW, BLOCK>(s, o, a);  // syncs
  tile_sort<KW, EW, BLOCK, IPT, DBITS, SLOTS>(s, n, varying, slot_bits, om);
  for (uint32_t i = tid; i < n; i += BLOCK) store_kv<KW, EW>(out_k, out_v, (size_t)g0 + i, s.el[i]);
  __syncthreads();
}

// ---- range sort: one block, a range longer than a tile ----------------------

// Shared memory of the range sort (8-bit digits).
template <int EW, int BLOCK, int IPT>
struct alignas(16) RangeSmem {
  static constexpr int NW = BLOCK / kWave;
  static constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  Elem<EW> el[TILE];
  uint32_t hist[2][NW][kRadixBins];
  uint32_t ghist[8][kRadixBins];  // all-digit histogram of the range
  uint32_t cursor[kRadixBins];    // next output position of each digit
  uint32_t gbase[kRadixBins];     // output position - chunk position, per digit
  uint32_t cstart[kRadixBins];
  uint32_t ctot[kRadixBins];
  uint32_t misc[4];
};

// Source or destination of a range pass: separate key/value arrays or an
// element (AoS) array, indexed from the range start.
struct RangeIo {
  const void* k;
  const void* v;
  uint32_t aos;
};

template <int KW, int EW>
__device__ inline Elem<EW> range_load(const RangeIo& io, size_t i) {
  if (io.aos) return ((const Elem<EW>*)io.k)[i];
  return load_kv<KW, EW>(io.k, io.v, i);
}

template <int KW, int EW>
__device__ inline void range_store(const RangeIo& io, size_t i, const Elem<EW>& e) {
  if (io.aos) ((Elem<EW>*)io.k)[i] = e;
  else store_kv<KW, EW>((void*)io.k, (void*)io.v, i, e);
}

// Stable sort of n elements (n > 0) from `in` into `out` by one block:
// an all-digit histogram, then one chunked counting pass per digit position
// that varies, ping-ponging through the element buffers s0 / s1 (n elements
// each). All pointers are already offset to the range start.
template <int KW, int EW, int BLOCK, int IPT>
__device__ __forceinline__ void range_sort(RangeSmem<EW, BLOCK, IPT>& s, RangeIo in, RangeIo out, Elem<EW>* s0,
                           Elem<EW>* s1, uint32_t n, OrderMap om) {
  using S = RangeSmem<EW, BLOCK, IPT>;
  constexpr int NW = S::NW;
  constexpr uint32_t TILE = S::TILE;
  constexpr int P = 4 * KW;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  for (uint32_t i = tid; i < 8u * kRadixBins; i += BLOCK) (&s.ghist[0][0])[i] = 0;
  for (uint32_t i = tid; i < (uint32_t)(2 * NW * kRadixBins); i += BLOCK) (&s.hist[0][0][0])[i] = 0;
  __syncthreads();
  // Unrolled so that several loads are in flight before their atomics.
#pragma unroll 8
  for (uint32_t i = tid; i < n; i += BLOCK) {
    const auto w = order_word(elem_word<KW, EW>(range_load<KW, EW>(in, i)), om);
#pragma unroll
    for (int p = 0; p < P; ++p) atomicAdd(&s.ghist[p][(uint32_t)(w >> (8 * p)) & 0xffu], 1u);
  }
  __syncthreads();
  // Passes whose digit is not constant over the range.
  uint32_t active_mask = 0;
  for (int p = 0; p < P; ++p) {
    const int triv = __syncthreads_or(s.ghist[p][tid & 255u] == n && tid < 256u);
    if (!triv) active_mask |= 1u << p;
  }
  const int nactive = __builtin_popcount(active_mask);
  if (nactive == 0) {
    for (uint32_t i = tid; i < n; i += BLOCK) range_store<KW, EW>(out, i, range_load<KW, EW>(in, i));
    __syncthreads();
    return;
  }
  int j = 0, buf = 0;
  for (int p = 0; p < P; ++p) {
    if (!(active_mask & (1u << p))) continue;
    const RangeIo src = j == 0 ? in : RangeIo{(j - 1) & 1 ? s1 : s0, nullptr, 1u};
    const RangeIo dst = j == nactive - 1 ? out : RangeIo{j & 1 ? s1 : s0, nullptr, 1u};
    const uint32_t shift = 8u * (uint32_t)p;
    // Cursors: exclusive scan of this digit position's histogram.
    if (wid == 0) {
      uint32_t carry = 0;
#pragma unroll
      for (int k = 0; k < kRadixBins / kWave; ++k) {
        const uint32_t d = (uint32_t)k * kWave + lane;
        const uint32_t c = s.ghist[p][d];
        const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(c);
        s.cursor[d] = carry + incl - c;
        carry += readlane(incl, kWave - 1);
      }
    }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n; c0 += TILE) {
      const uint32_t cn = min(TILE, n - c0);
      const uint32_t wbase = wid * (uint32_t)(kWave * IPT);
      Elem<EW> el[IPT];
      uint32_t dg[IPT], rk[IPT];
      // Unconditional array writes (clamped indices, selects), as in tile_pass.
#pragma unroll
      for (int q = 0; q < IPT; ++q) {
        const uint32_t i = wbase + (uint32_t)q * kWave + lane;
        el[q] = range_load<KW, EW>(src, (size_t)c0 + (i < cn ? i : 0u));
        const uint32_t dv = (uint32_t)(order_word(elem_word<KW, EW>(el[q]), om) >> shift) & 0xffu;
        dg[q] = i < cn ? dv : (uint32_t)kRadixBins;
      }
      uint32_t(*h)[kRadixBins] = s.hist[buf];
      const uint64_t lt = lanemask_lt();
#pragma unroll
      for (int q = 0; q < IPT; ++q) {
        const uint32_t d = dg[q];
        const bool valid = d < (uint32_t)kRadixBins;
        const uint64_t peers = match_digit<8>(d, valid);
        const uint32_t before = h[wid][valid ? d : 0u];
        const uint32_t below = (uint32_t)__builtin_popcountll(peers & lt);
        if (valid && below == 0) h[wid][d] = before + (uint32_t)__builtin_popcountll(peers);
        rk[q] = valid ? before + below : 0u;
      }
      __syncthreads();
      for (uint32_t i = tid; i < (uint32_t)(NW * kRadixBins); i += BLOCK)
        (&s.hist[buf ^ 1][0][0])[i] = 0;
      if (wid == 0) {
        digit_scan_wave0<NW, kRadixBins>(h, s.cstart, s.ctot);
#pragma unroll
        for (int k = 0; k < kRadixBins / kWave; ++k) {
          const uint32_t d = (uint32_t)k * kWave + lane;
          s.gbase[d] = s.cursor[d] - s.cstart[d];
          s.cursor[d] += s.ctot[d];
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < IPT; ++q)
        if (dg[q] < (uint32_t)kRadixBins) s.el[h[wid][dg[q]] + rk[q]] = el[q];
      __syncthreads();
      for (uint32_t i = tid; i < cn; i += BLOCK) {
        const Elem<EW> e = s.el[i];
        const uint32_t d = (uint32_t)(order_word(elem_word<KW, EW>(e), om) >> shift) & 0xffu;
        range_store<KW, EW>(dst, (size_t)s.gbase[d] + i, e);
      }
      buf ^= 1;
    }
    __syncthreads();
    ++j;
  }
}

}  // namespace msv
