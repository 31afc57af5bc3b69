// Long segments of seg::Sort (longer than a block's tile) sorted by many
// blocks at once: an LSD radix sort of 8-bit digits over all of them
// together, in which every segment is its own sorting problem.
//
//  1. layout (one block): each long segment is cut into tiles of kHugeTile
//     elements; tile_begin[h] is segment h's first tile and tile_seg[t] the
//     segment of tile t;
//  2. histograms: per segment, the counts of every digit at every digit
//     position (hist[h][p][d]), each block counting a run of consecutive
//     tiles;
//  3. starts (one block per segment step): hist[h][p][] becomes its exclusive
//     scan, the start of each digit's run inside the segment in pass p. A
//     digit position on which some segment has keys with more than one value
//     sets that pass's flag; passes with no flag are skipped for every
//     segment (all of them keep their order on that digit);
//  4. one pass per flagged digit position: blocks claim tiles in order from
//     the pass's counter (so every earlier tile is already held by a running
//     block), rank the tile's keys by digit, publish the tile's digit counts,
//     and look back over the earlier tiles *of the same segment* only (the
//     segment's first tile publishes an inclusive prefix at once). A digit's
//     run of the tile goes to segment start + digit start + prefix.
// Ties keep their input order (tiles in order, ranks in order inside a tile),
// so each segment's sort is stable; elements move whole, so -0.0 and +0.0 keep
// their bits under the `<` ordering. Between passes elements are kept
// interleaved in the scratch arrays s0 / s1 at their global positions.
#pragma once

#include "tile_sort.hpp"

namespace msv {

constexpr int kHugeBlock = 256;
constexpr int kHugeIpt = 16;
constexpr uint32_t kHugeTile = (uint32_t)kHugeBlock * kHugeIpt;

// ctl words
enum : uint32_t {
  HUGE_NSEG = 0,     // long segments listed (written by the classifying kernel)
  HUGE_NTILES = 1,   // tiles over all of them (layout)
  HUGE_RESERVED = 2,
  HUGE_FLAGS = 4,    // [4, 12): pass p needed
  HUGE_QUEUE = 12,   // [12, 20): tile queue of pass p
  HUGE_CTL_WORDS = 20,
};

struct HugeArgs {
  const void* keys;
  const void* vals;
  void* out_k;
  void* out_v;
  const uint32_t* offsets;
  const uint32_t* list;   // the long segments' ids
  uint32_t* ctl;
  uint32_t* tile_begin;   // per long segment
  uint32_t* tile_seg;     // per tile
  uint32_t* hist;         // per long segment: P x 256 counts, then starts
  void* s0;               // element scratch, indexed by global position
  void* s1;
  OrderMap om;
};

// 1. Layout and cleared counters (one block of 256 threads). Also zeroes the
//    histograms of the listed segments.
template <int P>
__global__ __launch_bounds__(256) void huge_layout_kernel(HugeArgs a) {
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_carry;
  const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
  const uint32_t nh = a.ctl[HUGE_NSEG];
  if (t == 0) s_carry = 0;
  if (t >= HUGE_RESERVED && t < HUGE_CTL_WORDS) a.ctl[t] = 0;
  __syncthreads();
  for (uint32_t h0 = 0; h0 < nh; h0 += 256) {
    const uint32_t h = h0 + t;
    uint32_t tiles = 0;
    if (h < nh) {
      const uint32_t s = a.list[h];
      tiles = (a.offsets[s + 1] - a.offsets[s] + kHugeTile - 1) / kHugeTile;
    }
    const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(tiles);
    if (lane == 63) s_wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = s_carry, tot = 0;
    for (uint32_t k = 0; k < 4; ++k) {
      if (k < wid) woff += s_wsum[k];
      tot += s_wsum[k];
    }
    const uint32_t first = woff + incl - tiles;
    if (h < nh) {
      a.tile_begin[h] = first;
      for (uint32_t k = 0; k < tiles; ++k) a.tile_seg[first + k] = h;
    }
    __syncthreads();
    if (t == 0) s_carry += tot;
    __syncthreads();
  }
  if (t == 0) a.ctl[HUGE_NTILES] = s_carry;
}

// Zeroes hist[0 .. nh * P * 256) (grid-stride; the count is read on device).
template <int P>
__global__ void huge_zero_kernel(HugeArgs a) {
  const size_t words = (size_t)a.ctl[HUGE_NSEG] * P * kRadixBins;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x)
    a.hist[i] = 0;
}

__device__ inline void huge_tile_range(const HugeArgs& a, uint32_t t, uint32_t* h, size_t* g0,
                                       uint32_t* L, uint32_t* k) {
  *h = a.tile_seg[t];
  const uint32_t s = a.list[*h];
  *g0 = a.offsets[s];
  *L = a.offsets[s + 1] - (uint32_t)*g0;
  *k = t - a.tile_begin[*h];
}

// 2. Digit histograms of every long segment, all digit positions at once.
//    Block b counts the consecutive tiles [b * chunk, (b + 1) * chunk) into
//    per-wave LDS histograms and adds them to the segment's histogram when the
//    segment changes and at the end (tiles of one segment are consecutive), so
//    a long segment's counters take one global add per block, not per tile.
template <int KW, int EW>
__global__ __launch_bounds__(kHugeBlock) void huge_hist_kernel(HugeArgs a) {
  constexpr int P = KW * 4;
  constexpr int NW = kHugeBlock / kWave;
  __shared__ uint32_t s_hw[NW][P * kRadixBins];
  const uint32_t tid = threadIdx.x;
  const uint32_t ntiles = a.ctl[HUGE_NTILES];
  const uint32_t chunk = (ntiles + gridDim.x - 1) / gridDim.x;
  const uint32_t t0 = blockIdx.x * chunk;
  const uint32_t t1 = min(ntiles, t0 + chunk);
  if (t0 >= t1) return;
  uint32_t* s_h = s_hw[tid / kWave];
  const RangeIo in{a.keys, a.vals, 0u};
  for (uint32_t i = tid; i < (uint32_t)(NW * P * kRadixBins); i += kHugeBlock) (&s_hw[0][0])[i] = 0;
  __syncthreads();
  for (uint32_t t = t0; t < t1; ++t) {
    uint32_t h, L, k;
    size_t g0;
    huge_tile_range(a, t, &h, &g0, &L, &k);
    const uint32_t lo = k * kHugeTile;
    KeyWord<KW> w[kHugeIpt];
#pragma unroll
    for (int j = 0; j < kHugeIpt; ++j) {
      const uint32_t i = lo + (uint32_t)j * kHugeBlock + tid;
      w[j] = i < L ? order_word(elem_word<KW, EW>(range_load<KW, EW>(in, g0 + i)), a.om)
                   : (KeyWord<KW>)0;
    }
#pragma unroll
    for (int j = 0; j < kHugeIpt; ++j) {
      const uint32_t i = lo + (uint32_t)j * kHugeBlock + tid;
      if (i < L) {
#pragma unroll
        for (int p = 0; p < P; ++p)
          atomicAdd(&s_h[p * kRadixBins + ((uint32_t)(w[j] >> (8 * p)) & 0xffu)], 1u);
      }
    }
    // Flush at the segment's last tile in this chunk.
    if (t + 1 == t1 || a.tile_seg[t + 1] != h) {
      __syncthreads();
      uint32_t* gh = a.hist + (size_t)h * P * kRadixBins;
      for (uint32_t i = tid; i < (uint32_t)(P * kRadixBins); i += kHugeBlock) {
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          c += s_hw[q][i];
          s_hw[q][i] = 0;
        }
        if (c) atomicAdd(&gh[i], c);
      }
      __syncthreads();
    }
  }
}

// 3. Per long segment: digit starts of every pass, and the pass flags.
template <int P>
__global__ __launch_bounds__(kRadixBins) void huge_starts_kernel(HugeArgs a) {
  __shared__ uint32_t s_wsum[kRadixBins / kWave];
  const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
  const uint32_t nh = a.ctl[HUGE_NSEG];
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint32_t s = a.list[h];
    const uint32_t L = a.offsets[s + 1] - a.offsets[s];
    uint32_t* gh = a.hist + (size_t)h * P * kRadixBins;
    for (int p = 0; p < P; ++p) {
      const uint32_t c = gh[p * kRadixBins + t];
      const int trivial = __syncthreads_or(c == L);
      if (!trivial && t == 0) atomicOr(&a.ctl[HUGE_FLAGS + p], 1u);
      const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(c);
      if (lane == 63) s_wsum[wid] = incl;
      __syncthreads();
      uint32_t woff = 0;
      for (uint32_t k = 0; k < wid; ++k) woff += s_wsum[k];
      gh[p * kRadixBins + t] = woff + incl - c;
      __syncthreads();
    }
  }
}

// 4. One digit pass over every long segment.
template <int KW, int EW>
__global__ __launch_bounds__(kHugeBlock) void huge_pass_kernel(HugeArgs a, int pass,
                                                               uint64_t* __restrict__ status,
                                                               uint32_t gen) {
  constexpr int P = KW * 4;
  constexpr int NW = kHugeBlock / kWave;
  constexpr uint32_t TILE = kHugeTile;
  constexpr int IPT = kHugeIpt;
  __shared__ Elem<EW> s_el[TILE];
  __shared__ uint32_t s_hist[NW][kRadixBins];   // per-wave counts, then wave offsets
  __shared__ uint32_t s_tile_start[kRadixBins];  // in-tile start of each digit
  __shared__ uint32_t s_goff[kRadixBins];        // global position - in-tile position
  __shared__ uint32_t s_wsum[kRadixBins / kWave];
  __shared__ uint32_t s_cnt[kRadixBins];
  __shared__ uint32_t s_tile;

  // Which passes run (a pass no segment needs is skipped; with none needed
  // the last one still moves the data), and where this one reads and writes.
  uint32_t mask = 0;
#pragma unroll
  for (int p = 0; p < P; ++p) mask |= a.ctl[HUGE_FLAGS + p] ? (1u << p) : 0u;
  if (mask == 0) mask = 1u << (P - 1);
  if (!(mask & (1u << pass))) return;
  const int j = __builtin_popcount(mask & ((1u << pass) - 1u));
  const int nactive = __builtin_popcount(mask);
  const RangeIo src = j == 0 ? RangeIo{a.keys, a.vals, 0u}
                             : RangeIo{((j - 1) & 1) ? a.s1 : a.s0, nullptr, 1u};
  const RangeIo dst = j == nactive - 1 ? RangeIo{a.out_k, a.out_v, 0u}
                                       : RangeIo{(j & 1) ? a.s1 : a.s0, nullptr, 1u};
  const uint32_t shift = 8u * (uint32_t)pass;
  const uint32_t ntiles = a.ctl[HUGE_NTILES];

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const bool digit_thread = tid < (uint32_t)kRadixBins;
  for (;;) {
    for (uint32_t i = tid; i < NW * kRadixBins; i += kHugeBlock) (&s_hist[0][0])[i] = 0;
    if (digit_thread) s_cnt[tid] = 0;
    if (tid == 0) s_tile = atomicAdd(&a.ctl[HUGE_QUEUE + pass], 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= ntiles) break;
    uint32_t h, L, k;
    size_t g0;
    huge_tile_range(a, t, &h, &g0, &L, &k);
    const uint32_t lo = k * TILE;
    const uint32_t tile_n = min(TILE, L - lo);
    // Wave w holds in-tile positions [w * 64 * IPT, (w + 1) * 64 * IPT).
    const uint32_t wbase = wid * (uint32_t)(kWave * IPT);

    Elem<EW> el[IPT];
    uint32_t dig[IPT];
    uint32_t rank[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t i = wbase + (uint32_t)q * kWave + lane;
      if (i < tile_n) {
        el[q] = range_load<KW, EW>(src, g0 + lo + i);
        dig[q] = (uint32_t)(order_word(elem_word<KW, EW>(el[q]), a.om) >> shift) & 0xffu;
      } else {
        dig[q] = 0x100u;
      }
    }
    uint32_t dstart = 0;
    if (digit_thread) dstart = a.hist[((size_t)h * P + pass) * kRadixBins + tid];
#pragma unroll
    for (int q = 0; q < IPT; ++q)
      if (dig[q] < 0x100u) atomicAdd(&s_cnt[dig[q]], 1u);
    __syncthreads();
    uint64_t* my_status = status + (size_t)t * kRadixBins + tid;
    uint32_t cnt = 0;
    if (digit_thread) {
      cnt = s_cnt[tid];
      st_agent(my_status, pack_status(gen, k == 0 ? kFlagInc : kFlagAgg, 0, cnt));
    }

    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t d = dig[q];
      const bool valid = d < 0x100u;
      const uint64_t peers = match_digit<8>(d, valid);
      uint32_t before = 0;
      if (valid) before = s_hist[wid][d];
      const uint32_t below = (uint32_t)__builtin_popcountll(peers & lt);
      if (valid && below == 0) s_hist[wid][d] = before + (uint32_t)__builtin_popcountll(peers);
      rank[q] = before + below;
    }
    __syncthreads();

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
      for (uint32_t q = 0; q < wid; ++q) woff += s_wsum[q];
      const uint32_t tstart = woff + incl - cnt;
      s_tile_start[tid] = tstart;
      // Look back over the segment's earlier tiles (t - k is its first).
      uint32_t prefix = 0;
      if (k > 0) {
        // All tiles of a long segment run at once, so inclusive prefixes
        // spread from its first tile one look-back step at a time: read 16
        // predecessors per step.
        constexpr int B = 16;
        const int64_t first = (int64_t)t - k;
        int64_t q = (int64_t)t - 1;
        for (;;) {
          uint64_t w[B];
#pragma unroll
          for (int b = 0; b < B; ++b)
            w[b] = q - b >= first ? ld_agent(status + (size_t)(q - b) * kRadixBins + tid) : 0ull;
          int used = 0;
          bool done = false;
#pragma unroll
          for (int b = 0; b < B; ++b) {
            if (done || used < b) continue;
            const uint32_t f = status_flag(w[b], gen);
            if (f == 0) continue;
            prefix += (uint32_t)w[b];
            used = b + 1;
            done = f == kFlagInc;
          }
          if (done) break;
          q -= used;
          if (used == 0) __builtin_amdgcn_s_sleep(1);
        }
        st_agent(my_status, pack_status(gen, kFlagInc, 0, prefix + cnt));
      }
      s_goff[tid] = dstart + prefix - tstart;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IPT; ++q)
      if (dig[q] < 0x100u) s_el[s_tile_start[dig[q]] + s_hist[wid][dig[q]] + rank[q]] = el[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t i = tid + (uint32_t)q * kHugeBlock;
      if (i < tile_n) {
        const Elem<EW> e = s_el[i];
        const uint32_t d = (uint32_t)(order_word(elem_word<KW, EW>(e), a.om) >> shift) & 0xffu;
        range_store<KW, EW>(dst, g0 + (size_t)s_goff[d] + i, e);
      }
    }
    __syncthreads();
  }
}

}  // namespace msv
