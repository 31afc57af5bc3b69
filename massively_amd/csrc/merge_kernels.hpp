// Merge-path kernels for gfx950, generic over the ordering: merge and
// merge_by_key, the ordered multiset operations, and the comparison merge sort
// (block sort + pairwise merge passes) that sorts by a user comparator.
//
// An ordering is a type Cmp with `static bool less(K a, K b)`: a comes
// strictly before b. The built-in Less / Greater compare order words
// (OrdLess); a user BinaryPredicateOp becomes such a type in code compiled at
// run time (hiprtc), which instantiates these same templates.
//
// merge / merge_by_key (massively/src/api/algorithm/merge.rs:109-282,
// core/merge.rs): stable; on equal keys the left element comes first
// (core/merge.rs:182-195, oracle vector.rs:493-513). A partition kernel finds
// the merge-path split of every 8192-output tile with one binary search per
// tile; each block loads its two input windows into LDS, every thread finds
// its own split inside them and merges 16 outputs, and the outputs (and the
// values, read at their source positions) leave through LDS as contiguous
// runs.
//
// set_union / set_intersection / set_difference (api/algorithm/set.rs:44-195,
// multiset rules core/set.rs:18-46, oracle vector.rs:595-664): the oracle walks
// both inputs once, pairing equal keys (the k-th occurrence on the left with
// the k-th on the right). Splitting that walk is valid at any point that does
// not separate a matched pair, so tile splits are taken from the merge path
// and then moved, inside a run of equal keys, to the nearest pair-aligned
// point (both sides consumed equally, or one side's run exhausted). Inside a
// tile that starts and ends pair-aligned, the walk's decision for an element
// depends only on counts within the tile: a left copy of rank r among the
// tile's left copies of its key is paired (kept by the intersection, dropped
// by the difference) iff r is below the number of right copies in the tile,
// and a right copy of rank r is unpaired (kept by the union) iff r reaches
// the number of left copies. Threads therefore walk plain merge-path ranges of
// the tile (ties: left first) and flag each element from the neighbouring
// keys the walk already holds; the emitted counts are compacted with a block
// scan plus a decoupled lookback over tiles, and the block of the last tile
// stores the output's device-resident length.
//
// Comparison sort (sort / sort_by_key with a user comparator; reference:
// core/ordering/mod.rs:325-748, stable, ties keep input order): a block sorts
// a 4096-element tile in LDS (each thread sorts 8 elements in registers by
// odd-even transposition, then 9 rounds of merge-path merges of neighbouring
// runs in LDS), then passes of the merge kernel merge neighbouring runs of
// width 4096, 8192, ... until one run remains. Every merge takes the left run
// on ties, so the sort is stable.
#pragma once

#include "device_common.hpp"

namespace msv {

// Merge: 512-thread blocks (4.5% faster than 256 at 2 x 2^27 u32).
constexpr int kMergeBlock = 512;
constexpr int kMergeIpt = 16;
constexpr uint32_t kMergeTile = kMergeBlock * kMergeIpt;
// Set operations emit through a decoupled look-back, so they use larger
// tiles (fewer tiles in the look-back chain): 512 x 16 = 8192 elements, at
// 60 VGPRs and 34 KB of LDS (8 waves per SIMD). 256 x 32 (the same tile at
// 86-90 VGPRs, 4 waves per SIMD) was 4-5% slower at 2 x 2^27 u32, 256 x 16
// (half the tile) 2-8% slower, 512 x 32 30% slower.
constexpr int kSetBlock = 512;
constexpr int kSetIpt = 16;
constexpr uint32_t kSetTile = kSetBlock * kSetIpt;
// Tile diagonals are kSetSpan apart: moving a split to a pair-aligned point
// lowers its diagonal by at most one, so a tile holds at most kSetTile inputs.
constexpr uint32_t kSetSpan = kSetTile - 1;
// Block sort of the comparison sort: 512 threads x 8 elements. Two of its
// tiles make one merge tile, so every merge-pass tile lies inside one pair
// of runs.
constexpr int kSortBlock = 512;
constexpr int kSortIpt = 8;
constexpr uint32_t kSortTile = kSortBlock * kSortIpt;
static_assert(2 * kSortTile == kMergeTile, "a merge tile is one pair of block-sorted runs");

// LDS windows are indexed with one padding word per 32 (x -> x + x/32):
// every thread walks its own range, so the lanes of a wave read at a stride
// of IPT/2 .. IPT elements, which without padding would hit 2-4 of the 32 LDS
// banks (16- to 32-way conflicts on every step of the walk).
__device__ __forceinline__ uint32_t lds_pad(uint32_t x) { return x + (x >> 5); }
template <uint32_t N>
struct LdsPadded { static constexpr uint32_t value = N + N / 32 + 1; };

// A window of LDS keys starting at logical index `off`, read through lds_pad.
template <typename K>
struct PadWindow {
  const K* base;
  uint32_t off;
  __device__ __forceinline__ K operator[](uint32_t i) const { return base[lds_pad(off + i)]; }
};

// Loads the two input windows A[0, la) ++ B[0, L - la) into padded LDS:
// ITEMS loads per thread, all issued before the first LDS store (a loop with
// a store after every load waits for each load in turn).
template <typename K, int ITEMS, int BLOCK>
__device__ __forceinline__ void load_windows(const K* __restrict__ A, uint32_t la,
                                             const K* __restrict__ B, uint32_t L, K* s_keys,
                                             uint32_t tid) {
  K v[ITEMS];
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const uint32_t k = tid + (uint32_t)q * BLOCK;
    v[q] = k < la ? A[k] : (k < L ? B[k - la] : K());
  }
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const uint32_t k = tid + (uint32_t)q * BLOCK;
    if (k < L) s_keys[lds_pad(k)] = v[q];
  }
}

// The built-in orderings: keys compared through their order words
// (order_bits: Less and Greater, -0.0 == +0.0 for floats).
template <typename K, int MODE>
struct OrdLess {
  __device__ static inline bool less(K a, K b) {
    return order_bits<K, MODE>(a) < order_bits<K, MODE>(b);
  }
};

// Number of A elements among the first d outputs of the stable merge
// (ties: A first). Works on global or LDS windows.
template <typename K, typename Cmp, typename AA, typename BA>
__device__ inline uint32_t merge_path(const AA& A, uint32_t na, const BA& B, uint32_t nb, uint32_t d) {
  uint32_t lo = d > nb ? d - nb : 0;
  uint32_t hi = d < na ? d : na;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (!Cmp::less(B[d - 1 - mid], A[mid])) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// First index in [lo, hi) whose element is not before v.
template <typename K, typename Cmp, typename AA>
__device__ inline uint32_t lower_bound_c(const AA& A, uint32_t lo, uint32_t hi, K v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (Cmp::less(A[mid], v)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// First index in [lo, hi) whose element is after v.
template <typename K, typename Cmp, typename AA>
__device__ inline uint32_t upper_bound_c(const AA& A, uint32_t lo, uint32_t hi, K v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (!Cmp::less(v, A[mid])) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename K, typename Cmp>
__device__ inline bool equivalent(K a, K b) {
  return !Cmp::less(a, b) && !Cmp::less(b, a);
}

// The same bounds found by galloping outward from a position known to hold v
// (the run of keys equivalent to v around a split): O(log run length) steps
// instead of O(log n), so a split inside a short run costs a few loads.
// First index in [0, i) from which every element up to i is not before v.
template <typename K, typename Cmp, typename AA>
__device__ inline uint32_t gallop_lower(const AA& A, uint32_t i, K v) {
  uint32_t lo = i, step = 1;  // A[lo, i) are not before v
  while (lo > 0) {
    const uint32_t probe = lo > step ? lo - step : 0u;
    if (Cmp::less(A[probe], v)) return lower_bound_c<K, Cmp>(A, probe + 1, lo, v);
    lo = probe;
    step *= 2;
  }
  return 0;
}
// First index in [i, n) whose element is after v.
template <typename K, typename Cmp, typename AA>
__device__ inline uint32_t gallop_upper(const AA& A, uint32_t i, uint32_t n, K v) {
  uint32_t hi = i, step = 1;  // A[i, hi) are not after v
  while (hi < n) {
    const uint32_t probe = n - hi > step ? hi + step : n;  // exclusive end of the next window
    if (Cmp::less(v, A[probe - 1])) return upper_bound_c<K, Cmp>(A, hi, probe - 1, v);
    hi = probe;
    step *= 2;
  }
  return n;
}

// Merge-path split at diagonal d, moved to the nearest pair-aligned point of
// the oracle's set walk when it falls inside a run of equal keys.
template <typename K, typename Cmp, typename AA, typename BA>
__device__ inline void set_split(const AA& A, uint32_t na, const BA& B, uint32_t nb, uint32_t d,
                                 uint32_t& i_out, uint32_t& j_out) {
  uint32_t i = merge_path<K, Cmp>(A, na, B, nb, d);
  uint32_t j = d - i;
  // Previous element in merge order (ties: B after A) and next element (ties: A first).
  const bool has_prev = (i > 0) || (j > 0);
  const bool has_next = (i < na) || (j < nb);
  if (has_prev && has_next) {
    K prev, next;
    if (i > 0 && j > 0) {
      const K a = A[i - 1], b = B[j - 1];
      prev = Cmp::less(a, b) ? b : (Cmp::less(b, a) ? a : b);
    } else {
      prev = i > 0 ? A[i - 1] : B[j - 1];
    }
    if (i < na && j < nb) {
      const K a = A[i], b = B[j];
      next = Cmp::less(b, a) ? b : a;
    } else {
      next = i < na ? A[i] : B[j];
    }
    if (equivalent<K, Cmp>(prev, next)) {
      const K v = next;
      const uint32_t lba = gallop_lower<K, Cmp>(A, i, v);
      const uint32_t lbb = gallop_lower<K, Cmp>(B, j, v);
      const uint32_t uba = gallop_upper<K, Cmp>(A, i, na, v);
      const uint32_t ubb = gallop_upper<K, Cmp>(B, j, nb, v);
      const uint32_t ma = uba - lba, mb = ubb - lbb;
      const uint32_t mp = ma < mb ? ma : mb;
      const uint32_t x = (i - lba) + (j - lbb);
      uint32_t ra, rb;
      if (x <= 2 * mp) {
        ra = rb = x / 2;
      } else if (ma > mb) {
        rb = mb;
        ra = x - mb;
      } else {
        ra = ma;
        rb = x - ma;
      }
      i = lba + ra;
      j = lbb + rb;
    }
  }
  i_out = i;
  j_out = j;
}

// ---- partition ------------------------------------------------------------
template <typename K, typename Cmp, bool SET>
__global__ void merge_partition_kernel(const K* __restrict__ A, uint32_t na,
                                       const K* __restrict__ B, uint32_t nb, uint32_t ntiles,
                                       uint32_t tile, uint32_t* __restrict__ sa,
                                       uint32_t* __restrict__ sb) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const uint64_t total = (uint64_t)na + nb;
  const uint32_t d = (uint32_t)min((uint64_t)t * tile, total);
  uint32_t i, j;
  if constexpr (SET) {
    set_split<K, Cmp>(A, na, B, nb, d, i, j);
  } else {
    i = merge_path<K, Cmp>(A, na, B, nb, d);
    j = d - i;
  }
  sa[t] = i;
  sb[t] = j;
}

template <int VS>
struct MergeValue { using T = uint32_t; };
template <>
struct MergeValue<8> { using T = uint64_t; };

// ---- merge of one window ----------------------------------------------------
// Merges A[0, la) and B[0, lb) (global windows, at most kMergeTile in all)
// into out_k[0, la + lb); with VS > 0 the values Av / Bv (of VS bytes, at the
// same positions as their keys) follow into out_v. s_keys / s_src are the
// block's LDS (LdsPadded<kMergeTile> entries each).
template <typename K, typename Cmp, int VS>
__device__ inline void merge_window(const K* __restrict__ A, uint32_t la, const K* __restrict__ B,
                                    uint32_t lb, const void* __restrict__ Av,
                                    const void* __restrict__ Bv, K* __restrict__ out_k,
                                    void* __restrict__ out_v, K* s_keys, uint32_t* s_src) {
  using VT = typename MergeValue<VS>::T;
  const uint32_t tid = threadIdx.x;
  const uint32_t L = la + lb;
  load_windows<K, kMergeIpt, kMergeBlock>(A, la, B, L, s_keys, tid);
  __syncthreads();
  const PadWindow<K> LA{s_keys, 0};
  const PadWindow<K> LB{s_keys, la};
  const uint32_t d = min(tid * (uint32_t)kMergeIpt, L);
  uint32_t ia = merge_path<K, Cmp>(LA, la, LB, lb, d);
  uint32_t jb = d - ia;
  K ok[kMergeIpt];
  uint32_t os[kMergeIpt];
#pragma unroll
  for (int q = 0; q < kMergeIpt; ++q) {
    if (d + q < L) {
      bool take_a;
      if (ia >= la) take_a = false;
      else if (jb >= lb) take_a = true;
      else take_a = !Cmp::less(LB[jb], LA[ia]);
      if (take_a) {
        ok[q] = LA[ia];
        os[q] = ia;
        ++ia;
      } else {
        ok[q] = LB[jb];
        os[q] = la + jb;
        ++jb;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kMergeIpt; ++q) {
    if (d + q < L) {
      s_keys[lds_pad(d + q)] = ok[q];
      if constexpr (VS > 0) s_src[lds_pad(d + q)] = os[q];
    }
  }
  __syncthreads();
  if constexpr (VS > 0) {
    // Values are read at their source positions: all loads first, then the stores.
    VT vv[kMergeIpt];
#pragma unroll
    for (int q = 0; q < kMergeIpt; ++q) {
      const uint32_t k = tid + (uint32_t)q * kMergeBlock;
      if (k < L) {
        const uint32_t s = s_src[lds_pad(k)];
        vv[q] = s < la ? ((const VT*)Av)[s] : ((const VT*)Bv)[s - la];
      }
    }
#pragma unroll
    for (int q = 0; q < kMergeIpt; ++q) {
      const uint32_t k = tid + (uint32_t)q * kMergeBlock;
      if (k < L) {
        out_k[k] = s_keys[lds_pad(k)];
        ((VT*)out_v)[k] = vv[q];
      }
    }
  } else {
    for (uint32_t k = tid; k < L; k += kMergeBlock) out_k[k] = s_keys[lds_pad(k)];
  }
}

// ---- merge ----------------------------------------------------------------
template <typename K, typename Cmp, int VS>
__global__ __launch_bounds__(kMergeBlock) void merge_tile_kernel(
    const K* __restrict__ A, uint32_t na, const K* __restrict__ B, uint32_t nb,
    const void* __restrict__ Av, const void* __restrict__ Bv, const uint32_t* __restrict__ sa,
    const uint32_t* __restrict__ sb, K* __restrict__ out_k, void* __restrict__ out_v) {
  using VT = typename MergeValue<VS>::T;
  constexpr uint32_t CAP = LdsPadded<kMergeTile>::value;
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_src[VS ? CAP : 1];
  (void)na;
  (void)nb;
  const uint32_t b = blockIdx.x;
  const uint32_t i0 = sa[b], i1 = sa[b + 1], j0 = sb[b], j1 = sb[b + 1];
  const size_t o = (size_t)i0 + j0;
  merge_window<K, Cmp, VS>(A + i0, i1 - i0, B + j0, j1 - j0,
                           VS ? (const void*)((const VT*)Av + i0) : nullptr,
                           VS ? (const void*)((const VT*)Bv + j0) : nullptr, out_k + o,
                           VS ? (void*)((VT*)out_v + o) : nullptr, s_keys, s_src);
}

// ---- comparison sort ----------------------------------------------------------
// Sorts each kSortTile-element tile of keys (and moves the values with
// them) into out_k / out_v: stable under Cmp.
template <typename K, typename Cmp, int VS>
__global__ __launch_bounds__(kSortBlock) void block_sort_kernel(const K* __restrict__ keys,
                                                                const void* __restrict__ vals,
                                                                uint32_t n, K* __restrict__ out_k,
                                                                void* __restrict__ out_v) {
  using VT = typename MergeValue<VS>::T;
  constexpr uint32_t CAP = LdsPadded<kSortTile>::value;
  __shared__ K s_k[CAP];
  __shared__ uint32_t s_i[CAP];
  const uint32_t tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * kSortTile;
  const uint32_t cnt = (uint32_t)min((size_t)kSortTile, (size_t)n - base);
  {
    K v[kSortIpt];
#pragma unroll
    for (int q = 0; q < kSortIpt; ++q) {
      const uint32_t k = tid + (uint32_t)q * kSortBlock;
      if (k < cnt) v[q] = keys[base + k];
    }
#pragma unroll
    for (int q = 0; q < kSortIpt; ++q) {
      const uint32_t k = tid + (uint32_t)q * kSortBlock;
      if (k < cnt) s_k[lds_pad(k)] = v[q];
    }
  }
  __syncthreads();
  // This thread's kSortIpt consecutive elements, sorted in registers by
  // odd-even transposition (kSortIpt rounds; a swap needs a strict
  // inversion, so equal keys keep their order).
  const uint32_t my0 = tid * kSortIpt;
  const uint32_t myn = my0 < cnt ? min((uint32_t)kSortIpt, cnt - my0) : 0u;
  K r[kSortIpt];
  uint32_t ri[kSortIpt];
#pragma unroll
  for (int q = 0; q < kSortIpt; ++q) {
    r[q] = (uint32_t)q < myn ? s_k[lds_pad(my0 + q)] : K();
    ri[q] = my0 + q;
  }
#pragma unroll
  for (int round = 0; round < kSortIpt; ++round) {
#pragma unroll
    for (int q = round & 1; q + 1 < kSortIpt; q += 2) {
      if ((uint32_t)q + 1 < myn && Cmp::less(r[q + 1], r[q])) {
        const K tk = r[q];
        r[q] = r[q + 1];
        r[q + 1] = tk;
        const uint32_t ti = ri[q];
        ri[q] = ri[q + 1];
        ri[q + 1] = ti;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kSortIpt; ++q) {
    if ((uint32_t)q < myn) {
      s_k[lds_pad(my0 + q)] = r[q];
      s_i[lds_pad(my0 + q)] = ri[q];
    }
  }
  __syncthreads();
  // Merge rounds: groups of 2s threads merge two runs of s * kSortIpt.
#pragma unroll 1
  for (uint32_t s = 1; s < (uint32_t)kSortBlock; s <<= 1) {
    const uint32_t g0 = (tid & ~(2 * s - 1)) * kSortIpt;  // group start (elements)
    const uint32_t a0 = min(g0, cnt), a1 = min(g0 + s * kSortIpt, cnt);
    const uint32_t b1 = min(g0 + 2 * s * kSortIpt, cnt);
    const uint32_t la = a1 - a0, lb = b1 - a1;
    const uint32_t d = min((tid & (2 * s - 1)) * (uint32_t)kSortIpt, la + lb);
    const PadWindow<K> WA{s_k, a0};
    const PadWindow<K> WB{s_k, a1};
    uint32_t ia = merge_path<K, Cmp>(WA, la, WB, lb, d);
    uint32_t jb = d - ia;
#pragma unroll
    for (int q = 0; q < kSortIpt; ++q) {
      if (d + q < la + lb) {
        bool take_a;
        if (ia >= la) take_a = false;
        else if (jb >= lb) take_a = true;
        else take_a = !Cmp::less(WB[jb], WA[ia]);
        const uint32_t src = take_a ? a0 + ia : a1 + jb;
        r[q] = s_k[lds_pad(src)];
        ri[q] = s_i[lds_pad(src)];
        ia += take_a ? 1u : 0u;
        jb += take_a ? 0u : 1u;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kSortIpt; ++q) {
      if (d + q < la + lb) {
        s_k[lds_pad(a0 + d + q)] = r[q];
        s_i[lds_pad(a0 + d + q)] = ri[q];
      }
    }
    __syncthreads();
  }
  for (uint32_t k = tid; k < cnt; k += kSortBlock) {
    out_k[base + k] = s_k[lds_pad(k)];
    if constexpr (VS > 0) ((VT*)out_v)[base + k] = ((const VT*)vals)[base + s_i[lds_pad(k)]];
  }
}

// Split of merge tile t of one merge-sort pass (runs of `width` elements,
// merged in pairs): the number of elements the tile's start takes from the
// left run of its pair. 2 * width is a multiple of kMergeTile.
template <typename K, typename Cmp>
__global__ void sort_pass_partition_kernel(const K* __restrict__ keys, uint32_t n,
                                           uint32_t width, uint32_t ntiles,
                                           uint32_t* __restrict__ sa) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint64_t d = (uint64_t)t * kMergeTile;
  const uint64_t pair = d / (2ull * width) * (2ull * width);
  const uint32_t na = (uint32_t)min((uint64_t)width, (uint64_t)n - pair);
  const uint32_t nb = (uint32_t)min((uint64_t)width, (uint64_t)n - pair - na);
  sa[t] = merge_path<K, Cmp>(keys + pair, na, keys + pair + na, nb, (uint32_t)(d - pair));
}

// One merge-sort pass: merge tile t of the pair of runs it lies in.
template <typename K, typename Cmp, int VS>
__global__ __launch_bounds__(kMergeBlock) void sort_pass_kernel(
    const K* __restrict__ keys, const void* __restrict__ vals, uint32_t n, uint32_t width,
    const uint32_t* __restrict__ sa, K* __restrict__ out_k, void* __restrict__ out_v) {
  using VT = typename MergeValue<VS>::T;
  constexpr uint32_t CAP = LdsPadded<kMergeTile>::value;
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_src[VS ? CAP : 1];
  const uint32_t t = blockIdx.x;
  const uint64_t d0 = (uint64_t)t * kMergeTile;
  const uint64_t pair = d0 / (2ull * width) * (2ull * width);
  const uint32_t na = (uint32_t)min((uint64_t)width, (uint64_t)n - pair);
  const uint32_t nb = (uint32_t)min((uint64_t)width, (uint64_t)n - pair - na);
  const uint64_t pair_end = pair + na + nb;
  const uint32_t i0 = sa[t], j0 = (uint32_t)(d0 - pair) - i0;
  uint32_t i1, j1;
  if (d0 + kMergeTile >= pair_end) {
    i1 = na;
    j1 = nb;
  } else {
    i1 = sa[t + 1];
    j1 = (uint32_t)(d0 + kMergeTile - pair) - i1;
  }
  const size_t pa = (size_t)pair + i0, pb = (size_t)pair + na + j0;
  merge_window<K, Cmp, VS>(keys + pa, i1 - i0, keys + pb, j1 - j0,
                           VS ? (const void*)((const VT*)vals + pa) : nullptr,
                           VS ? (const void*)((const VT*)vals + pb) : nullptr, out_k + d0,
                           VS ? (void*)((VT*)out_v + d0) : nullptr, s_keys, s_src);
}

// ---- ordered set operations ---------------------------------------------------
enum SetKind : int { SET_UNION = 0, SET_INTERSECTION = 1, SET_DIFFERENCE = 2 };

// Output count of set tile q, computed by one wave from global memory (the
// self-help path of the lookback, for a predecessor that is not resident):
// lane l walks the pair-aligned slice [32 l, 32 l + 32) of the tile.
template <typename K, typename Cmp, int KIND>
__device__ uint32_t set_tile_count(const K* __restrict__ A, const K* __restrict__ B,
                                   const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb,
                                   uint32_t q) {
  constexpr uint32_t SLICE = kSetTile / kWave;
  const uint32_t lane = __lane_id();
  const uint32_t i0 = sa[q], j0 = sb[q];
  const uint32_t la = sa[q + 1] - i0, lb = sb[q + 1] - j0, L = la + lb;
  const K* TA = A + i0;
  const K* TB = B + j0;
  uint32_t ia, jb, ia_e, jb_e;
  set_split<K, Cmp>(TA, la, TB, lb, min(lane * SLICE, L), ia, jb);
  set_split<K, Cmp>(TA, la, TB, lb, min(lane * SLICE + SLICE, L), ia_e, jb_e);
  uint32_t cnt = 0;
  for (uint32_t q2 = 0; q2 < SLICE + 2 && (ia < ia_e || jb < jb_e); ++q2) {
    int step;
    if (ia < ia_e && jb < jb_e) {
      const K a = TA[ia], b = TB[jb];
      step = Cmp::less(a, b) ? 0 : (Cmp::less(b, a) ? 1 : 2);
    } else {
      step = ia < ia_e ? 0 : 1;
    }
    if (step == 0) {
      cnt += KIND != SET_INTERSECTION ? 1u : 0u;
      ++ia;
    } else if (step == 1) {
      cnt += KIND == SET_UNION ? 1u : 0u;
      ++jb;
    } else {
      cnt += KIND != SET_DIFFERENCE ? 1u : 0u;
      ++ia;
      ++jb;
    }
  }
  return wave_reduce<uint32_t, OP_ADD>(cnt);
}

// Rank of v among the left copies of its run below `hi`, and the number of
// copies from `lo` on: searches for the rare case of a key repeated inside
// one side that also occurs on the other (out of line: the walk stays small).
template <typename K, typename Cmp, typename W>
__device__ __attribute__((noinline)) uint32_t set_run_lower(const W& X, uint32_t hi, K v) {
  return lower_bound_c<K, Cmp>(X, 0, hi, v);
}
template <typename K, typename Cmp, typename W>
__device__ __attribute__((noinline)) uint32_t set_run_upper(const W& X, uint32_t lo, uint32_t hi,
                                                           K v) {
  return upper_bound_c<K, Cmp>(X, lo, hi, v);
}

// Merge walk of outputs [d, d + kSetIpt) of the tile's merge (ties: left
// first); step q stores the element it visits in ok[q] and sets bit q of
// `emit` if the operation outputs it. A left element a has right copies in
// the tile iff the right head equals it (every smaller right key was taken
// before it); its rank is 0 unless it repeats the previous left key. A right
// element b has left copies iff the last left key taken equals it. Only a key
// repeated on one side and present on the other needs searches, and their
// result carries along the run.
template <typename K, typename Cmp, int KIND>
__device__ inline void set_flag_walk(const PadWindow<K>& LA, uint32_t la, const PadWindow<K>& LB,
                                     uint32_t lb, uint32_t d, uint32_t L, K* ok, uint64_t& emit) {
  uint32_t ia = merge_path<K, Cmp>(LA, la, LB, lb, d);
  uint32_t jb = d - ia;
  K a = LA[ia < la ? ia : 0u], b = LB[jb < lb ? jb : 0u];
  bool hpa = ia > 0, hpb = jb > 0;
  K pa = LA[hpa ? ia - 1 : 0u], pb = LB[hpb ? jb - 1 : 0u];
  // (rank, other side's count) of the run of pa / pb, valid when va / vb.
  uint32_t ra = 0, ca = 0, rb = 0, cb = 0;
  bool va = false, vb = false;
  emit = 0;
#pragma unroll
  for (int q = 0; q < kSetIpt; ++q) {
    const bool act = d + q < L;
    const bool ha = ia < la, hb = jb < lb;
    const bool take_a = ha && (!hb || !Cmp::less(b, a));
    bool keep;
    if (take_a) {
      const bool eqb = hb && !Cmp::less(a, b);    // the tile's right side holds a copy of a
      const bool dup = hpa && !Cmp::less(pa, a);  // a repeats the previous left key
      uint32_t r = 0, cnt = eqb ? 1u : 0u;         // 1 stands for "at least one"
      if (eqb && dup) {
        if (va) {
          r = ra + 1;
          cnt = ca;
        } else {
          r = ia - set_run_lower<K, Cmp>(LA, ia, a);
          cnt = set_run_upper<K, Cmp>(LB, jb, lb, a) - jb;
        }
      }
      va = !eqb || dup;  // exact unless a first copy met a right copy
      ra = r;
      ca = cnt;
      keep = KIND == SET_UNION ? true : (KIND == SET_INTERSECTION ? r < cnt : r >= cnt);
      ok[q] = a;
      if (act) {
        pa = a;
        hpa = true;
        ++ia;
        a = LA[ia < la ? ia : 0u];
      }
    } else {
      keep = false;
      if constexpr (KIND == SET_UNION) {
        const bool eqa = hpa && !Cmp::less(pa, b);  // the tile's left side holds a copy of b
        const bool dup = hpb && !Cmp::less(pb, b);
        uint32_t r = 0, cnt = eqa ? 1u : 0u;
        if (eqa && dup) {
          if (vb) {
            r = rb + 1;
            cnt = cb;
          } else {
            r = jb - set_run_lower<K, Cmp>(LB, jb, b);
            cnt = ia - set_run_lower<K, Cmp>(LA, ia, b);
          }
        }
        vb = !eqa || dup;
        rb = r;
        cb = cnt;
        keep = r >= cnt;
      }
      ok[q] = b;
      if (act) {
        pb = b;
        hpb = true;
        ++jb;
        b = LB[jb < lb ? jb : 0u];
      }
    }
    emit |= (uint64_t)((act && keep) ? 1u : 0u) << q;
  }
}

template <typename K, typename Cmp, int KIND>
__global__ __launch_bounds__(kSetBlock) void set_tile_kernel(
    const K* __restrict__ A, uint32_t na, const K* __restrict__ B, uint32_t nb,
    const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb, K* __restrict__ out,
    uint32_t* __restrict__ out_len, uint64_t* counter, uint64_t* status, uint32_t gen,
    uint32_t ntiles) {
  constexpr int NW = kSetBlock / kWave;
  constexpr uint32_t CAP = LdsPadded<kSetTile>::value;
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_wave[NW];
  __shared__ uint32_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  (void)counter;
  (void)na;
  (void)nb;
  const uint32_t b = blockIdx.x;
  const uint32_t i0 = sa[b], i1 = sa[b + 1], j0 = sb[b], j1 = sb[b + 1];
  // la + lb <= kSetTile by the kSetSpan spacing; the clamps only keep every
  // LDS access inside the tile.
  const uint32_t la = min(i1 - i0, kSetTile);
  const uint32_t lb = min(j1 - j0, kSetTile - la);
  const uint32_t L = la + lb;
  load_windows<K, kSetIpt, kSetBlock>(A + i0, la, B + j0, L, s_keys, tid);
  __syncthreads();
  const PadWindow<K> LA{s_keys, 0};
  const PadWindow<K> LB{s_keys, la};
  K ok[kSetIpt];
  uint64_t emit;
  set_flag_walk<K, Cmp, KIND>(LA, la, LB, lb, min(tid * (uint32_t)kSetIpt, L), L, ok, emit);
  const uint32_t cnt = (uint32_t)__builtin_popcountll(emit);
  // Block scan of emitted counts, then lookback over tiles.
  const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(cnt);
  if (lane == 63) s_wave[wid] = incl;
  __syncthreads();
  uint32_t woff = 0, agg = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    if (w < (int)wid) woff += s_wave[w];
    agg += s_wave[w];
  }
  if (wid == 0) {
    TileStatus<uint32_t> st{status, gen};
    uint32_t excl = 0;
    if (b == 0) {
      if (lane == 0) st.publish(0, kFlagInc, 0, agg);
    } else {
      if (lane == 0) st.publish(b, kFlagAgg, 0, agg);
      excl = wave_lookback_sh<uint32_t, OP_ADD>(st, b, [&](uint32_t q) {
        return set_tile_count<K, Cmp, KIND>(A, B, sa, sb, q);
      });
      if (lane == 0) st.publish(b, kFlagInc, 0, excl + agg);
    }
    if (lane == 0) {
      s_prefix = excl;
      if (b == ntiles - 1) *out_len = excl + agg;
    }
  }
  __syncthreads();  // everyone is done reading s_keys as input
  uint32_t my = woff + incl - cnt;
#pragma unroll
  for (int q = 0; q < kSetIpt; ++q) {
    if ((emit >> q) & 1u) {
      s_keys[lds_pad(my)] = ok[q];
      ++my;
    }
  }
  __syncthreads();
  const size_t o = s_prefix;
  for (uint32_t k = tid; k < agg; k += kSetBlock) out[o + k] = s_keys[lds_pad(k)];
}

}  // namespace msv
