// Merge-path merge, merge_by_key and ordered multiset operations for gfx950.
//
// merge / merge_by_key (massively/src/api/algorithm/merge.rs:109-282,
// core/merge.rs): stable; on equal keys the left element comes first
// (core/merge.rs:182-195, oracle vector.rs:493-513). A partition kernel finds
// the merge-path split of every 2048-output tile with one binary search per
// tile; each block loads its two input windows into LDS, every thread finds
// its own split inside them and merges 8 outputs, and the outputs (and the
// values, read at their source positions) leave through LDS as contiguous runs.
//
// set_union / set_intersection / set_difference (api/algorithm/set.rs:44-195,
// multiset rules core/set.rs:18-46, oracle vector.rs:595-664): the oracle walks
// both inputs once, pairing equal keys (the k-th occurrence on the left with
// the k-th on the right). Splitting that walk is valid at any point that does
// not separate a matched pair, so tile and thread splits are taken from the
// merge path and then moved, inside a run of equal keys, to the nearest
// pair-aligned point (both sides consumed equally, or one side's run
// exhausted). Each thread then runs the oracle's loop on its own range; the
// emitted counts are compacted with a block scan plus a decoupled lookback
// over tiles, and the block of the last tile stores the output's
// device-resident length.
#include "device_common.hpp"
#include "msv_internal.hpp"

namespace msv {

// Merge: 512-thread blocks (4.5% faster than 256 at 2 x 2^27 u32); the set
// kernel keeps 256 (512 was 30% slower there).
constexpr int kMergeBlock = 512;
constexpr int kSetBlock = 256;
constexpr int kMergeIpt = 16;
constexpr uint32_t kMergeTile = kMergeBlock * kMergeIpt;
// Set operations emit through a decoupled look-back, so they use larger
// tiles (fewer tiles in the look-back chain).
constexpr int kSetIpt = 32;
constexpr uint32_t kSetTile = kSetBlock * kSetIpt;

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

// Loads the two input windows A[i0, i0 + la) ++ B[j0, j0 + lb) into padded
// LDS: ITEMS loads per thread, all issued before the first LDS store (a loop
// with a store after every load waits for each load in turn).
template <typename K, int ITEMS, int BLOCK>
__device__ __forceinline__ void load_windows(const K* __restrict__ A, uint32_t i0, uint32_t la,
                                             const K* __restrict__ B, uint32_t j0, uint32_t L,
                                             K* s_keys, uint32_t tid) {
  K v[ITEMS];
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const uint32_t k = tid + (uint32_t)q * BLOCK;
    v[q] = k < la ? A[i0 + k] : (k < L ? B[j0 + k - la] : K());
  }
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const uint32_t k = tid + (uint32_t)q * BLOCK;
    if (k < L) s_keys[lds_pad(k)] = v[q];
  }
}

// Keys are compared through their order words (order_bits).
template <typename K, int MODE>
struct Ord {
  __device__ static inline typename Bits<K>::U w(K k) { return order_bits<K, MODE>(k); }
};

// Number of A elements among the first d outputs of the stable merge
// (ties: A first). Works on global or LDS pointers.
template <typename K, int MODE, typename AA, typename BA>
__device__ inline uint32_t merge_path(const AA& A, uint32_t na, const BA& B, uint32_t nb, uint32_t d) {
  uint32_t lo = d > nb ? d - nb : 0;
  uint32_t hi = d < na ? d : na;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (!(Ord<K, MODE>::w(B[d - 1 - mid]) < Ord<K, MODE>::w(A[mid]))) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename K, int MODE, typename AA>
__device__ inline uint32_t lower_bound_w(const AA& A, uint32_t lo, uint32_t hi,
                                         typename Bits<K>::U v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (Ord<K, MODE>::w(A[mid]) < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename K, int MODE, typename AA>
__device__ inline uint32_t upper_bound_w(const AA& A, uint32_t lo, uint32_t hi,
                                         typename Bits<K>::U v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (!(v < Ord<K, MODE>::w(A[mid]))) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Merge-path split at diagonal d, moved to the nearest pair-aligned point of
// the oracle's set walk when it falls inside a run of equal keys.
template <typename K, int MODE, typename AA, typename BA>
__device__ inline void set_split(const AA& A, uint32_t na, const BA& B, uint32_t nb, uint32_t d,
                                 uint32_t& i_out, uint32_t& j_out) {
  using U = typename Bits<K>::U;
  uint32_t i = merge_path<K, MODE>(A, na, B, nb, d);
  uint32_t j = d - i;
  // Previous element in merge order (ties: B after A) and next element (ties: A first).
  const bool has_prev = (i > 0) || (j > 0);
  const bool has_next = (i < na) || (j < nb);
  if (has_prev && has_next) {
    U prev, next;
    if (i > 0 && j > 0) {
      const U a = Ord<K, MODE>::w(A[i - 1]), b = Ord<K, MODE>::w(B[j - 1]);
      prev = a < b ? b : (b < a ? a : b);
    } else {
      prev = i > 0 ? Ord<K, MODE>::w(A[i - 1]) : Ord<K, MODE>::w(B[j - 1]);
    }
    if (i < na && j < nb) {
      const U a = Ord<K, MODE>::w(A[i]), b = Ord<K, MODE>::w(B[j]);
      next = b < a ? b : a;
    } else {
      next = i < na ? Ord<K, MODE>::w(A[i]) : Ord<K, MODE>::w(B[j]);
    }
    if (prev == next) {
      const U v = next;
      const uint32_t lba = lower_bound_w<K, MODE>(A, 0, i, v);
      const uint32_t lbb = lower_bound_w<K, MODE>(B, 0, j, v);
      const uint32_t uba = upper_bound_w<K, MODE>(A, i, na, v);
      const uint32_t ubb = upper_bound_w<K, MODE>(B, j, nb, v);
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
template <typename K, int MODE, bool SET>
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
    set_split<K, MODE>(A, na, B, nb, d, i, j);
  } else {
    i = merge_path<K, MODE>(A, na, B, nb, d);
    j = d - i;
  }
  sa[t] = i;
  sb[t] = j;
}

template <int VS>
struct MergeValue { using T = uint32_t; };
template <>
struct MergeValue<8> { using T = uint64_t; };

// ---- merge ----------------------------------------------------------------
template <typename K, int MODE, int VS>
__global__ __launch_bounds__(kMergeBlock) void merge_tile_kernel(
    const K* __restrict__ A, uint32_t na, const K* __restrict__ B, uint32_t nb,
    const void* __restrict__ Av, const void* __restrict__ Bv, const uint32_t* __restrict__ sa,
    const uint32_t* __restrict__ sb, K* __restrict__ out_k, void* __restrict__ out_v) {
  using VT = typename MergeValue<VS>::T;
  constexpr uint32_t CAP = LdsPadded<kMergeTile>::value;
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_src[VS ? CAP : 1];
  const uint32_t tid = threadIdx.x, b = blockIdx.x;
  const uint32_t i0 = sa[b], i1 = sa[b + 1], j0 = sb[b], j1 = sb[b + 1];
  const uint32_t la = i1 - i0, lb = j1 - j0, L = la + lb;
  load_windows<K, kMergeIpt, kMergeBlock>(A, i0, la, B, j0, L, s_keys, tid);
  __syncthreads();
  const PadWindow<K> LA{s_keys, 0};
  const PadWindow<K> LB{s_keys, la};
  const uint32_t d = min(tid * (uint32_t)kMergeIpt, L);
  uint32_t ia = merge_path<K, MODE>(LA, la, LB, lb, d);
  uint32_t jb = d - ia;
  K ok[kMergeIpt];
  uint32_t os[kMergeIpt];
#pragma unroll
  for (int q = 0; q < kMergeIpt; ++q) {
    if (d + q < L) {
      bool take_a;
      if (ia >= la) take_a = false;
      else if (jb >= lb) take_a = true;
      else take_a = !(Ord<K, MODE>::w(LB[jb]) < Ord<K, MODE>::w(LA[ia]));
      if (take_a) {
        ok[q] = LA[ia];
        os[q] = i0 + ia;
        ++ia;
      } else {
        ok[q] = LB[jb];
        os[q] = na + j0 + jb;
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
  const size_t o = (size_t)i0 + j0;
  if constexpr (VS > 0) {
    // Values are read at their source positions: all loads first, then the stores.
    VT vv[kMergeIpt];
#pragma unroll
    for (int q = 0; q < kMergeIpt; ++q) {
      const uint32_t k = tid + (uint32_t)q * kMergeBlock;
      if (k < L) {
        const uint32_t s = s_src[lds_pad(k)];
        vv[q] = s < na ? ((const VT*)Av)[s] : ((const VT*)Bv)[s - na];
      }
    }
#pragma unroll
    for (int q = 0; q < kMergeIpt; ++q) {
      const uint32_t k = tid + (uint32_t)q * kMergeBlock;
      if (k < L) {
        out_k[o + k] = s_keys[lds_pad(k)];
        ((VT*)out_v)[o + k] = vv[q];
      }
    }
  } else {
    for (uint32_t k = tid; k < L; k += kMergeBlock) out_k[o + k] = s_keys[lds_pad(k)];
  }
}

// ---- ordered set operations ---------------------------------------------------
enum SetKind : int { SET_UNION = 0, SET_INTERSECTION = 1, SET_DIFFERENCE = 2 };

// Output count of set tile q, computed by one wave from global memory (the
// self-help path of the lookback, for a predecessor that is not resident):
// lane l walks the pair-aligned slice [32 l, 32 l + 32) of the tile.
template <typename K, int MODE, int KIND>
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
  set_split<K, MODE>(TA, la, TB, lb, min(lane * SLICE, L), ia, jb);
  set_split<K, MODE>(TA, la, TB, lb, min(lane * SLICE + SLICE, L), ia_e, jb_e);
  uint32_t cnt = 0;
  for (uint32_t q2 = 0; q2 < SLICE + 2 && (ia < ia_e || jb < jb_e); ++q2) {
    int step;
    if (ia < ia_e && jb < jb_e) {
      const auto wa = Ord<K, MODE>::w(TA[ia]), wb = Ord<K, MODE>::w(TB[jb]);
      step = wa < wb ? 0 : (wb < wa ? 1 : 2);
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

template <typename K, int MODE, int KIND>
__global__ __launch_bounds__(kSetBlock) void set_tile_kernel(
    const K* __restrict__ A, uint32_t na, const K* __restrict__ B, uint32_t nb,
    const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb, K* __restrict__ out,
    uint32_t* __restrict__ out_len, uint64_t* counter, uint64_t* status, uint32_t gen,
    uint32_t ntiles) {
  constexpr int NW = kSetBlock / kWave;
  constexpr uint32_t CAP = LdsPadded<kSetTile + 2>::value;
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_split[kSetBlock + 1][2];
  __shared__ uint32_t s_wave[NW];
  __shared__ uint32_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  (void)counter;
  const uint32_t b = blockIdx.x;
  const uint32_t i0 = sa[b], i1 = sa[b + 1], j0 = sb[b], j1 = sb[b + 1];
  const uint32_t la = i1 - i0, lb = j1 - j0, L = la + lb;
  // A set tile holds up to kSetTile + 2 inputs (pair-aligned ends).
  load_windows<K, kSetIpt, kSetBlock>(A, i0, la, B, j0, min(L, (uint32_t)kSetTile), s_keys, tid);
  for (uint32_t k = kSetTile + tid; k < L; k += kSetBlock)
    s_keys[lds_pad(k)] = k < la ? A[i0 + k] : B[j0 + k - la];
  __syncthreads();
  const PadWindow<K> LA{s_keys, 0};
  const PadWindow<K> LB{s_keys, la};
  {
    uint32_t ia, jb;
    set_split<K, MODE>(LA, la, LB, lb, min(tid * (uint32_t)kSetIpt, L), ia, jb);
    s_split[tid][0] = ia;
    s_split[tid][1] = jb;
    if (tid == 0) {
      s_split[kSetBlock][0] = la;
      s_split[kSetBlock][1] = lb;
    }
  }
  __syncthreads();
  uint32_t ia = s_split[tid][0], jb = s_split[tid][1];
  const uint32_t ia_e = s_split[tid + 1][0], jb_e = s_split[tid + 1][1];
  // The oracle's loop (vector.rs:595-664) on this thread's range.
  // A thread range holds at most kSetIpt + 2 steps (both of its ends can
  // move by one when they are pair-aligned). Step q always writes ok[q] (the
  // element it visits) and records in `emit` whether it is an output, so the
  // register array is indexed statically (a running output index would put
  // it in scratch memory).
  constexpr int STEPS = kSetIpt + 2;
  static_assert(STEPS <= 64, "emit mask width");
  K ok[STEPS];
  uint64_t emit = 0;
#pragma unroll
  for (int q = 0; q < STEPS; ++q) {
    const bool ha = ia < ia_e, hb = jb < jb_e;
    const K a = LA[ha ? ia : 0u];
    const K b = LB[hb ? jb : 0u];
    int step;  // 0: left only, 1: right only, 2: matched pair
    if (ha && hb) {
      const auto wa = Ord<K, MODE>::w(a), wb = Ord<K, MODE>::w(b);
      step = wa < wb ? 0 : (wb < wa ? 1 : 2);
    } else {
      step = ha ? 0 : 1;
    }
    const bool act = ha || hb;
    const bool out = act && (step == 0 ? KIND != SET_INTERSECTION
                             : step == 1 ? KIND == SET_UNION
                                         : KIND != SET_DIFFERENCE);
    ok[q] = step == 1 ? b : a;
    emit |= (uint64_t)(out ? 1u : 0u) << q;
    ia += (act && step != 1) ? 1u : 0u;
    jb += (act && step != 0) ? 1u : 0u;
  }
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
        return set_tile_count<K, MODE, KIND>(A, B, sa, sb, q);
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
  for (int q = 0; q < STEPS; ++q) {
    if ((emit >> q) & 1u) {
      s_keys[lds_pad(my)] = ok[q];
      ++my;
    }
  }
  __syncthreads();
  const size_t o = s_prefix;
  for (uint32_t k = tid; k < agg; k += kSetBlock) out[o + k] = s_keys[lds_pad(k)];
}

__global__ void set_len_zero_kernel(uint32_t* p) { *p = 0; }

// ---- host ---------------------------------------------------------------------
template <typename K, int MODE, bool SET>
static msv_status partition(Ctx& c, const K* A, uint32_t na, const K* B, uint32_t nb,
                            uint32_t ntiles, uint32_t tile, uint32_t** sa, uint32_t** sb) {
  *sa = (uint32_t*)c.scratch(sizeof(uint32_t) * (ntiles + 1));
  *sb = (uint32_t*)c.scratch(sizeof(uint32_t) * (ntiles + 1));
  if (!*sa || !*sb) return MSV_ERR_OUT_OF_MEMORY;
  hipLaunchKernelGGL((merge_partition_kernel<K, MODE, SET>), dim3((ntiles + 1 + 255) / 256),
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
  hipLaunchKernelGGL((merge_tile_kernel<K, MODE, VS>), dim3(ntiles), dim3(kMergeBlock), 0,
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
  const uint32_t ntiles = (uint32_t)((total + kSetTile - 1) / kSetTile);
  uint32_t *sa, *sb;
  MSV_TRY((partition<K, MODE, true>(c, (const K*)A, na, (const K*)B, nb, ntiles, kSetTile, &sa,
                                    &sb)));
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
#define MSV_SET_LAUNCH(KIND)                                                                    \
  hipLaunchKernelGGL((set_tile_kernel<K, MODE, KIND>), dim3(ntiles), dim3(kSetBlock), 0,      \
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
