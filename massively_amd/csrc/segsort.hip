// seg::ForEachSegment(Sort): every segment sorted stably on its own, with
// the values moved alongside (massively/src/seg/control.rs:262-330; order of
// the oracle: verification/oracle/src/seg.rs:113-140).
//
// Work items by segment length L, with W = half a tile (TILE = BLOCK * IPT):
//  * windows: the segments with L <= W whose start lies in [k W, (k+1) W)
//    form one contiguous range of fewer than 2 W = TILE elements (a longer
//    segment can only be the last one starting in a window, so excluding it
//    keeps the rest contiguous). The range is sorted in LDS by key and then
//    by segment slot, which restores the segment grouping;
//  * big segments (W < L <= TILE): one segment per tile in LDS;
//  * long segments (L > max(TILE, 4096)): many blocks per segment, an LSD
//    radix sort over all long segments at once (segsort_huge.hpp); with a
//    tile smaller than 4096, segments in between are sorted by one block
//    through global memory, a tile-sized chunk at a time per digit pass.
// A plan kernel finds every window's segments by binary search over the
// offsets and appends the big segments to one list and the long ones to
// another. The sort then runs
// either as one persistent launch (blocks take items from a queue, list items
// first because they are the longest) or as one launch per size class. The
// tile shape (threads, keys per thread, digit width) is one of a few compiled
// candidates, chosen per (key width, element width, mean segment length
// class) by timing them on first use.
#include <algorithm>
#include <atomic>

#include "msv_internal.hpp"
#include "segsort_huge.hpp"
#include "segsort_small.hpp"
#include "tile_sort.hpp"

namespace {

// Resident blocks per CU of one kernel instantiation. The caller keeps one
// cache per instantiation; concurrent first calls (scheduler threads, user
// threads) compute and store the same value, so a relaxed atomic is enough.
template <typename Kern>
int blocks_per_cu(std::atomic<int>& cache, Kern kern, int block) {
  int pc = cache.load(std::memory_order_relaxed);
  if (pc == 0) {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, block, 0) != hipSuccess) b = 1;
    pc = std::max(1, b);
    cache.store(pc, std::memory_order_relaxed);
  }
  return pc;
}

}  // namespace

namespace msv {

struct SegSortArgs {
  const void* keys;
  const void* vals;
  void* out_k;
  void* out_v;
  const uint32_t* offsets;
  const uint2* win;       // per window: [first, end) segment range
  const uint32_t* big;    // segments longer than W
  uint32_t* counters;     // [0] number of big segments, [1] work queue
  void* s0;               // element scratch for huge segments (n elements)
  void* s1;
  uint32_t nwin;
  uint32_t klass;         // SEG_ITEMS_ALL / SEG_ITEMS_WINDOWS / SEG_ITEMS_BIG
  OrderMap om;
};

enum : uint32_t { SEG_ITEMS_ALL = 0, SEG_ITEMS_WINDOWS = 1, SEG_ITEMS_BIG = 2 };

// First index s in [0, m) with a[s] >= v (m if none).
__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t m, uint64_t v) {
  uint32_t lo = 0, hi = m;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void segsort_plan_kernel(const uint32_t* __restrict__ offsets, uint32_t nseg,
                                    uint32_t W, uint32_t nwin, uint32_t long_min,
                                    uint2* __restrict__ win, uint32_t* __restrict__ big,
                                    uint32_t* __restrict__ counters, uint32_t* __restrict__ longs,
                                    uint32_t* __restrict__ long_count) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nwin) return;
  const uint64_t lo = (uint64_t)k * W;
  const uint32_t first = lower_bound_u32(offsets, nseg, lo);
  uint32_t end = lower_bound_u32(offsets, nseg, lo + W);
  if (end > first) {
    const uint32_t last = end - 1;
    const uint32_t L = offsets[last + 1] - offsets[last];
    if (L > W) {
      if (L >= long_min) longs[atomicAdd(long_count, 1u)] = last;
      else big[atomicAdd(&counters[0], 1u)] = last;
      end = last;
    }
  }
  win[k] = make_uint2(first, end);
}

template <int KW, int EW, int BLOCK, int IPT, int DBITS>
struct SegSortSmem {
  using T = TileSmem<EW, BLOCK, IPT, DBITS>;
  using R = RangeSmem<EW, BLOCK, IPT>;
  static constexpr size_t BYTES = sizeof(T) > sizeof(R) ? sizeof(T) : sizeof(R);
};

template <int KW, int EW, int BLOCK, int IPT, int DBITS>
__device__ __forceinline__ void segsort_item(char* smem, const SegSortArgs& a, uint32_t item, uint32_t nbig) {
  using T = TileSmem<EW, BLOCK, IPT, DBITS>;
  using R = RangeSmem<EW, BLOCK, IPT>;
  constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  constexpr int VW = EW - KW;
  if (item < nbig) {
    const uint32_t sg = a.big[item];
    const uint32_t g0 = a.offsets[sg];
    const uint32_t L = a.offsets[sg + 1] - g0;
    if (L <= TILE) {
      tile_sort_range<KW, EW, BLOCK, IPT, DBITS, false>(*(T*)smem, a.keys, a.vals, a.out_k,
                                                         a.out_v, g0, L, a.offsets, 0, 0, a.om);
    } else {
      const RangeIo in{(const char*)a.keys + (size_t)g0 * KW * 4,
                       VW ? (const char*)a.vals + (size_t)g0 * VW * 4 : nullptr, 0u};
      const RangeIo out{(char*)a.out_k + (size_t)g0 * KW * 4,
                        VW ? (char*)a.out_v + (size_t)g0 * VW * 4 : nullptr, 0u};
      range_sort<KW, EW, BLOCK, IPT>(*(R*)smem, in, out, (Elem<EW>*)a.s0 + g0,
                                     (Elem<EW>*)a.s1 + g0, L, a.om);
    }
    return;
  }
  const uint2 w = a.win[item - nbig];
  if (w.y <= w.x) return;
  const uint32_t g0 = a.offsets[w.x];
  const uint32_t n = a.offsets[w.y] - g0;
  if (n == 0) return;
  tile_sort_range<KW, EW, BLOCK, IPT, DBITS, true>(*(T*)smem, a.keys, a.vals, a.out_k, a.out_v,
                                                    g0, n, a.offsets, w.x, w.y, a.om);
}

template <int KW, int EW, int BLOCK, int IPT, int DBITS>
__global__ __launch_bounds__(BLOCK) void segsort_kernel(SegSortArgs a) {
  using SM = SegSortSmem<KW, EW, BLOCK, IPT, DBITS>;
  __shared__ __align__(16) char smem[SM::BYTES];
  __shared__ uint32_t s_item;
  const uint32_t nbig = a.counters[0];
  const uint32_t total = nbig + a.nwin;
  // Per-class launches take one item per block; the persistent launch loops
  // over the queue. One call site keeps the item code inlined once.
  for (uint32_t round = 0;; ++round) {
    uint32_t item = total;
    if (a.klass == SEG_ITEMS_ALL) {
      if (threadIdx.x == 0) s_item = atomicAdd(&a.counters[1], 1u);
      __syncthreads();
      item = s_item;
      __syncthreads();
    } else if (round == 0) {
      if (a.klass == SEG_ITEMS_WINDOWS) item = nbig + blockIdx.x;
      else if (blockIdx.x < nbig) item = blockIdx.x;
    }
    if (item >= total) break;
    segsort_item<KW, EW, BLOCK, IPT, DBITS>(smem, a, item, nbig);
  }
}

// ---- host ---------------------------------------------------------------------

struct SegSortConfig {
  int block, ipt, dbits;
};
// The candidates, in the order of MSV_OPT_SEGSORT_CONFIG.
constexpr SegSortConfig kSegSortConfigs[] = {
    {256, 8, 8}, {256, 16, 8}, {512, 8, 8}, {512, 16, 8}, {256, 16, 4}, {256, 4, 8},
};
constexpr uint32_t kSegSortNumConfigs = sizeof(kSegSortConfigs) / sizeof(kSegSortConfigs[0]);
constexpr int kSegSortDefault = 1;

// A tile of `tile` elements of EW dwords must fit the LDS budget of a block.
constexpr bool segsort_fits(int ew, int tile) { return (size_t)tile * ew * 4 <= 65536; }

// Long segments (listed in `longs`, their number in ctl[HUGE_NSEG] on the
// device) sorted by many blocks each (segsort_huge.hpp). `long_min` is the
// shortest length that can be listed, which bounds the lists' sizes. The
// plan holds everything the launches need, set up on the host first (scratch
// and one look-back generation per pass), so that the launches can go to a
// forked stream.
template <int KW>
struct LongPlan {
  static constexpr int P = KW * 4;
  HugeArgs h;
  uint64_t* status[P];
  uint32_t gen[P];
  uint32_t tiles_max = 0;
  uint32_t nlong_max = 0;
};

template <int KW, int EW>
static msv_status long_prepare(Ctx& c, const SegSortArgs& base, uint32_t n, uint32_t long_min,
                               const uint32_t* longs, uint32_t* ctl, void* s0, void* s1,
                               LongPlan<KW>* lp) {
  constexpr int P = KW * 4;
  lp->nlong_max = n / long_min;
  if (lp->nlong_max == 0) return MSV_OK;
  lp->tiles_max = div_up(n, kHugeTile) + lp->nlong_max;
  uint32_t* tile_begin = (uint32_t*)c.scratch(sizeof(uint32_t) * lp->nlong_max);
  uint32_t* tile_seg = (uint32_t*)c.scratch(sizeof(uint32_t) * lp->tiles_max);
  uint32_t* hist = (uint32_t*)c.scratch(sizeof(uint32_t) * (size_t)lp->nlong_max * P * kRadixBins);
  if (!tile_begin || !tile_seg || !hist) return MSV_ERR_OUT_OF_MEMORY;
  HugeArgs& h = lp->h;
  h.keys = base.keys;
  h.vals = base.vals;
  h.out_k = base.out_k;
  h.out_v = base.out_v;
  h.offsets = base.offsets;
  h.list = longs;
  h.ctl = ctl;
  h.tile_begin = tile_begin;
  h.tile_seg = tile_seg;
  h.hist = hist;
  h.s0 = s0;
  h.s1 = s1;
  h.om = base.om;
  for (int p = 0; p < P; ++p) {
    uint64_t* counter;
    MSV_TRY(c.lookback((size_t)lp->tiles_max * kRadixBins, &counter, &lp->status[p], &lp->gen[p]));
  }
  return MSV_OK;
}

template <int KW, int EW>
static msv_status long_launch(Ctx& c, hipStream_t st, const LongPlan<KW>& lp) {
  constexpr int P = KW * 4;
  if (lp.nlong_max == 0) return MSV_OK;
  const HugeArgs& h = lp.h;
  hipLaunchKernelGGL((huge_layout_kernel<P>), dim3(1), dim3(256), 0, st, h);
  const uint32_t zgrid = std::min<uint32_t>(div_up(lp.nlong_max * P * kRadixBins, 256u * 8u), 1024u);
  hipLaunchKernelGGL((huge_zero_kernel<P>), dim3(std::max(1u, zgrid)), dim3(256), 0, st, h);
  const uint32_t ncu = (uint32_t)c.num_cus();
  // About 8 consecutive tiles per histogram block (fewer when there are
  // fewer tiles than blocks the device holds at once).
  hipLaunchKernelGGL((huge_hist_kernel<KW, EW>),
                     dim3(std::max(1u, std::min(div_up(lp.tiles_max, 8u), ncu * 4u))),
                     dim3(kHugeBlock), 0, st, h);
  hipLaunchKernelGGL((huge_starts_kernel<P>), dim3(std::min(lp.nlong_max, ncu)), dim3(kRadixBins),
                     0, st, h);
  MSV_HIP(hipGetLastError());
  static std::atomic<int> per_cu_cache{0};  // resident blocks per CU of this instantiation
  const int per_cu = blocks_per_cu(per_cu_cache, huge_pass_kernel<KW, EW>, kHugeBlock);
  const uint32_t pgrid = std::min<uint32_t>(lp.tiles_max, (uint32_t)per_cu * ncu);
  for (int p = 0; p < P; ++p) {
    hipLaunchKernelGGL((huge_pass_kernel<KW, EW>), dim3(pgrid), dim3(kHugeBlock), 0, st, h, p,
                       lp.status[p], lp.gen[p]);
    MSV_HIP(hipGetLastError());
  }
  return MSV_OK;
}

// Shortest segment of the long path for a block tile of `tile` elements.
constexpr uint32_t seg_long_min(uint32_t tile) {
  return (tile > kHugeTile ? tile : kHugeTile) + 1u;
}

// Segments of up to 1024 elements sorted inside one wave
// (segsort_small.hpp), those up to the long limit by the block path (one tile
// per segment, or the one-block range sort), the long ones by the long path,
// all from lists the classify and fill kernels build on the device.
template <int KW, int EW, int BLOCK, int IPT, int DBITS>
static msv_status segsort_wave_go(Ctx& c, const SegSortArgs& base, uint32_t n, uint32_t nseg,
                                  const uint32_t* offsets) {
  constexpr int VW = EW - KW;
  if constexpr (!segsort_fits(EW, BLOCK * IPT)) {
    set_error(MSV_ERR_UNSUPPORTED, 0, 0, "tile shape does not fit this element width");
    return MSV_ERR_UNSUPPORTED;
  } else {
    constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
    const uint32_t long_min = seg_long_min(TILE);
    // ctl: [0, 13) class counts, [16, 29) cursors ([27] = the block path's
    // count, [28] its queue, never a cursor), [32, 52) the long path's words.
    uint32_t* ctl = (uint32_t*)c.scratch(sizeof(uint32_t) * 64);
    uint2* list = (uint2*)c.scratch(sizeof(uint2) * (size_t)nseg);
    uint32_t* big = (uint32_t*)c.scratch(sizeof(uint32_t) * (size_t)nseg);
    const uint32_t nlong_max = n / long_min;
    uint32_t* longs = (uint32_t*)c.scratch(sizeof(uint32_t) * std::max(1u, nlong_max));
    if (!ctl || !list || !big || !longs) return MSV_ERR_OUT_OF_MEMORY;
    uint32_t* cursors = ctl + 16;
    uint32_t* hctl = ctl + 32;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, c.stream(), ctl, 64u, 0u);
    const uint32_t sblocks = div_up(nseg, 256u * kClassifySpt);
    hipLaunchKernelGGL(seg_classify_kernel, dim3(sblocks), dim3(256), 0, c.stream(), offsets, nseg,
                       long_min, ctl);
    hipLaunchKernelGGL(seg_fill_kernel, dim3(sblocks), dim3(256), 0, c.stream(), offsets, nseg,
                       long_min, ctl, cursors, list, big, longs, hctl + HUGE_NSEG);
    MSV_HIP(hipGetLastError());
    WaveSortArgs w;
    w.keys = base.keys;
    w.vals = base.vals;
    w.out_k = base.out_k;
    w.out_v = base.out_v;
    w.list = list;
    w.counts = ctl;
    w.om = base.om;
    // Segments longer than 1024: the block path over its list (no windows)
    // and the long path, on a forked stream beside the wave sorts (they have
    // few, long work items; the wave sorts have many short ones).
    const bool longer = (uint64_t)n > kWaveSortMaxLen;
    SegSortArgs b = base;
    LongPlan<KW> lp;
    if (longer) {
      void* s0 = c.scratch((size_t)n * EW * 4);
      void* s1 = c.scratch((size_t)n * EW * 4);
      if (!s0 || !s1) return MSV_ERR_OUT_OF_MEMORY;
      b.win = nullptr;
      b.big = big;
      b.counters = cursors + kSegClassBlock;  // [0] block-path segments, [1] queue
      b.s0 = s0;
      b.s1 = s1;
      b.nwin = 0;
      b.klass = SEG_ITEMS_ALL;
      MSV_TRY((long_prepare<KW, EW>(c, base, n, long_min, longs, hctl, s0, s1, &lp)));
    }
    auto kern = segsort_kernel<KW, EW, BLOCK, IPT, DBITS>;
    static std::atomic<int> per_cu_cache{0};  // resident blocks per CU of this instantiation
    const int per_cu = longer ? blocks_per_cu(per_cu_cache, kern, BLOCK) : 1;
    hipStream_t aux = nullptr;
    if (longer) {
      aux = c.fork();
      if (!aux) return MSV_ERR_LAUNCH;
      const uint32_t most = (uint32_t)(n / (kWaveSortMaxLen + 1));  // segments longer than 1024
      const uint32_t bgrid = std::min<uint32_t>(std::max(1u, most), (uint32_t)(per_cu * c.num_cus()));
      hipLaunchKernelGGL(kern, dim3(bgrid), dim3(BLOCK), 0, aux, b);
      MSV_HIP(hipGetLastError());
      const msv_status ls = long_launch<KW, EW>(c, aux, lp);
      if (ls != MSV_OK) {
        (void)c.join();
        return ls;
      }
    }
    // Enough waves to fill the device; each walks its share of the units.
    // The multi-register classes run on the main stream; the one-slot classes
    // follow the block and long paths on the forked stream, which otherwise
    // ends well before the main one.
    const uint32_t ncu = (uint32_t)c.num_cus();
    const uint32_t sgrid = std::max(1u, std::min(div_up(nseg, kWaveSortWaves), ncu * 8u));
    if constexpr (KW == 1) {
      // 4-byte keys: a pair of waves per segment (5 waves per SIMD); with
      // 8-byte keys its LDS would allow fewer waves than the one-wave form.
      const uint32_t pgrid = std::max(1u, std::min(nseg, ncu * 10u));
      hipLaunchKernelGGL((wave_sort_pair_kernel<KW, VW>), dim3(pgrid), dim3(kPairBlock), 0,
                         c.stream(), w);
    } else {
      hipLaunchKernelGGL((wave_sort_large_kernel<KW, VW, 16>), dim3(sgrid), dim3(kWaveSortBlock),
                         0, c.stream(), w);
    }
    hipLaunchKernelGGL((wave_sort_large_kernel<KW, VW, 8>), dim3(sgrid), dim3(kWaveSortBlock), 0,
                       c.stream(), w);
    hipLaunchKernelGGL((wave_sort_large_kernel<KW, VW, 4>), dim3(sgrid), dim3(kWaveSortBlock), 0,
                       c.stream(), w);
    hipLaunchKernelGGL((wave_sort_large_kernel<KW, VW, 2>), dim3(sgrid), dim3(kWaveSortBlock), 0,
                       c.stream(), w);
    hipStream_t small = longer ? aux : c.stream();
    hipLaunchKernelGGL((wave_sort_small_kernel<KW, VW, 6, 6>), dim3(sgrid), dim3(kWaveSortBlock), 0,
                       small, w);
    hipLaunchKernelGGL((wave_sort_small_kernel<KW, VW, 5, 5>), dim3(sgrid), dim3(kWaveSortBlock), 0,
                       small, w);
    hipLaunchKernelGGL((wave_sort_small_kernel<KW, VW, 0, 4>), dim3(sgrid), dim3(kWaveSortBlock), 0,
                       small, w);
    const hipError_t le = hipGetLastError();
    if (longer) MSV_TRY(c.join());
    MSV_HIP(le);
    return MSV_OK;
  }
}

template <int KW, int EW, int BLOCK, int IPT, int DBITS>
static msv_status segsort_go(Ctx& c, const SegSortArgs& base, uint32_t n, uint32_t nseg,
                             const uint32_t* offsets) {
  if (c.ex()->opt_segsort_wave)
    return segsort_wave_go<KW, EW, BLOCK, IPT, DBITS>(c, base, n, nseg, offsets);
  if constexpr (!segsort_fits(EW, BLOCK * IPT)) {
    set_error(MSV_ERR_UNSUPPORTED, 0, 0, "tile shape does not fit this element width");
    return MSV_ERR_UNSUPPORTED;
  } else {
    constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
    constexpr uint32_t W = TILE / 2;
    const uint32_t long_min = seg_long_min(TILE);
    const uint32_t nwin = (uint32_t)(((uint64_t)n + W - 1) / W);
    const uint32_t nlong_max = n / long_min;
    uint2* win = (uint2*)c.scratch(sizeof(uint2) * nwin);
    uint32_t* big = (uint32_t*)c.scratch(sizeof(uint32_t) * nwin);
    // [0] big segments, [1] queue, [32, 52) the long path's words.
    uint32_t* counters = (uint32_t*)c.scratch(sizeof(uint32_t) * 64);
    uint32_t* longs = (uint32_t*)c.scratch(sizeof(uint32_t) * std::max(1u, nlong_max));
    void* s0 = c.scratch((size_t)n * EW * 4);
    void* s1 = c.scratch((size_t)n * EW * 4);
    if (!win || !big || !counters || !longs || !s0 || !s1) return MSV_ERR_OUT_OF_MEMORY;
    uint32_t* hctl = counters + 32;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, c.stream(), counters, 64u, 0u);
    hipLaunchKernelGGL(segsort_plan_kernel, dim3((nwin + 255) / 256), dim3(256), 0, c.stream(),
                       offsets, nseg, W, nwin, long_min, win, big, counters, longs,
                       hctl + HUGE_NSEG);
    MSV_HIP(hipGetLastError());
    SegSortArgs a = base;
    a.win = win;
    a.big = big;
    a.counters = counters;
    a.s0 = s0;
    a.s1 = s1;
    a.nwin = nwin;
    auto kern = segsort_kernel<KW, EW, BLOCK, IPT, DBITS>;
    if (c.ex()->opt_seg_launch == MSV_SEG_LAUNCH_PER_CLASS) {
      a.klass = SEG_ITEMS_BIG;
      hipLaunchKernelGGL(kern, dim3(nwin), dim3(BLOCK), 0, c.stream(), a);
      a.klass = SEG_ITEMS_WINDOWS;
      hipLaunchKernelGGL(kern, dim3(nwin), dim3(BLOCK), 0, c.stream(), a);
    } else {
      static std::atomic<int> per_cu_cache{0};  // resident blocks per CU of this instantiation
      const int per_cu = blocks_per_cu(per_cu_cache, kern, BLOCK);
      a.klass = SEG_ITEMS_ALL;
      const uint32_t grid = std::min<uint32_t>(2 * nwin, (uint32_t)(per_cu * c.num_cus()));
      hipLaunchKernelGGL(kern, dim3(std::max(1u, grid)), dim3(BLOCK), 0, c.stream(), a);
    }
    MSV_HIP(hipGetLastError());
    LongPlan<KW> lp;
    MSV_TRY((long_prepare<KW, EW>(c, base, n, long_min, longs, hctl, s0, s1, &lp)));
    return long_launch<KW, EW>(c, c.stream(), lp);
  }
}

template <int KW, int EW>
static msv_status segsort_config_go(Ctx& c, int k, const SegSortArgs& a, uint32_t n,
                                    uint32_t nseg, const uint32_t* offsets) {
  switch (k) {
    case 0: return segsort_go<KW, EW, 256, 8, 8>(c, a, n, nseg, offsets);
    case 1: return segsort_go<KW, EW, 256, 16, 8>(c, a, n, nseg, offsets);
    case 2: return segsort_go<KW, EW, 512, 8, 8>(c, a, n, nseg, offsets);
    case 3: return segsort_go<KW, EW, 512, 16, 8>(c, a, n, nseg, offsets);
    case 4: return segsort_go<KW, EW, 256, 16, 4>(c, a, n, nseg, offsets);
    case 5: return segsort_go<KW, EW, 256, 4, 8>(c, a, n, nseg, offsets);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, (uint64_t)k, 0, "bad seg sort config %d", k);
  return MSV_ERR_INVALID_ARGUMENT;
}

// Autotuner key (tuning table family "segsort"): key words, element words,
// launch mode, log2 of the mean segment length, and whether short segments
// are sorted inside a wave.
static uint32_t tune_key(int kw, int ew, int64_t launch, uint32_t n, uint32_t nseg, int64_t wave) {
  const uint32_t mean = nseg ? n / nseg : n;
  const uint32_t cls = mean ? 31u - (uint32_t)__builtin_clz(mean) : 0u;
  return (uint32_t)kw | ((uint32_t)ew << 4) | ((uint32_t)launch << 8) | (cls << 12) |
         ((wave ? 1u : 0u) << 20);
}

template <int KW, int EW>
static msv_status segsort_dispatch(Ctx& c, const SegSortArgs& a, uint32_t n, uint32_t nseg,
                                   const uint32_t* offsets) {
  msv_executor* ex = c.ex();
  if (ex->opt_segsort_config >= 0)
    return segsort_config_go<KW, EW>(c, (int)ex->opt_segsort_config, a, n, nseg, offsets);
  const uint32_t key = tune_key(KW, EW, ex->opt_seg_launch, n, nseg, ex->opt_segsort_wave);
  const int tuned = tune_get(ex, "segsort", key);
  if (tuned >= 0 && tuned < (int)kSegSortNumConfigs &&
      segsort_fits(EW, kSegSortConfigs[tuned].block * kSegSortConfigs[tuned].ipt))
    return segsort_config_go<KW, EW>(c, tuned, a, n, nseg, offsets);
  // Timing needs host synchronisation, so a graph capture uses the default.
  if (!ex->opt_autotune || n < (1u << 16) || ex->capturing)
    return segsort_config_go<KW, EW>(c, kSegSortDefault, a, n, nseg, offsets);
  // Time every candidate that fits (each run is a complete, correct sort of
  // this input; the last one timed leaves its result in the output).
  hipEvent_t e0, e1;
  MSV_HIP(hipEventCreate(&e0));
  MSV_HIP(hipEventCreate(&e1));
  int best = kSegSortDefault;
  float best_ms = 1e30f;
  for (uint32_t k = 0; k < kSegSortNumConfigs; ++k) {
    if (!segsort_fits(EW, kSegSortConfigs[k].block * kSegSortConfigs[k].ipt)) continue;
    {
      Ctx warm(ex);  // scratch goes back to the executor's cache per candidate
      MSV_TRY((segsort_config_go<KW, EW>(warm, (int)k, a, n, nseg, offsets)));
    }
    MSV_HIP(hipEventRecord(e0, c.stream()));
    {
      Ctx timed(ex);
      MSV_TRY((segsort_config_go<KW, EW>(timed, (int)k, a, n, nseg, offsets)));
    }
    MSV_HIP(hipEventRecord(e1, c.stream()));
    MSV_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MSV_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = (int)k;
    }
  }
  MSV_HIP(hipEventDestroy(e0));
  MSV_HIP(hipEventDestroy(e1));
  tune_put(ex, "segsort", key, best);
  return segsort_config_go<KW, EW>(c, best, a, n, nseg, offsets);
}

msv_status seg_sort_run(Ctx& c, msv_dtype kdt, int mode, const void* keys, const void* vals,
                        size_t vsize, uint32_t n, const uint32_t* offsets, uint32_t nseg,
                        void* out_k, void* out_v) {
  if (n == 0 || nseg == 0) return MSV_OK;
  SegSortArgs a{};
  a.keys = keys;
  a.vals = vals;
  a.out_k = out_k;
  a.out_v = out_v;
  a.offsets = offsets;
  a.om.kind = dtype_is_float(kdt) ? KEY_FLOAT
              : (kdt == MSV_I32 || kdt == MSV_I64) ? KEY_SIGNED
                                                   : KEY_UNSIGNED;
  a.om.mode = mode;
  const int kw = (int)dtype_size(kdt) / 4;
  const int vw = (vals && out_v) ? (int)vsize / 4 : 0;
  if (kw == 1) {
    if (vw == 0) return segsort_dispatch<1, 1>(c, a, n, nseg, offsets);
    if (vw == 1) return segsort_dispatch<1, 2>(c, a, n, nseg, offsets);
    return segsort_dispatch<1, 3>(c, a, n, nseg, offsets);
  }
  if (vw == 0) return segsort_dispatch<2, 2>(c, a, n, nseg, offsets);
  if (vw == 1) return segsort_dispatch<2, 3>(c, a, n, nseg, offsets);
  return segsort_dispatch<2, 4>(c, a, n, nseg, offsets);
}

}  // namespace msv

extern "C" {

uint32_t msv_segsort_config_count(void) { return msv::kSegSortNumConfigs; }

int32_t msv_segsort_tuned_config(uint32_t key_bytes, uint32_t value_bytes, int64_t launch,
                                 uint32_t n, uint32_t nseg) {
  const int kw = (int)key_bytes / 4, ew = kw + (int)value_bytes / 4;
  return msv::tune_get_current("segsort", msv::tune_key(kw, ew, launch, n, nseg, 1));
}

int32_t msv_segsort_tuned_config_windows(uint32_t key_bytes, uint32_t value_bytes, int64_t launch,
                                         uint32_t n, uint32_t nseg) {
  const int kw = (int)key_bytes / 4, ew = kw + (int)value_bytes / 4;
  return msv::tune_get_current("segsort", msv::tune_key(kw, ew, launch, n, nseg, 0));
}

msv_status msv_segsort_config_info(uint32_t k, uint32_t* block, uint32_t* items_per_thread,
                                   uint32_t* digit_bits) {
  msv::clear_error();
  if (k >= msv::kSegSortNumConfigs) {
    msv::set_error(MSV_ERR_INVALID_ARGUMENT, k, 0, "bad seg sort config %u", k);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (block) *block = (uint32_t)msv::kSegSortConfigs[k].block;
  if (items_per_thread) *items_per_thread = (uint32_t)msv::kSegSortConfigs[k].ipt;
  if (digit_bits) *digit_bits = (uint32_t)msv::kSegSortConfigs[k].dbits;
  return MSV_OK;
}

}  // extern "C"
