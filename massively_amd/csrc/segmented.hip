// seg:: (ForEachSegment InclusiveScan / ExclusiveScan / Reduce / Sort) and the
// Segmentation constructors for gfx950.
//
// Reference: massively/src/seg/mod.rs:379-425 (Sort), :458-572 (scans),
// :676-806, 1314-1336 (Reduce, empty segments -> init), seg/control.rs
// (offsets -> head flags), core/segmented_fixed.rs (head-flag segmented scan),
// seg/segmentation.rs:10-52, 187-264 (validation, offsets from lengths).
//
// Scans and reductions run over the flat value array in tiles of 512 threads
// x 4 rows x one 16-byte vector (8192 u32 / 4096 u64). Each tile finds its
// segment heads directly from the CSR offsets: one binary search for the
// first offset at or after the tile start, then the threads mark the heads
// that fall in the tile in an LDS bitmap (empty segments share their start
// with the next segment). The scan is the head-flag segmented scan
// (f1,v1).(f2,v2) = (f1|f2, f2 ? v2 : v1 (+) v2): vector-local, wave (DPP),
// rows x waves in LDS, then across tiles. Integer types resolve the carry
// across tiles with a decoupled lookback whose status word carries the head
// bit (the walk stops at the nearest tile containing a head). Float types use
// fixed-shape passes (tile aggregates, one-block exclusive scan of the
// aggregates, tile pass with its carry) so results are reproducible bit for
// bit. Reduce keeps the tile's inclusive values in LDS and writes, for every
// segment whose last element lies in the tile, init (+) that value; empty
// segments receive init from the tile holding their start (trailing empty
// segments from the last tile).
#include "device_common.hpp"
#include "msv_internal.hpp"

namespace msv {

constexpr int kSegBlock = 512;

enum SegOut : int { SEG_INCL = 0, SEG_EXCL = 1, SEG_REDUCE = 2, SEG_RUNS = 3 };

// Rows of 16-byte vectors per thread (512-thread blocks): 4 (8192 u32 / 4096
// u64 per tile); 2 for SEG_RUNS, whose LDS staging would otherwise limit
// occupancy. 512-thread blocks ran the seg reduce and reduce_by_key 3-7%
// faster than 256-thread blocks with the same tiles. 1024-thread blocks with
// half the rows (the same tiles at 38-52 VGPRs, 8 waves per SIMD instead of
// 4-5) were slower: seg scan 0.153 -> 0.182 ms, reduce_by_key 0.997 -> 1.532
// ms at 2^26 / 2^28 u32.
__host__ __device__ constexpr int seg_rows(int out) { return out == SEG_RUNS ? 2 : 4; }
enum SegMode : int { SEG_LOOKBACK = 0, SEG_AGGREGATE = 1, SEG_PREFIXED = 2 };

template <typename T, int OP>
__device__ inline void seg_combine(uint32_t& f, T& v, uint32_t f2, T v2) {
  v = f2 ? v2 : Combine<T, OP>::apply(v, v2);
  f = f | f2;
}

__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi,
                                           uint32_t v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Extra arguments of the by-key forms. HK (the key type) != void: segment
// heads are run heads, elements whose key differs (key_equal) from the
// previous element's, found from the keys inside each tile. SEG_RUNS
// (reduce_by_key) also counts the heads: each run head writes its key to
// out_keys[run], each run end (an element followed by a head, or the last
// element) writes init (+) its segmented inclusive value to out[run], and the
// last tile stores the run count in both device lengths.
struct RunsArgs {
  const void* keys = nullptr;
  void* out_keys = nullptr;
  uint32_t* len_keys = nullptr;
  uint32_t* len_vals = nullptr;
  uint64_t* cstatus = nullptr;  // look-back form: run-count tile status
  uint32_t* tile_c = nullptr;   // fixed-shape form: tile run counts, then their exclusive prefixes
  const uint32_t* tile_first = nullptr;  // offsets form: per tile, the last segment starting before it
};

// V consecutive elements from p + i (16-byte aligned when full).
template <typename K, int V>
__device__ __forceinline__ void load_run(const K* __restrict__ p, size_t i, uint32_t n, bool full,
                                         K (&v)[V]) {
  if (full) {
    struct alignas(16) Q { K v[V]; };
    const Q q = *reinterpret_cast<const Q*>(p + i);
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = q.v[k];
  } else {
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = (i + k < n) ? p[i + k] : K();
  }
}

template <typename T, int OP, int MODE, int OUT, typename HK = void>
__global__ __launch_bounds__(kSegBlock) void seg_tile_kernel(
    const T* __restrict__ in, uint32_t n, const uint32_t* __restrict__ offsets, uint32_t nseg,
    T* __restrict__ out, T init, bool aligned, uint64_t* status, uint32_t gen, uint32_t ntiles,
    uint32_t* __restrict__ tile_f, T* __restrict__ tile_v, RunsArgs ra) {
  using C = Combine<T, OP>;
  constexpr bool KEYS = !std::is_same<HK, void>::value;
  constexpr bool RUNS = OUT == SEG_RUNS;
  static_assert(!RUNS || KEYS, "runs need keys");
  static_assert(OUT != SEG_REDUCE, "per-segment reduce: seg_reduce_kernel");
  using KT = typename std::conditional<KEYS, HK, uint32_t>::type;
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = kSegBlock / kWave;
  constexpr int kSegRows = seg_rows(OUT);
  constexpr uint32_t TILE = (uint32_t)kSegBlock * kSegRows * V;
  __shared__ uint32_t s_heads[KEYS ? 1 : TILE / 32];
  __shared__ uint32_t s_wf[kSegRows][NW];
  __shared__ T s_wv[kSegRows][NW];
  __shared__ uint32_t s_wc[RUNS ? kSegRows : 1][NW];
  __shared__ uint32_t s_cf;
  __shared__ T s_cv;
  __shared__ uint32_t s_cc;
  __shared__ uint32_t s_first;
  // SEG_RUNS: each thread parks its own keys in s_keys. Each wave's part of
  // a row (a chunk of 64 vectors) later compacts its run keys in place in
  // its part of s_keys and its run values in the same part of s_vals, in run
  // order, and the chunks are then stored contiguously.
  __shared__ KT s_keys[RUNS ? TILE : 1];
  __shared__ T s_rvals[RUNS ? TILE : 1];
  __shared__ uint4 s_chunk[RUNS ? kSegRows : 1][RUNS ? NW : 1];  // run base, heads, h0, ends

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * TILE;
  const uint32_t tile_end = (uint32_t)min((size_t)n, base + TILE);
  const KT* __restrict__ hk = (const KT*)ra.keys;
  const bool vec_ok = aligned && (base + TILE <= n);
  T x[kSegRows][V];
  // The values first: their latency overlaps the search for the tile's
  // segment starts below.
#pragma unroll
  for (int r = 0; r < kSegRows; ++r)
    load_run<T, V>(in, base + ((uint32_t)r * kSegBlock + tid) * V, n, vec_ok, x[r]);
  uint32_t s0 = 0;
  if constexpr (!KEYS) {
    for (uint32_t i = tid; i < TILE / 32; i += kSegBlock) s_heads[i] = 0;
    if (ra.tile_first) {
      s0 = ra.tile_first[tile];
    } else {
      if (wid == 0) {
        const uint32_t lb = wave_lower_bound(offsets, 0, nseg + 1, (uint32_t)base);
        if (lane == 0) s_first = lb > 0 ? lb - 1 : 0;
      }
    }
    // The bitmap is cleared before any thread marks a start in it.
    __syncthreads();
    if (!ra.tile_first) s0 = s_first;
    // Mark the segment starts that fall inside the tile.
    for (uint32_t c = 0;; c += kSegBlock) {
      const uint32_t s = s0 + c + tid;
      bool more = false;
      if (s < nseg) {
        const uint32_t st = offsets[s];
        if (st < tile_end) {
          more = tid == kSegBlock - 1;  // (sorted starts)
          if (st >= base) atomicOr(&s_heads[(st - base) >> 5], 1u << ((st - base) & 31u));
        }
      }
      if (!__syncthreads_or(more)) break;
    }
  }

  const T id = C::identity();
  uint32_t hb[kSegRows];    // bit k: element k of the thread's vector in row r starts a segment
  uint32_t ends[kSegRows];  // SEG_RUNS: bit k = element k of the vector ends a run
#pragma unroll
  for (int r = 0; r < kSegRows; ++r) {
    const uint32_t li = ((uint32_t)r * kSegBlock + tid) * V;
    const size_t gi = base + li;
    if constexpr (KEYS) {
      KT kv[V];
      load_run<KT, V>(hk, gi, n, vec_ok, kv);
      // The key before the vector: the previous lane's last key (lane 0
      // reads it; consecutive threads hold consecutive vectors).
      const KT before = shift_up1<KT>(kv[V - 1], (lane == 0 && gi > 0 && gi <= n) ? hk[gi - 1] : KT());
      uint32_t b = 0;
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const KT pk = k == 0 ? before : kv[k - 1];
        b |= ((gi + k < n) && (gi + k == 0 || !key_equal(pk, kv[k])) ? 1u : 0u) << k;
      }
      hb[r] = b;
      if constexpr (RUNS) {
#pragma unroll
        for (int k = 0; k < V; ++k) s_keys[li + k] = kv[k];
        const KT after = shift_down1<KT>(kv[0], (lane == 63 && gi + V < n) ? hk[gi + V] : KT());
        uint32_t e = 0;
#pragma unroll
        for (int k = 0; k < V; ++k) {
          const KT nk = k == V - 1 ? after : kv[k + 1];
          const bool end = (gi + k < n) && (gi + k + 1 == n || !key_equal(kv[k], nk));
          e |= (end ? 1u : 0u) << k;
        }
        ends[r] = e;
      }
    } else {
      // li is a multiple of V and V divides 32: the vector's bits share a word.
      hb[r] = (s_heads[li >> 5] >> (li & 31u)) & ((1u << V) - 1u);
    }
  }

  // Vector totals and wave scans of them.
  uint32_t wf[kSegRows];
  T wv[kSegRows];
  uint32_t wc[kSegRows];
#pragma unroll
  for (int r = 0; r < kSegRows; ++r) {
    uint32_t f = hb[r] & 1u;
    T v = x[r][0];
#pragma unroll
    for (int k = 1; k < V; ++k) seg_combine<T, OP>(f, v, (hb[r] >> k) & 1u, x[r][k]);
    wf[r] = f;
    wv[r] = v;
    wave_inclusive_seg_scan<T, OP>(wf[r], wv[r]);
    if (lane == 63) {
      s_wf[r][wid] = wf[r];
      s_wv[r][wid] = wv[r];
    }
    if constexpr (RUNS) {
      wc[r] = wave_inclusive_scan<uint32_t, OP_ADD>((uint32_t)__builtin_popcount(hb[r]));
      if (lane == 63) s_wc[r][wid] = wc[r];
    }
  }
  __syncthreads();
  // Exclusive carry of (row r, wave wid) and the tile aggregate: every wave
  // scans the kSegRows x NW partials in (row, wave) order, one per lane, and
  // takes its own with readlane (wave-uniform values).
  static_assert(kSegRows * NW <= kWave, "one partial per lane");
  constexpr int NP = kSegRows * NW;
  uint32_t qf = 0, qc = 0;
  T qv = id;
  if (lane < (uint32_t)NP) {
    qf = s_wf[lane / NW][lane % NW];
    qv = s_wv[lane / NW][lane % NW];
    if constexpr (RUNS) qc = s_wc[lane / NW][lane % NW];
  }
  wave_inclusive_seg_scan<T, OP>(qf, qv);
  const uint32_t af = readlane(qf, NP - 1);
  const T av = readlane(qv, NP - 1);
  const uint32_t xf = shift_up1<uint32_t>(qf, 0u);
  const T xv = shift_up1<T>(qv, id);
  uint32_t ac = 0, xc = 0;
  if constexpr (RUNS) {
    qc = wave_inclusive_scan<uint32_t, OP_ADD>(qc);
    ac = readlane(qc, NP - 1);
    xc = shift_up1<uint32_t>(qc, 0u);
  }
  const int w0 = __builtin_amdgcn_readfirstlane((int)wid);
  uint32_t pf[kSegRows];
  T pv[kSegRows];
  uint32_t pc[kSegRows];
#pragma unroll
  for (int r = 0; r < kSegRows; ++r) {
    pf[r] = readlane(xf, r * NW + w0);
    pv[r] = readlane(xv, r * NW + w0);
    pc[r] = RUNS ? readlane(xc, r * NW + w0) : 0u;
  }

  // Segmented aggregate of tile q from the input (the elements from its last
  // segment start, or all of it if none): look-back self-help.
  auto agg_of = [&](uint32_t q, uint32_t& hf, T& a) {
    const size_t qb = (size_t)q * TILE;
    const uint32_t qe = (uint32_t)min((size_t)n, qb + TILE);
    size_t from = qb;
    hf = 0;
    if constexpr (KEYS) {
      // Last run head of the tile, searched from the end.
      for (int64_t wb = (int64_t)qe - kWave; wb + kWave > (int64_t)qb; wb -= kWave) {
        const int64_t i = wb + (int64_t)lane;
        const bool hd = i >= (int64_t)qb && (i == 0 || !key_equal(hk[i - 1], hk[i]));
        const uint64_t any = ballot(hd);
        if (any) {
          from = (size_t)(wb + 63 - (int64_t)__builtin_clzll(any));
          hf = 1;
          break;
        }
      }
    } else {
      const uint32_t j = lower_bound_u32(offsets, 0, nseg + 1, qe);
      if (j > 0 && offsets[j - 1] >= qb) {
        from = offsets[j - 1];
        hf = 1;
      }
    }
    T acc = id;
    for (size_t i = from + lane; i < qe; i += kWave) acc = C::apply(acc, in[i]);
    a = wave_reduce<T, OP>(acc);
  };

  if constexpr (MODE == SEG_AGGREGATE) {
    if (tid == 0) {
      tile_f[tile] = af;
      tile_v[tile] = av;
      if constexpr (RUNS) ra.tile_c[tile] = ac;
    }
    return;
  } else if constexpr (MODE == SEG_LOOKBACK) {
    if (wid == 0) {
      TileStatus<T> st{status, gen};
      uint32_t cf = 0;
      T cv = id;
      if (tile == 0) {
        if (lane == 0) st.publish(0, kFlagInc, af, av);
      } else {
        if (lane == 0) st.publish(tile, kFlagAgg, af, av);
        wave_lookback_seg_sh<T, OP>(st, tile, cf, cv, agg_of);
        uint32_t nf = cf;
        T nv = cv;
        seg_combine<T, OP>(nf, nv, af, av);
        if (lane == 0) st.publish(tile, kFlagInc, nf, nv);
      }
      if (lane == 0) {
        s_cf = cf;
        s_cv = cv;
      }
    }
    if (RUNS && wid == 1) {
      // The run count before this tile, on its own status words.
      TileStatus<uint32_t> cs{ra.cstatus, gen};
      uint32_t excl = 0;
      if (tile == 0) {
        if (lane == 0) cs.publish(0, kFlagInc, 0, ac);
      } else {
        if (lane == 0) cs.publish(tile, kFlagAgg, 0, ac);
        auto count_of = [&](uint32_t q) -> uint32_t {
          const size_t qb = (size_t)q * TILE;
          const size_t qe = min((size_t)n, qb + TILE);
          uint32_t cnt = 0;
          for (size_t i = qb + lane; i < qe; i += kWave)
            cnt += (i == 0 || !key_equal(hk[i - 1], hk[i])) ? 1u : 0u;
          return wave_reduce<uint32_t, OP_ADD>(cnt);
        };
        excl = wave_lookback_sh<uint32_t, OP_ADD>(cs, tile, count_of);
        if (lane == 0) cs.publish(tile, kFlagInc, 0, excl + ac);
      }
      if (lane == 0) s_cc = excl;
    }
  } else {
    if (tid == 0) {
      s_cf = tile > 0 ? tile_f[tile] : 0u;
      s_cv = tile > 0 ? tile_v[tile] : id;
      if constexpr (RUNS) s_cc = ra.tile_c[tile];
    }
  }
  __syncthreads();
  if constexpr (RUNS) {
    if (tid == 0 && tile == ntiles - 1) {
      *ra.len_keys = s_cc + ac;
      *ra.len_vals = s_cc + ac;
    }
  }

#pragma unroll
  for (int r = 0; r < kSegRows; ++r) {
    // Carry into this thread's vector: tile carry . row/wave carry . wave-exclusive.
    uint32_t cf = s_cf;
    T cv = s_cv;
    seg_combine<T, OP>(cf, cv, pf[r], pv[r]);
    const uint32_t ef = shift_up1<uint32_t>(wf[r], 0u);
    const T ev = shift_up1<T>(wv[r], id);
    seg_combine<T, OP>(cf, cv, ef, ev);
    const uint32_t li = ((uint32_t)r * kSegBlock + tid) * V;
    // The segmented inclusive values continue from the carry (the value of
    // the element before the vector).
    T y[V];
    T run = cv;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const bool head = (hb[r] >> k) & 1u;
      if constexpr (OUT == SEG_EXCL) y[k] = head ? init : C::apply(init, run);
      run = head ? x[r][k] : C::apply(run, x[r][k]);
      if constexpr (OUT != SEG_EXCL) y[k] = run;
    }
    if constexpr (RUNS) {
      // Runs started before this vector, then one per head; consecutive lanes
      // write consecutive runs.
      // This wave's chunk of the row: its heads take runs [wb, wb + wh) and
      // its ends the runs from wb - 1 + h0 on (h0 = 1 if the chunk starts a
      // run). Keys are read before any lane writes, so compacting them in
      // place (rank <= position) is safe.
      const uint32_t cbase = ((uint32_t)r * kSegBlock + wid * kWave) * V;
      uint32_t rank = shift_up1<uint32_t>(wc[r], 0u);
      const uint32_t h0 = readlane(hb[r], 0) & 1u;
      KT kk[V];
#pragma unroll
      for (int k = 0; k < V; ++k) kk[k] = s_keys[li + k];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const uint32_t head = (hb[r] >> k) & 1u;
        if (head) s_keys[cbase + rank] = kk[k];
        rank += head;
        if ((ends[r] >> k) & 1u) s_rvals[cbase + rank - h0] = C::apply(init, y[k]);
      }
      // Runs that meet the chunk: wh + 1 - h0; all but the last end in it
      // unless its last element ends a run (always so for the chunk that
      // holds element n - 1; none meet a chunk past n).
      const size_t cstart = base + cbase;
      const size_t cend = cstart + (size_t)kWave * V;
      const uint32_t wh = readlane(wc[r], kWave - 1);
      const uint32_t elast = readlane(ends[r], kWave - 1) >> (V - 1);
      const uint32_t last = cend >= n ? (cstart < n ? 1u : 0u) : elast;
      if (lane == 0) s_chunk[r][wid] = make_uint4(s_cc + pc[r], wh, h0, wh - h0 + last);
    } else {
      if (vec_ok) {
        struct alignas(16) Q { T v[V]; };
        Q q;
#pragma unroll
        for (int k = 0; k < V; ++k) q.v[k] = y[k];
        *reinterpret_cast<Q*>(out + base + li) = q;
      } else {
#pragma unroll
        for (int k = 0; k < V; ++k)
          if (base + li + k < n) out[base + li + k] = y[k];
      }
    }
  }

  if constexpr (RUNS) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegRows; ++r) {
      const uint4 ch = s_chunk[r][wid];
      const uint32_t cbase = ((uint32_t)r * kSegBlock + wid * kWave) * V;
      KT* ok = (KT*)ra.out_keys + ch.x;
      for (uint32_t i = lane; i < ch.y; i += kWave) ok[i] = s_keys[cbase + i];
      T* ov = out + (ch.x + ch.z - 1u);
      for (uint32_t i = lane; i < ch.w; i += kWave) ov[i] = s_rvals[cbase + i];
    }
  }
}

// One-block exclusive segmented scan of tile aggregates (float path), with
// an optional exclusive sum of per-tile run counts beside it.
template <typename T, int OP>
__global__ __launch_bounds__(1024) void seg_pairs_exclusive_kernel(
    const uint32_t* __restrict__ f_in, const T* __restrict__ v_in, uint32_t m,
    uint32_t* __restrict__ f_out, T* __restrict__ v_out, const uint32_t* __restrict__ c_in,
    uint32_t* __restrict__ c_out) {
  using C = Combine<T, OP>;
  constexpr int NW = 1024 / kWave;
  __shared__ uint32_t s_f[NW];
  __shared__ T s_v[NW];
  __shared__ uint32_t s_c[NW];
  __shared__ uint32_t s_cf, s_cc;
  __shared__ T s_cv;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  if (tid == 0) {
    s_cf = 0;
    s_cc = 0;
    s_cv = C::identity();
  }
  __syncthreads();
  for (uint32_t base = 0; base < m; base += 1024) {
    const uint32_t i = base + tid;
    uint32_t f = i < m ? f_in[i] : 0u;
    T v = i < m ? v_in[i] : C::identity();
    const uint32_t cnt = (c_in && i < m) ? c_in[i] : 0u;
    wave_inclusive_seg_scan<T, OP>(f, v);
    const uint32_t ci = wave_inclusive_scan<uint32_t, OP_ADD>(cnt);
    if (lane == 63) {
      s_f[wid] = f;
      s_v[wid] = v;
      s_c[wid] = ci;
    }
    __syncthreads();
    uint32_t cf = s_cf, cc = s_cc;
    T cv = s_cv;
    for (uint32_t w = 0; w < wid; ++w) {
      seg_combine<T, OP>(cf, cv, s_f[w], s_v[w]);
      cc += s_c[w];
    }
    const uint32_t ef = shift_up1<uint32_t>(f, 0u);
    const T ev = shift_up1<T>(v, C::identity());
    uint32_t xf = cf;
    T xv = cv;
    seg_combine<T, OP>(xf, xv, ef, ev);
    if (i < m) {
      f_out[i] = xf;
      v_out[i] = xv;
      if (c_out) c_out[i] = cc + ci - cnt;
    }
    __syncthreads();
    if (tid == 1023) {
      seg_combine<T, OP>(cf, cv, f, v);
      s_cf = cf;
      s_cv = cv;
      s_cc = cc + ci;
    }
    __syncthreads();
  }
}

// seg Reduce in one read of the values, for every element type, one tile
// per block. The tile combines its elements from an empty carry (head-flag
// segmented scan: vectors, waves by DPP, rows x waves), writes init (+) value
// for every segment that starts and ends in it (init for empty ones), and
// leaves four records: whether it holds a segment start (f), the combination
// after its last start (v: all of it without one), the combination before its
// first start (lead), and the segment entering it if that one ends in it
// (span). seg_span_walk_kernel then completes the segments that cross tiles.
// The offsets of the tile's first 512 segments, read once to mark the starts,
// stay in registers for the result writes.
template <typename T, int OP>
__global__ __launch_bounds__(kSegBlock) void seg_reduce_kernel(
    const T* __restrict__ in, uint32_t n, const uint32_t* __restrict__ offsets, uint32_t nseg,
    T* __restrict__ out, T init, bool aligned, uint32_t ntiles,
    const uint32_t* __restrict__ tile_first, uint32_t* __restrict__ rec_f, T* __restrict__ rec_v,
    T* __restrict__ rec_lead, uint32_t* __restrict__ rec_span) {
  using C = Combine<T, OP>;
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = kSegBlock / kWave;
  constexpr int R = 4;
  constexpr uint32_t TILE = (uint32_t)kSegBlock * R * V;
  constexpr int NP = R * NW;
  static_assert(NP <= kWave, "one partial per lane");
  __shared__ uint32_t s_heads[TILE / 32];
  __shared__ uint32_t s_wf[R][NW];
  __shared__ T s_wv[R][NW];
  __shared__ T s_vals[TILE];
  __shared__ uint32_t s_more;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const T id = C::identity();
  const uint32_t t = blockIdx.x;
  const size_t base = (size_t)t * TILE;
  const uint32_t tile_end = (uint32_t)min((size_t)n, base + TILE);
  T x[R][V];
  {
    const bool full = aligned && base + TILE <= n;
#pragma unroll
    for (int r = 0; r < R; ++r)
      load_run<T, V>(in, base + ((uint32_t)r * kSegBlock + tid) * V, n, full, x[r]);
  }
  const uint32_t s0 = tile_first[t];
  for (uint32_t i = tid; i < TILE / 32; i += kSegBlock) s_heads[i] = 0;
  __syncthreads();
  // Segment starts inside the tile; starts are sorted, so only a chunk whose
  // last start is inside can be followed by more. The first chunk's start
  // and end offsets are kept.
  uint32_t st0 = n, en0 = n;
  {
    const uint32_t s = s0 + tid;
    if (s < nseg) {
      st0 = offsets[s];
      en0 = offsets[s + 1];
      if (st0 >= base && st0 < tile_end)
        atomicOr(&s_heads[(st0 - base) >> 5], 1u << ((st0 - base) & 31u));
    }
  }
  const bool more = __syncthreads_or(tid == kSegBlock - 1 && st0 < tile_end);
  if (more) {
    for (uint32_t c = kSegBlock;; c += kSegBlock) {
      const uint32_t s = s0 + c + tid;
      bool m = false;
      if (s < nseg) {
        const uint32_t st = offsets[s];
        if (st < tile_end) {
          m = tid == kSegBlock - 1;
          if (st >= base) atomicOr(&s_heads[(st - base) >> 5], 1u << ((st - base) & 31u));
        }
      }
      if (!__syncthreads_or(m)) break;
    }
  }
  uint32_t hb[R];
  uint32_t wf[R];
  T wv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t li = ((uint32_t)r * kSegBlock + tid) * V;
    hb[r] = (s_heads[li >> 5] >> (li & 31u)) & ((1u << V) - 1u);
    uint32_t f = hb[r] & 1u;
    T v = x[r][0];
#pragma unroll
    for (int k = 1; k < V; ++k) seg_combine<T, OP>(f, v, (hb[r] >> k) & 1u, x[r][k]);
    wf[r] = f;
    wv[r] = v;
    wave_inclusive_seg_scan<T, OP>(wf[r], wv[r]);
    if (lane == 63) {
      s_wf[r][wid] = wf[r];
      s_wv[r][wid] = wv[r];
    }
  }
  __syncthreads();
  // (row, wave) partials scanned by every wave, one per lane.
  uint32_t qf = 0;
  T qv = id;
  if (lane < (uint32_t)NP) {
    qf = s_wf[lane / NW][lane % NW];
    qv = s_wv[lane / NW][lane % NW];
  }
  wave_inclusive_seg_scan<T, OP>(qf, qv);
  const uint32_t af = readlane(qf, NP - 1);
  const T av = readlane(qv, NP - 1);
  const uint32_t xf = shift_up1<uint32_t>(qf, 0u);
  const T xv = shift_up1<T>(qv, id);
  const int w0 = __builtin_amdgcn_readfirstlane((int)wid);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    uint32_t cf = readlane(xf, r * NW + w0);
    T cv = readlane(xv, r * NW + w0);
    seg_combine<T, OP>(cf, cv, shift_up1<uint32_t>(wf[r], 0u), shift_up1<T>(wv[r], id));
    const uint32_t li = ((uint32_t)r * kSegBlock + tid) * V;
    T run = cv;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      run = ((hb[r] >> k) & 1u) ? x[r][k] : C::apply(run, x[r][k]);
      s_vals[li + k] = run;
    }
  }
  __syncthreads();
  // One result per segment that starts in the tile and ends in it.
  auto write = [&](uint32_t s, uint32_t st, uint32_t en) {
    if (st >= base && st < tile_end) {
      if (en == st) out[s] = init;
      else if (en <= tile_end) out[s] = C::apply(init, s_vals[en - 1 - base]);
    }
  };
  if (s0 + tid < nseg) write(s0 + tid, st0, en0);
  if (more) {
    for (uint32_t c = kSegBlock;; c += kSegBlock) {
      const uint32_t s = s0 + c + tid;
      bool m = false;
      if (s < nseg) {
        const uint32_t st = offsets[s];
        if (st < tile_end) {
          m = tid == kSegBlock - 1;
          write(s, st, offsets[s + 1]);
        }
      }
      if (!__syncthreads_or(m)) break;
    }
  }
  if (tid == 0) {
    rec_f[t] = af;
    rec_v[t] = av;
    uint32_t sp = ~0u;
    T ld = av;
    // No start at the tile's first element: segment s0 enters the tile and
    // ends at its end offset (en0 of thread 0), in this tile or later.
    if (!(s_heads[0] & 1u) && en0 <= tile_end) {
      sp = s0;
      ld = s_vals[en0 - 1 - base];
    }
    rec_span[t] = sp;
    rec_lead[t] = ld;
  }
  if (t == ntiles - 1) {
    // Trailing empty segments (start == n).
    if (tid == 0) s_more = lower_bound_u32(offsets, 0, nseg + 1, n);
    __syncthreads();
    for (uint32_t s = s_more + tid; s < nseg; s += kSegBlock) out[s] = init;
  }
}

// Completes the segments that enter a tile from before and end in it: one
// thread per such tile walks back over the earlier tiles' records, 16 per
// step, to the tile where the segment starts. Within a step the partials are
// folded right to left; the steps' results (lead[t] and v[t-1] first) are
// combined pairwise by a binary cascade (level k holds the combination of 2^k
// consecutive step results; a new result, to the left of all held ones,
// merges upward with equal-sized blocks). So a segment's result is a fixed
// tree over its partials whose depth grows with 16 + log2 of the number of
// tiles it spans.
template <typename T, int OP>
__global__ void seg_span_walk_kernel(const uint32_t* __restrict__ rec_f, const T* __restrict__ rec_v,
                                     const T* __restrict__ rec_lead,
                                     const uint32_t* __restrict__ rec_span, uint32_t ntiles,
                                     T* __restrict__ out, T init) {
  using C = Combine<T, OP>;
  constexpr int kLevels = 17;  // 32-bit lengths: at most 2^20 tiles, 2^16 steps
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t s = rec_span[t];
  if (s == ~0u) return;
  // Common case: the segment starts in the tile before (t > 0: tile 0
  // starts with a segment).
  const uint32_t f1 = rec_f[t - 1];
  const T v1 = rec_v[t - 1];
  const T r1 = C::apply(v1, rec_lead[t]);
  if (f1) {
    out[s] = C::apply(init, r1);
    return;
  }
  T lvl[kLevels];
  uint32_t have = 0;
  // Branch-free in the level (no early exit), so every level stays a register.
  auto push = [&](T x) {
    bool placed = false;
#pragma unroll
    for (int k = 0; k < kLevels; ++k) {
      const bool occupied = (have >> k) & 1u;
      if (!placed && occupied) x = C::apply(x, lvl[k]);  // x lies left of level k's block
      if (!placed && !occupied) lvl[k] = x;
      if (!placed) have ^= 1u << k;
      placed = placed || !occupied;
    }
  };
  push(r1);
  bool done = false;
  for (int64_t q = (int64_t)t - 2; !done && q >= 0; q -= 16) {
    uint32_t f[16];
    T v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      f[k] = q - k >= 0 ? rec_f[q - k] : 1u;
      v[k] = q - k >= 0 ? rec_v[q - k] : C::identity();
    }
    T acc = v[0];
    done = f[0] != 0;
#pragma unroll
    for (int k = 1; k < 16; ++k) {
      if (!done) acc = C::apply(v[k], acc);
      done = done || f[k] != 0;
    }
    push(acc);
  }
  // The held blocks, left (lowest level) to right.
  bool any = false;
  T acc = C::identity();
#pragma unroll
  for (int k = 0; k < kLevels; ++k) {
    if (have & (1u << k)) {
      acc = any ? C::apply(acc, lvl[k]) : lvl[k];
      any = true;
    }
  }
  out[s] = C::apply(init, acc);
}

// first[t] = the last segment starting before tile t (the non-empty segment
// holding element t * tile - 1), first[0] = 0: one thread per segment writes
// the tiles whose start lies in (its start, its end].
__global__ void seg_tile_first_kernel(const uint32_t* __restrict__ offsets, uint32_t nseg,
                                      uint32_t tile, uint32_t ntiles, uint32_t* __restrict__ first) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) first[0] = 0;
  if (s >= nseg) return;
  const uint32_t lo = offsets[s], hi = offsets[s + 1];
  if (lo == hi) return;
  const uint32_t t1 = min(hi / tile, ntiles - 1);
  for (uint32_t t = lo / tile + 1; t <= t1; ++t) first[t] = s;
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, uint32_t n, T v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- Segmentation ---------------------------------------------------------------
// status[0]: invalid flag (max), status[1]: last offset.
__global__ void validate_offsets_kernel(const uint32_t* __restrict__ offsets, uint32_t m,
                                        uint32_t* __restrict__ status) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t o = offsets[i];
  if ((i == 0 && o != 0) || (i > 0 && o < offsets[i - 1])) atomicMax(&status[0], 1u);
  if (i + 1 == m) status[1] = o;
}

// ---- host -------------------------------------------------------------------------
// seg Reduce over CSR offsets: one read of the values (seg_reduce_kernel),
// then the segments that cross tiles (seg_span_walk_kernel).
template <typename T, int OP>
static msv_status seg_reduce_run(Ctx& c, const T* in, uint32_t n, const uint32_t* offsets,
                                 uint32_t nseg, T* out, T init) {
  constexpr uint32_t TILE = (uint32_t)kSegBlock * 4 * (16 / sizeof(T));
  const uint32_t ntiles = div_up(n, TILE);
  uint32_t* first = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
  uint32_t* rec_f = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
  uint32_t* rec_span = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
  T* rec_v = (T*)c.scratch(sizeof(T) * ntiles);
  T* rec_lead = (T*)c.scratch(sizeof(T) * ntiles);
  if (!first || !rec_f || !rec_span || !rec_v || !rec_lead) return MSV_ERR_OUT_OF_MEMORY;
  hipLaunchKernelGGL(seg_tile_first_kernel, dim3(div_up(nseg, 256u)), dim3(256), 0, c.stream(),
                     offsets, nseg, TILE, ntiles, first);
  const bool aligned = ((uintptr_t)in & 15u) == 0;
  hipLaunchKernelGGL((seg_reduce_kernel<T, OP>), dim3(ntiles), dim3(kSegBlock), 0, c.stream(), in, n,
                     offsets, nseg, out, init, aligned, ntiles, first, rec_f, rec_v, rec_lead,
                     rec_span);
  hipLaunchKernelGGL((seg_span_walk_kernel<T, OP>), dim3(div_up(ntiles, 256u)), dim3(256), 0,
                     c.stream(), rec_f, rec_v, rec_lead, rec_span, ntiles, out, init);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

// Segmented scans / reduce over CSR offsets (HK = void) or over runs of equal
// keys (HK = the key type, ra.keys set).
template <typename T, int OP, typename HK = void>
static msv_status seg_scan_impl(Ctx& c, int outk, const T* in, uint32_t n,
                                const uint32_t* offsets, uint32_t nseg, T* out, T init,
                                RunsArgs ra = RunsArgs()) {
  constexpr bool KEYS = !std::is_same<HK, void>::value;
  const uint32_t TILE = (uint32_t)kSegBlock * seg_rows(outk) * (16 / sizeof(T));
  if (n == 0) {
    if (outk == SEG_REDUCE && nseg) {
      hipLaunchKernelGGL(fill_kernel<T>, dim3((nseg + 255) / 256), dim3(256), 0, c.stream(), out,
                         nseg, init);
      MSV_HIP(hipGetLastError());
    }
    return MSV_OK;
  }
  if constexpr (!KEYS) {
    if (outk == SEG_REDUCE) return seg_reduce_run<T, OP>(c, in, n, offsets, nseg, out, init);
  }
  const uint32_t ntiles = div_up(n, TILE);
  // Per-element outputs are stored as vectors; the others only need the
  // inputs aligned.
  const bool elem_out = outk == SEG_INCL || outk == SEG_EXCL;
  const uintptr_t vec_ptrs = (uintptr_t)in | (elem_out ? (uintptr_t)out : 0) | (uintptr_t)ra.keys;
  const bool aligned = (vec_ptrs & 15u) == 0;
#define MSV_SEG_LAUNCH(MODE, OUTK, ST, TF, TV)                                                   \
  hipLaunchKernelGGL((seg_tile_kernel<T, OP, MODE, OUTK, HK>), dim3(ntiles), dim3(kSegBlock), 0, \
                     c.stream(), in, n, offsets, nseg, out, init, aligned, ST, gen, ntiles, TF, TV, ra)
#define MSV_SEG_BY_OUT(MODE, ST, TF, TV)                                  \
  if (outk == SEG_INCL) MSV_SEG_LAUNCH(MODE, SEG_INCL, ST, TF, TV);        \
  else if (outk == SEG_EXCL) MSV_SEG_LAUNCH(MODE, SEG_EXCL, ST, TF, TV);   \
  else if constexpr (KEYS) MSV_SEG_LAUNCH(MODE, SEG_RUNS, ST, TF, TV);
  if constexpr (!KEYS) {
    if (nseg > 0 && ntiles > 1) {
      uint32_t* first = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
      if (!first) return MSV_ERR_OUT_OF_MEMORY;
      hipLaunchKernelGGL(seg_tile_first_kernel, dim3(div_up(nseg, 256u)), dim3(256), 0, c.stream(),
                         offsets, nseg, TILE, ntiles, first);
      ra.tile_first = first;
    }
  }
  uint32_t gen = 0;
  if constexpr (!IsFloat<T>::value) {
    uint64_t *counter, *status;
    const size_t words = (size_t)ntiles * TileStatus<T>::kWords;
    const size_t cwords = outk == SEG_RUNS ? (size_t)ntiles : 0;
    MSV_TRY(c.lookback(words + cwords, &counter, &status, &gen));
    if (cwords) ra.cstatus = status + words;
    MSV_SEG_BY_OUT(SEG_LOOKBACK, status, (uint32_t*)nullptr, (T*)nullptr)
  } else {
    uint32_t* af = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
    T* av = (T*)c.scratch(sizeof(T) * ntiles);
    uint32_t* pf = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
    T* pv = (T*)c.scratch(sizeof(T) * ntiles);
    uint32_t *acnt = nullptr, *pcnt = nullptr;
    if (outk == SEG_RUNS) {
      acnt = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
      pcnt = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
      if (!acnt || !pcnt) return MSV_ERR_OUT_OF_MEMORY;
    }
    if (!af || !av || !pf || !pv) return MSV_ERR_OUT_OF_MEMORY;
    RunsArgs agg = ra;
    agg.tile_c = acnt;
    // The aggregate pass computes the plain segmented aggregates (its output
    // kind only matters for the run counts).
    if (outk == SEG_RUNS) {
      if constexpr (KEYS)
        hipLaunchKernelGGL((seg_tile_kernel<T, OP, SEG_AGGREGATE, SEG_RUNS, HK>), dim3(ntiles),
                           dim3(kSegBlock), 0, c.stream(), in, n, offsets, nseg, out, init, aligned,
                           (uint64_t*)nullptr, 0u, ntiles, af, av, agg);
    } else {
      hipLaunchKernelGGL((seg_tile_kernel<T, OP, SEG_AGGREGATE, SEG_INCL, HK>), dim3(ntiles),
                         dim3(kSegBlock), 0, c.stream(), in, n, offsets, nseg, out, init, aligned,
                         (uint64_t*)nullptr, 0u, ntiles, af, av, agg);
    }
    MSV_HIP(hipGetLastError());
    hipLaunchKernelGGL((seg_pairs_exclusive_kernel<T, OP>), dim3(1), dim3(1024), 0, c.stream(),
                       af, av, ntiles, pf, pv, acnt, pcnt);
    MSV_HIP(hipGetLastError());
    ra.tile_c = pcnt;
    MSV_SEG_BY_OUT(SEG_PREFIXED, (uint64_t*)nullptr, pf, pv)
  }
#undef MSV_SEG_BY_OUT
#undef MSV_SEG_LAUNCH
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename T>
static msv_status seg_scan_by_op(Ctx& c, msv_op op, int outk, const void* in, uint32_t n,
                                 const uint32_t* offsets, uint32_t nseg, void* out,
                                 const void* init_host) {
  T init = T();
  if (init_host) memcpy(&init, init_host, sizeof(T));
  switch (op) {
    case MSV_OP_ADD: return seg_scan_impl<T, OP_ADD>(c, outk, (const T*)in, n, offsets, nseg, (T*)out, init);
    case MSV_OP_MIN: return seg_scan_impl<T, OP_MIN>(c, outk, (const T*)in, n, offsets, nseg, (T*)out, init);
    case MSV_OP_MAX: return seg_scan_impl<T, OP_MAX>(c, outk, (const T*)in, n, offsets, nseg, (T*)out, init);
    case MSV_OP_XOR:
      if constexpr (!IsFloat<T>::value)
        return seg_scan_impl<T, OP_XOR>(c, outk, (const T*)in, n, offsets, nseg, (T*)out, init);
      break;
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for this dtype", (int)op);
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status seg_scan_run(Ctx& c, msv_dtype dt, msv_op op, int outk, const void* in, uint32_t n,
                        const uint32_t* offsets, uint32_t nseg, void* out, const void* init) {
  switch (dt) {
    case MSV_U32: return seg_scan_by_op<uint32_t>(c, op, outk, in, n, offsets, nseg, out, init);
    case MSV_I32: return seg_scan_by_op<int32_t>(c, op, outk, in, n, offsets, nseg, out, init);
    case MSV_F32: return seg_scan_by_op<float>(c, op, outk, in, n, offsets, nseg, out, init);
    case MSV_U64: return seg_scan_by_op<uint64_t>(c, op, outk, in, n, offsets, nseg, out, init);
    case MSV_I64: return seg_scan_by_op<int64_t>(c, op, outk, in, n, offsets, nseg, out, init);
    case MSV_F64: return seg_scan_by_op<double>(c, op, outk, in, n, offsets, nseg, out, init);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

template <typename T, typename K>
static msv_status runs_scan_by_op(Ctx& c, msv_op op, int outk, const void* in, uint32_t n,
                                  void* out, const void* init_host, const RunsArgs& ra) {
  T init = T();
  if (init_host) memcpy(&init, init_host, sizeof(T));
  switch (op) {
    case MSV_OP_ADD: return seg_scan_impl<T, OP_ADD, K>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, ra);
    case MSV_OP_MIN: return seg_scan_impl<T, OP_MIN, K>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, ra);
    case MSV_OP_MAX: return seg_scan_impl<T, OP_MAX, K>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, ra);
    case MSV_OP_XOR:
      if constexpr (!IsFloat<T>::value)
        return seg_scan_impl<T, OP_XOR, K>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, ra);
      break;
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for this dtype", (int)op);
  return MSV_ERR_INVALID_ARGUMENT;
}

template <typename K>
static msv_status runs_scan_values(Ctx& c, msv_dtype vdt, msv_op op, int outk, const void* in,
                                   uint32_t n, void* out, const void* init, const RunsArgs& ra) {
  switch (vdt) {
    case MSV_U32: return runs_scan_by_op<uint32_t, K>(c, op, outk, in, n, out, init, ra);
    case MSV_I32: return runs_scan_by_op<int32_t, K>(c, op, outk, in, n, out, init, ra);
    case MSV_F32: return runs_scan_by_op<float, K>(c, op, outk, in, n, out, init, ra);
    case MSV_U64: return runs_scan_by_op<uint64_t, K>(c, op, outk, in, n, out, init, ra);
    case MSV_I64: return runs_scan_by_op<int64_t, K>(c, op, outk, in, n, out, init, ra);
    case MSV_F64: return runs_scan_by_op<double, K>(c, op, outk, in, n, out, init, ra);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status runs_scan_run(Ctx& c, msv_dtype kdt, const void* keys, msv_dtype vdt, msv_op op,
                         int kind, const void* in, uint32_t n, void* out, const void* init,
                         void* out_keys, uint32_t* len_keys, uint32_t* len_vals) {
  const int outk = kind == RUNS_INCLUSIVE ? SEG_INCL : kind == RUNS_EXCLUSIVE ? SEG_EXCL : SEG_RUNS;
  RunsArgs ra;
  ra.keys = keys;
  ra.out_keys = out_keys;
  ra.len_keys = len_keys;
  ra.len_vals = len_vals;
  switch (kdt) {
    case MSV_U32: return runs_scan_values<uint32_t>(c, vdt, op, outk, in, n, out, init, ra);
    case MSV_I32: return runs_scan_values<int32_t>(c, vdt, op, outk, in, n, out, init, ra);
    case MSV_F32: return runs_scan_values<float>(c, vdt, op, outk, in, n, out, init, ra);
    case MSV_U64: return runs_scan_values<uint64_t>(c, vdt, op, outk, in, n, out, init, ra);
    case MSV_I64: return runs_scan_values<int64_t>(c, vdt, op, outk, in, n, out, init, ra);
    case MSV_F64: return runs_scan_values<double>(c, vdt, op, outk, in, n, out, init, ra);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

}  // namespace msv

using namespace msv;

static msv_status seg_check(msv_executor* ex, const msv_segmentation* seg, const View& v) {
  if (!seg || !seg->offsets) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null segmentation");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (seg->offsets->owner_id != ex->id) {
    set_error(MSV_ERR_FOREIGN_EXECUTOR, 0, 0, "executor does not own this segmentation");
    return MSV_ERR_FOREIGN_EXECUTOR;
  }
  if (seg->value_count != v.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, seg->value_count, v.len, "input lengths differ: %u != %u",
              seg->value_count, v.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  return MSV_OK;
}

static bool op_ok(msv_dtype dt, int op) {
  if (op < MSV_OP_ADD || op > MSV_OP_XOR) return false;
  return !(op == MSV_OP_XOR && dtype_is_float(dt));
}

static msv_status seg_scan_api(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                               const void* init, msv_op op, int outk, msv_vec** out) {
  clear_error();
  View v;
  MSV_TRY(resolve(ex, values, &v));
  MSV_TRY(seg_check(ex, seg, v));
  if (!op_ok(v.dtype, op) || (outk != SEG_INCL && !init)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad operator or null init");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  const uint32_t cap = outk == SEG_REDUCE ? seg->segment_count : v.len;
  MSV_TRY(new_vec(ex, v.dtype, cap, false, out));
  Ctx c(ex);
  msv_status s = seg_scan_run(c, v.dtype, op, outk, v.ptr, v.len, (const uint32_t*)seg->offsets->ptr,
                              seg->segment_count, (*out)->ptr, init);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

extern "C" {

msv_status msv_segmentation_from_offsets(msv_executor* ex, msv_slice offsets,
                                         msv_segmentation** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "Segmentation::from_offsets"));
  View o;
  MSV_TRY(resolve(ex, offsets, &o));
  if (!out || o.dtype != MSV_U32) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "offsets must be u32 (MIndex)");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (o.len == 0) {
    set_error(MSV_ERR_INVALID_SEGMENTATION, 0, 0, "segmentation is not a valid ordered partition");
    return MSV_ERR_INVALID_SEGMENTATION;
  }
  msv_vec* copy = nullptr;
  MSV_TRY(new_vec(ex, MSV_U32, o.len, false, &copy));
  Ctx c(ex);
  uint32_t* status = (uint32_t*)c.scratch(8);
  if (!status) {
    msv_vec_release(copy);
    return MSV_ERR_OUT_OF_MEMORY;
  }
  uint32_t hs[2] = {0, 0};
  hipError_t e = hipMemcpyAsync(copy->ptr, o.ptr, (size_t)o.len * 4, hipMemcpyDeviceToDevice, ex->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(status, hs, 8, hipMemcpyHostToDevice, ex->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(validate_offsets_kernel, dim3((o.len + 255) / 256), dim3(256), 0, ex->stream,
                       (const uint32_t*)copy->ptr, o.len, status);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hs, status, 8, hipMemcpyDeviceToHost, ex->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ex->stream);
  if (e != hipSuccess) {
    msv_vec_release(copy);
    set_error(MSV_ERR_LAUNCH, 0, 0, "segmentation validation failed: %s", hipGetErrorString(e));
    return MSV_ERR_LAUNCH;
  }
  if (hs[0] != 0) {
    msv_vec_release(copy);
    set_error(MSV_ERR_INVALID_SEGMENTATION, 0, 0, "segmentation is not a valid ordered partition");
    return MSV_ERR_INVALID_SEGMENTATION;
  }
  msv_segmentation* sg = new msv_segmentation;
  sg->offsets = copy;
  sg->segment_count = o.len - 1;
  sg->value_count = hs[1];
  *out = sg;
  return MSV_OK;
}

msv_status msv_segmentation_from_lengths(msv_executor* ex, msv_slice lengths,
                                         msv_segmentation** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "Segmentation::from_lengths"));
  View l;
  MSV_TRY(resolve(ex, lengths, &l));
  if (!out || l.dtype != MSV_U32) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "lengths must be u32 (MIndex)");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (l.len == 0xffffffffu) {
    set_error(MSV_ERR_LENGTH_TOO_LARGE, 0x100000000ull, 0, "length does not fit in u32");
    return MSV_ERR_LENGTH_TOO_LARGE;
  }
  msv_vec* offs = nullptr;
  MSV_TRY(new_vec(ex, MSV_U32, l.len + 1, false, &offs));
  Ctx c(ex);
  uint32_t* status = (uint32_t*)c.scratch(8);
  uint32_t hs[2] = {0, 0};
  msv_status s = status ? MSV_OK : MSV_ERR_OUT_OF_MEMORY;
  hipError_t e = hipSuccess;
  if (s == MSV_OK) e = hipMemsetAsync(offs->ptr, 0, 4, ex->stream);
  if (s == MSV_OK && e == hipSuccess)
    s = scan_run(c, MSV_U32, MSV_OP_ADD, false, l.ptr, (uint32_t*)offs->ptr + 1, l.len, nullptr);
  if (s == MSV_OK && e == hipSuccess) e = hipMemcpyAsync(status, hs, 8, hipMemcpyHostToDevice, ex->stream);
  if (s == MSV_OK && e == hipSuccess) {
    // A wrapped running sum decreases at the wrap: validation flags it.
    hipLaunchKernelGGL(validate_offsets_kernel, dim3((l.len + 1 + 255) / 256), dim3(256), 0,
                       ex->stream, (const uint32_t*)offs->ptr, l.len + 1, status);
    e = hipGetLastError();
  }
  if (s == MSV_OK && e == hipSuccess) e = hipMemcpyAsync(hs, status, 8, hipMemcpyDeviceToHost, ex->stream);
  if (s == MSV_OK && e == hipSuccess) e = hipStreamSynchronize(ex->stream);
  if (s != MSV_OK || e != hipSuccess) {
    msv_vec_release(offs);
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "offsets from lengths failed: %s", hipGetErrorString(e));
      return MSV_ERR_LAUNCH;
    }
    return s;
  }
  if (hs[0] != 0) {
    msv_vec_release(offs);
    set_error(MSV_ERR_LENGTH_TOO_LARGE, 0x100000000ull, 0, "length does not fit in u32: %llu",
              0x100000000ull);
    return MSV_ERR_LENGTH_TOO_LARGE;
  }
  msv_segmentation* sg = new msv_segmentation;
  sg->offsets = offs;
  sg->segment_count = l.len;
  sg->value_count = hs[1];
  *out = sg;
  return MSV_OK;
}

msv_status msv_segmentation_destroy(msv_segmentation* s) {
  if (!s) return MSV_ERR_INVALID_ARGUMENT;
  msv_vec_release(s->offsets);
  delete s;
  return MSV_OK;
}

msv_status msv_segmentation_offsets(msv_segmentation* s, msv_vec** out) {
  clear_error();
  if (!s || !out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null segmentation or output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  msv_vec_retain(s->offsets);
  *out = s->offsets;
  return MSV_OK;
}

uint32_t msv_segmentation_segment_count(const msv_segmentation* s) { return s ? s->segment_count : 0; }
uint32_t msv_segmentation_value_count(const msv_segmentation* s) { return s ? s->value_count : 0; }

msv_status msv_seg_inclusive_scan(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                                  msv_op op, msv_vec** out) {
  return seg_scan_api(ex, values, seg, nullptr, op, SEG_INCL, out);
}

msv_status msv_seg_exclusive_scan(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                                  const void* init, msv_op op, msv_vec** out) {
  return seg_scan_api(ex, values, seg, init, op, SEG_EXCL, out);
}

msv_status msv_seg_reduce(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                          const void* init, msv_op op, msv_vec** out) {
  return seg_scan_api(ex, values, seg, init, op, SEG_REDUCE, out);
}

msv_status msv_seg_sort(msv_executor* ex, msv_slice keys, msv_slice values,
                        const msv_segmentation* seg, msv_cmp cmp, msv_vec** out_keys,
                        msv_vec** out_values) {
  clear_error();
  if (cmp != MSV_CMP_LESS && cmp != MSV_CMP_GREATER) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown comparator %d", (int)cmp);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  MSV_TRY(seg_check(ex, seg, k));
  const bool has_v = values.vec != nullptr;
  if (has_v) {
    MSV_TRY(resolve(ex, values, &v));
    if (v.len != k.len) {
      set_error(MSV_ERR_LENGTH_MISMATCH, k.len, v.len, "input lengths differ: %u != %u", k.len, v.len);
      return MSV_ERR_LENGTH_MISMATCH;
    }
    if (!out_values) {
      set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null value output");
      return MSV_ERR_INVALID_ARGUMENT;
    }
  }
  if (!out_keys) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null key output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, k.dtype, k.len, false, out_keys));
  msv_status s = MSV_OK;
  if (has_v) s = new_vec(ex, v.dtype, v.len, false, out_values);
  if (s == MSV_OK) {
    Ctx c(ex);
    s = seg_sort_run(c, k.dtype, cmp == MSV_CMP_LESS ? 1 : 2, k.ptr, has_v ? v.ptr : nullptr,
                     has_v ? dtype_size(v.dtype) : 0, k.len, (const uint32_t*)seg->offsets->ptr,
                     seg->segment_count, (*out_keys)->ptr, has_v ? (*out_values)->ptr : nullptr);
  }
  if (s != MSV_OK) {
    msv_vec_release(*out_keys);
    *out_keys = nullptr;
    if (has_v && *out_values) {
      msv_vec_release(*out_values);
      *out_values = nullptr;
    }
  }
  return s;
}

}  // extern "C"
