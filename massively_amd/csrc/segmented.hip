// seg:: (ForEachSegment InclusiveScan / ExclusiveScan / Reduce / Sort) and the
// Segmentation constructors for gfx950.
//
// Reference: massively/src/seg/mod.rs:379-425 (Sort), :458-572 (scans),
// :676-806, 1314-1336 (Reduce, empty segments -> init), seg/control.rs
// (offsets -> head flags), core/segmented_fixed.rs (head-flag segmented scan),
// seg/segmentation.rs:10-52, 187-264 (validation, offsets from lengths).
//
// Scans and reductions run over the flat value array in tiles of 256 threads
// x 4 rows x one 16-byte vector (4096 u32 / 2048 u64). Each tile finds its
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

constexpr int kSegBlock = 256;
constexpr int kSegRows = 8;

enum SegOut : int { SEG_INCL = 0, SEG_EXCL = 1, SEG_REDUCE = 2 };
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

// First index in [lo, hi) with a[i] >= v (hi if none), for a nondecreasing
// array, searched by one whole wave: each round samples 64 evenly spaced
// points and keeps the gap that holds the boundary, so a search over 2^20
// entries takes 4 dependent loads instead of 20.
__device__ inline uint32_t wave_lower_bound(const uint32_t* a, uint32_t lo, uint32_t hi,
                                            uint32_t v) {
  const uint32_t lane = __lane_id();
  while (hi - lo > (uint32_t)kWave) {
    const uint32_t step = (hi - lo + kWave - 1) / kWave;
    const uint32_t idx = lo + lane * step;
    const bool below = idx < hi && a[idx] < v;
    const uint32_t cnt = (uint32_t)__builtin_popcountll(ballot(below));
    // Points lo + k*step for k < cnt are below v; the answer lies after the
    // last of them and at or before the next sampled point.
    const uint32_t nlo = cnt ? lo + (cnt - 1) * step + 1 : lo;
    const uint32_t nhi = min(hi, lo + cnt * step);
    lo = nlo;
    hi = nhi;
  }
  const uint32_t idx = lo + lane;
  const bool below = idx < hi && a[idx] < v;
  return lo + (uint32_t)__builtin_popcountll(ballot(below));
}

// MASK: segment heads come from a bitmask (bit i set = element i starts a
// segment; the by-key scans build it from adjacent key comparisons) instead
// of CSR offsets.
template <typename T, int OP, int MODE, int OUT, bool MASK = false>
__global__ __launch_bounds__(kSegBlock) void seg_tile_kernel(
    const T* __restrict__ in, uint32_t n, const uint32_t* __restrict__ offsets, uint32_t nseg,
    T* __restrict__ out, T init, bool aligned, uint64_t* counter, uint64_t* status,
    uint32_t gen, uint32_t ntiles, uint32_t* __restrict__ tile_f, T* __restrict__ tile_v,
    const uint32_t* __restrict__ hmask) {
  using C = Combine<T, OP>;
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = kSegBlock / kWave;
  constexpr uint32_t TILE = (uint32_t)kSegBlock * kSegRows * V;
  __shared__ uint32_t s_heads[TILE / 32];
  __shared__ uint32_t s_wf[kSegRows][NW];
  __shared__ T s_wv[kSegRows][NW];
  __shared__ uint32_t s_cf;
  __shared__ T s_cv;
  __shared__ uint32_t s_first;
  __shared__ T s_vals[OUT == SEG_REDUCE ? TILE : 1];

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  (void)counter;
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * TILE;
  const uint32_t tile_end = (uint32_t)min((size_t)n, base + TILE);
  uint32_t s0 = 0;
  if constexpr (MASK) {
    // TILE is a multiple of 32, so the tile's heads are whole mask words.
    const uint32_t w0 = (uint32_t)(base >> 5), nw = (uint32_t)((n + 31) / 32);
    for (uint32_t i = tid; i < TILE / 32; i += kSegBlock) s_heads[i] = w0 + i < nw ? hmask[w0 + i] : 0u;
    __syncthreads();
  } else {
    for (uint32_t i = tid; i < TILE / 32; i += kSegBlock) s_heads[i] = 0;
    if (wid == 0) {
      const uint32_t lb = wave_lower_bound(offsets, 0, nseg + 1, (uint32_t)base);
      if (lane == 0) s_first = lb > 0 ? lb - 1 : 0;
    }
    __syncthreads();
    s0 = s_first;
    // Mark the segment starts that fall inside the tile.
    for (uint32_t c = 0;; c += kSegBlock) {
      const uint32_t s = s0 + c + tid;
      bool more = false;
      if (s < nseg) {
        const uint32_t st = offsets[s];
        if (st < tile_end) {
          more = true;
          if (st >= base) atomicOr(&s_heads[(st - base) >> 5], 1u << ((st - base) & 31u));
        }
      }
      if (!__syncthreads_or(more)) break;
    }
  }

  const bool vec_ok = aligned && (base + TILE <= n);
  const T id = C::identity();
  T x[kSegRows][V];
  uint32_t h[kSegRows][V];
#pragma unroll
  for (int r = 0; r < kSegRows; ++r) {
    const uint32_t li = ((uint32_t)r * kSegBlock + tid) * V;
    if (vec_ok) {
      struct alignas(16) Q { T v[V]; };
      const Q q = *reinterpret_cast<const Q*>(in + base + li);
#pragma unroll
      for (int k = 0; k < V; ++k) x[r][k] = q.v[k];
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) x[r][k] = (base + li + k < n) ? in[base + li + k] : id;
    }
#pragma unroll
    for (int k = 0; k < V; ++k) h[r][k] = (s_heads[(li + k) >> 5] >> ((li + k) & 31u)) & 1u;
  }

  // Vector-local segmented inclusive scan and wave scans of vector totals.
  T inc[kSegRows][V];
  uint32_t wf[kSegRows];
  T wv[kSegRows];
#pragma unroll
  for (int r = 0; r < kSegRows; ++r) {
    uint32_t f = h[r][0];
    T v = x[r][0];
    inc[r][0] = v;
#pragma unroll
    for (int k = 1; k < V; ++k) {
      seg_combine<T, OP>(f, v, h[r][k], x[r][k]);
      inc[r][k] = v;
    }
    wf[r] = f;
    wv[r] = v;
    wave_inclusive_seg_scan<T, OP>(wf[r], wv[r]);
    if (lane == 63) {
      s_wf[r][wid] = wf[r];
      s_wv[r][wid] = wv[r];
    }
  }
  __syncthreads();
  // Exclusive carry of (row r, wave wid) and the tile aggregate.
  uint32_t pf[kSegRows];
  T pv[kSegRows];
  uint32_t af = 0;
  T av = id;
#pragma unroll
  for (int r = 0; r < kSegRows; ++r)
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if (w == (int)wid) {
        pf[r] = af;
        pv[r] = av;
      }
      seg_combine<T, OP>(af, av, s_wf[r][w], s_wv[r][w]);
    }

  if constexpr (MODE == SEG_AGGREGATE) {
    if (tid == 0) {
      tile_f[tile] = af;
      tile_v[tile] = av;
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
        // Self-help: the segmented aggregate of tile q from the input (the
        // elements from its last segment start, or all of it if none).
        auto agg_of = [&](uint32_t q, uint32_t& hf, T& a) {
          const size_t qb = (size_t)q * TILE;
          const uint32_t qe = (uint32_t)min((size_t)n, qb + TILE);
          size_t from = qb;
          hf = 0;
          if constexpr (MASK) {
            // Last set bit of the tile's mask words, searched from the end.
            const uint32_t w0 = (uint32_t)(qb >> 5), w1 = (qe + 31) / 32;
            for (int64_t wb = (int64_t)w1 - kWave; wb + kWave > (int64_t)w0; wb -= kWave) {
              const int64_t w = wb + (int64_t)lane;
              const uint32_t word = (w >= (int64_t)w0 && w < (int64_t)w1) ? hmask[w] : 0u;
              const uint64_t any = ballot(word != 0);
              if (any) {
                const uint32_t hl = 63u - (uint32_t)__builtin_clzll(any);
                const uint32_t hw = readlane(word, (int)hl);
                from = (size_t)(wb + hl) * 32 + (31u - (uint32_t)__builtin_clz(hw));
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
  } else {
    if (tid == 0) {
      s_cf = tile > 0 ? tile_f[tile] : 0u;
      s_cv = tile > 0 ? tile_v[tile] : id;
    }
  }
  __syncthreads();

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
    T y[V];
    if constexpr (OUT == SEG_EXCL) {
      T run = cv;  // segmented inclusive value of the element before
#pragma unroll
      for (int k = 0; k < V; ++k) {
        y[k] = h[r][k] ? init : C::apply(init, run);
        run = h[r][k] ? x[r][k] : C::apply(run, x[r][k]);
      }
    } else {
      // Elements before the vector's first head continue the carried segment.
      uint32_t seen = 0;
#pragma unroll
      for (int k = 0; k < V; ++k) {
        seen |= h[r][k];
        y[k] = seen ? inc[r][k] : C::apply(cv, inc[r][k]);
      }
    }
    if constexpr (OUT == SEG_REDUCE) {
#pragma unroll
      for (int k = 0; k < V; ++k) s_vals[li + k] = y[k];
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

  if constexpr (OUT == SEG_REDUCE) {
    __syncthreads();
    // One result per segment whose last element lies in this tile.
    for (uint32_t c = 0;; c += kSegBlock) {
      const uint32_t s = s0 + c + tid;
      bool more = false;
      if (s < nseg) {
        const uint32_t st = offsets[s];
        if (st < tile_end) {
          more = true;
          const uint32_t en = offsets[s + 1];
          if (en == st) {
            if (st >= base) out[s] = init;
          } else if (en - 1 >= base && en - 1 < tile_end) {
            out[s] = C::apply(init, s_vals[en - 1 - base]);
          }
        }
      }
      if (!__syncthreads_or(more)) break;
    }
    // Trailing empty segments (start == n) belong to the last tile.
    if (tile == ntiles - 1) {
      if (tid == 0) s_first = lower_bound_u32(offsets, 0, nseg + 1, n);
      __syncthreads();
      for (uint32_t s = s_first + tid; s < nseg; s += kSegBlock) out[s] = init;
    }
  }
}

// One-block exclusive segmented scan of tile aggregates (float path).
template <typename T, int OP>
__global__ __launch_bounds__(1024) void seg_pairs_exclusive_kernel(const uint32_t* __restrict__ f_in,
                                                                   const T* __restrict__ v_in,
                                                                   uint32_t m,
                                                                   uint32_t* __restrict__ f_out,
                                                                   T* __restrict__ v_out) {
  using C = Combine<T, OP>;
  constexpr int NW = 1024 / kWave;
  __shared__ uint32_t s_f[NW];
  __shared__ T s_v[NW];
  __shared__ uint32_t s_cf;
  __shared__ T s_cv;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  if (tid == 0) {
    s_cf = 0;
    s_cv = C::identity();
  }
  __syncthreads();
  for (uint32_t base = 0; base < m; base += 1024) {
    const uint32_t i = base + tid;
    uint32_t f = i < m ? f_in[i] : 0u;
    T v = i < m ? v_in[i] : C::identity();
    wave_inclusive_seg_scan<T, OP>(f, v);
    if (lane == 63) {
      s_f[wid] = f;
      s_v[wid] = v;
    }
    __syncthreads();
    uint32_t cf = s_cf;
    T cv = s_cv;
    for (uint32_t w = 0; w < wid; ++w) seg_combine<T, OP>(cf, cv, s_f[w], s_v[w]);
    const uint32_t ef = shift_up1<uint32_t>(f, 0u);
    const T ev = shift_up1<T>(v, C::identity());
    uint32_t xf = cf;
    T xv = cv;
    seg_combine<T, OP>(xf, xv, ef, ev);
    if (i < m) {
      f_out[i] = xf;
      v_out[i] = xv;
    }
    __syncthreads();
    if (tid == 1023) {
      seg_combine<T, OP>(cf, cv, f, v);
      s_cf = cf;
      s_cv = cv;
    }
    __syncthreads();
  }
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
template <typename T, int OP, bool MASK = false>
static msv_status seg_scan_impl(Ctx& c, int outk, const T* in, uint32_t n,
                                const uint32_t* offsets, uint32_t nseg, T* out, T init,
                                const uint32_t* hmask = nullptr) {
  constexpr uint32_t TILE = (uint32_t)kSegBlock * kSegRows * (16 / sizeof(T));
  if (n == 0) {
    if (outk == SEG_REDUCE && nseg) {
      hipLaunchKernelGGL(fill_kernel<T>, dim3((nseg + 255) / 256), dim3(256), 0, c.stream(), out,
                         nseg, init);
      MSV_HIP(hipGetLastError());
    }
    return MSV_OK;
  }
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  // Reduce writes per segment, so only the input needs 16-byte alignment.
  const uintptr_t vec_ptrs = outk == SEG_REDUCE ? (uintptr_t)in : ((uintptr_t)in | (uintptr_t)out);
  const bool aligned = (vec_ptrs & 15u) == 0;
#define MSV_SEG_LAUNCH(MODE, OUTK, CNT, ST, TF, TV)                                               \
  hipLaunchKernelGGL((seg_tile_kernel<T, OP, MODE, OUTK, MASK>), dim3(ntiles), dim3(kSegBlock), 0, \
                     c.stream(), in, n, offsets, nseg, out, init, aligned, CNT, ST, gen, ntiles, TF, TV, \
                     hmask)
#define MSV_SEG_BY_OUT(MODE, CNT, ST, TF, TV)                  \
  if (outk == SEG_INCL) MSV_SEG_LAUNCH(MODE, SEG_INCL, CNT, ST, TF, TV);        \
  else if (outk == SEG_EXCL) MSV_SEG_LAUNCH(MODE, SEG_EXCL, CNT, ST, TF, TV);   \
  else if constexpr (!MASK) MSV_SEG_LAUNCH(MODE, SEG_REDUCE, CNT, ST, TF, TV);
  uint32_t gen = 0;
  if constexpr (!IsFloat<T>::value) {
    uint64_t *counter, *status;
    MSV_TRY(c.lookback((size_t)ntiles * TileStatus<T>::kWords, &counter, &status, &gen));
    MSV_SEG_BY_OUT(SEG_LOOKBACK, counter, status, (uint32_t*)nullptr, (T*)nullptr)
  } else {
    uint32_t* af = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
    T* av = (T*)c.scratch(sizeof(T) * ntiles);
    uint32_t* pf = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
    T* pv = (T*)c.scratch(sizeof(T) * ntiles);
    if (!af || !av || !pf || !pv) return MSV_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL((seg_tile_kernel<T, OP, SEG_AGGREGATE, SEG_INCL, MASK>), dim3(ntiles),
                       dim3(kSegBlock), 0, c.stream(), in, n, offsets, nseg, out, init, aligned,
                       (uint64_t*)nullptr, (uint64_t*)nullptr, 0u, ntiles, af, av, hmask);
    MSV_HIP(hipGetLastError());
    hipLaunchKernelGGL((seg_pairs_exclusive_kernel<T, OP>), dim3(1), dim3(1024), 0, c.stream(),
                       af, av, ntiles, pf, pv);
    MSV_HIP(hipGetLastError());
    MSV_SEG_BY_OUT(SEG_PREFIXED, (uint64_t*)nullptr, (uint64_t*)nullptr, pf, pv)
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

template <typename T>
static msv_status mask_scan_by_op(Ctx& c, msv_op op, int outk, const void* in, uint32_t n,
                                  const uint32_t* hmask, void* out, const void* init_host) {
  T init = T();
  if (init_host) memcpy(&init, init_host, sizeof(T));
  switch (op) {
    case MSV_OP_ADD: return seg_scan_impl<T, OP_ADD, true>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, hmask);
    case MSV_OP_MIN: return seg_scan_impl<T, OP_MIN, true>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, hmask);
    case MSV_OP_MAX: return seg_scan_impl<T, OP_MAX, true>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, hmask);
    case MSV_OP_XOR:
      if constexpr (!IsFloat<T>::value)
        return seg_scan_impl<T, OP_XOR, true>(c, outk, (const T*)in, n, nullptr, 0, (T*)out, init, hmask);
      break;
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for this dtype", (int)op);
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status mask_scan_run(Ctx& c, msv_dtype dt, msv_op op, bool exclusive, const void* in,
                         uint32_t n, const uint32_t* hmask, void* out, const void* init) {
  const int outk = exclusive ? SEG_EXCL : SEG_INCL;
  switch (dt) {
    case MSV_U32: return mask_scan_by_op<uint32_t>(c, op, outk, in, n, hmask, out, init);
    case MSV_I32: return mask_scan_by_op<int32_t>(c, op, outk, in, n, hmask, out, init);
    case MSV_F32: return mask_scan_by_op<float>(c, op, outk, in, n, hmask, out, init);
    case MSV_U64: return mask_scan_by_op<uint64_t>(c, op, outk, in, n, hmask, out, init);
    case MSV_I64: return mask_scan_by_op<int64_t>(c, op, outk, in, n, hmask, out, init);
    case MSV_F64: return mask_scan_by_op<double>(c, op, outk, in, n, hmask, out, init);
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
