// Scan and reduce tile kernels (device code), shared by scan.hip and the
// tuning probes. See scan.hip for the algorithm description.
#pragma once

#include "device_common.hpp"

namespace msv {

constexpr int kScanBlock = 256;
constexpr int kScanRowsSmall = 4;   // 16 KiB per tile (4096 u32)
// Long inputs use 32 / 16 rows (amortises the lookback): 32768 u32 (128 KiB) /
// 8192 u64 (64 KiB) per tile.
template <typename T>
constexpr int scan_rows_large() { return sizeof(T) == 4 ? 32 : 16; }
constexpr int kReduceRows = 16;
// Reduce blocks: 512 threads x 16 rows (32768 u32 per tile). On MI355X this
// read-only pass ran at 0.68 of HBM peak against 0.63 with 256-thread blocks.
constexpr int kReduceBlock = 512;
// Inputs at least this long use the large tile (>= 256 tiles).
constexpr uint32_t kScanLargeMin = 1u << 22;

template <typename T>
struct alignas(16) Vec16 {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
};

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// Streaming 16-byte access; NT marks the line as used once (non-temporal).
template <bool NT>
__device__ inline u32x4_t ld16(const void* p) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
  else return *reinterpret_cast<const u32x4_t*>(p);
}
template <bool NT>
__device__ inline void st16(void* p, u32x4_t v) {
  if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<u32x4_t*>(p));
  else *reinterpret_cast<u32x4_t*>(p) = v;
}

template <typename T, int BLOCK, int ROWS, bool NT = false>
__device__ inline void load_tile(const T* __restrict__ in, size_t base, uint32_t n, bool vec_ok,
                                 T (&x)[ROWS][16 / sizeof(T)], T fill) {
  constexpr int V = 16 / sizeof(T);
  const uint32_t tid = threadIdx.x;
  if (vec_ok) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const u32x4_t q = ld16<NT>(in + base + ((size_t)r * BLOCK + tid) * V);
      const Vec16<T> w = bit_cast<Vec16<T>>(q);
#pragma unroll
      for (int k = 0; k < V; ++k) x[r][k] = w.v[k];
    }
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const size_t i = base + ((size_t)r * BLOCK + tid) * V + k;
        x[r][k] = (i < n) ? in[i] : fill;
      }
    }
  }
}

// A lazy source: element i is in(i) (the same thread-to-element layout).
// A full tile reads V consecutive elements per row at once (in.load<V>).
template <typename T, int BLOCK, int ROWS, bool NT = false, typename Src,
          typename = std::enable_if_t<is_lazy_src<Src>>>
__device__ inline void load_tile(const Src& in, size_t base, uint32_t n, bool vec_ok,
                                 T (&x)[ROWS][16 / sizeof(T)], T fill) {
  constexpr int V = 16 / sizeof(T);
  const uint32_t tid = threadIdx.x;
  if (vec_ok) {  // the kernel has checked src_tile_ok over the tile
#pragma unroll
    for (int r = 0; r < ROWS; ++r) in.template load_fast<V>(base + ((size_t)r * BLOCK + tid) * V, x[r]);
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const size_t i = base + ((size_t)r * BLOCK + tid) * V + k;
        x[r][k] = (i < n) ? (T)in(i) : fill;
      }
    }
  }
}

template <typename T, int BLOCK, int ROWS, bool NT = false>
__device__ inline void store_tile(T* __restrict__ out, size_t base, uint32_t n, bool vec_ok,
                                  const T (&y)[ROWS][16 / sizeof(T)]) {
  constexpr int V = 16 / sizeof(T);
  const uint32_t tid = threadIdx.x;
  if (vec_ok) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      Vec16<T> q;
#pragma unroll
      for (int k = 0; k < V; ++k) q.v[k] = y[r][k];
      st16<NT>(out + base + ((size_t)r * BLOCK + tid) * V, bit_cast<u32x4_t>(q));
    }
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const size_t i = base + ((size_t)r * BLOCK + tid) * V + k;
        if (i < n) out[i] = y[r][k];
      }
    }
  }
}

// Aggregate of tile q computed by one wave straight from the input (the
// lookback's self-help path).
template <typename T, int OP, uint32_t TILE, typename Src>
__device__ inline T wave_tile_fold(const Src& in, uint32_t n, uint32_t q) {
  using C = Combine<T, OP>;
  const size_t b = (size_t)q * TILE;
  const size_t e = min((size_t)n, b + TILE);
  T acc = C::identity();
  for (size_t i = b + __lane_id(); i < e; i += kWave) acc = C::apply(acc, (T)src_at(in, i));
  return wave_reduce<T, OP>(acc);
}

enum ScanMode : int { SCAN_LOOKBACK = 0, SCAN_PREFIXED = 1 };

// One tile per block, tile = blockIdx.x. SCAN_LOOKBACK resolves the tile
// prefix by decoupled lookback (with self-help); SCAN_PREFIXED uses
// `prefixes[tile]` (or the identity when prefixes is null). Src: the input
// source (a pointer, or a lazy source compiled at run time).
template <typename T, int OP, bool EXCL, int MODE, int BLOCK, int ROWS,
          typename Src = const T* __restrict__>
__global__ __launch_bounds__(BLOCK) void scan_tile_kernel(Src in,
                                                          T* __restrict__ out, uint32_t n, T init,
                                                          bool aligned, uint64_t* status,
                                                          uint32_t gen,
                                                          const T* __restrict__ prefixes) {
  using C = Combine<T, OP>;
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * ROWS * V;
  constexpr bool NT = MODE == SCAN_LOOKBACK;
  __shared__ T s_wave[ROWS][NW];
  __shared__ T s_prefix;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (kWave - 1);
  const uint32_t wid = tid / kWave;
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * TILE;
  const bool vec_ok = aligned && (base + TILE <= n) && src_tile_ok(in, base, base + TILE);
  const T id = C::identity();

  T x[ROWS][V];
  load_tile<T, BLOCK, ROWS, NT>(in, base, n, vec_ok, x, id);

  // Vector-local inclusive scans; wave totals of the vector totals to LDS.
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
#pragma unroll
    for (int k = 1; k < V; ++k) x[r][k] = C::apply(x[r][k - 1], x[r][k]);
    const T wt = wave_reduce<T, OP>(x[r][V - 1]);
    if (lane == 0) s_wave[r][wid] = wt;
  }
  __syncthreads();

  T agg = id;  // tile total
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int w = 0; w < NW; ++w) agg = C::apply(agg, s_wave[r][w]);

  if constexpr (MODE == SCAN_LOOKBACK) {
    if (wid == 0) {
      TileStatus<T> st{status, gen};
      T excl = id;
      if (tile == 0) {
        if (lane == 0) st.publish(0, kFlagInc, 0, agg);
      } else {
        if (lane == 0) st.publish(tile, kFlagAgg, 0, agg);
        excl = wave_lookback_sh<T, OP>(
            st, tile, [&](uint32_t q) { return wave_tile_fold<T, OP, TILE, Src>(in, n, q); });
        if (lane == 0) st.publish(tile, kFlagInc, 0, C::apply(excl, agg));
      }
      if (lane == 0) s_prefix = excl;
    }
  } else {
    if (tid == 0) s_prefix = prefixes ? prefixes[tile] : id;
  }
  __syncthreads();

  T tp = s_prefix;
  if constexpr (EXCL) tp = C::apply(init, tp);

  // The wave scans and the (row, wave) prefixes are recomputed here rather
  // than kept live across the lookback: only the tile itself stays in VGPRs.
  T run = id;
  // Outputs are formed in place (keeps the register footprint to one tile).
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    T pre = run;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if (w == (int)wid) pre = run;
      run = C::apply(run, s_wave[r][w]);
    }
    const T wex = shift_up1<T>(wave_inclusive_scan<T, OP>(x[r][V - 1]), id);
    const T p = C::apply(C::apply(tp, pre), wex);
    if constexpr (EXCL) {
#pragma unroll
      for (int k = V - 1; k > 0; --k) x[r][k] = C::apply(p, x[r][k - 1]);
      x[r][0] = p;
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) x[r][k] = C::apply(p, x[r][k]);
    }
  }
  store_tile<T, BLOCK, ROWS, NT>(out, base, n, vec_ok, x);
}

// Pairwise tree over a register array (depth log2(N)).
template <typename T, int OP, int N>
__device__ inline T tree_fold(T (&v)[N]) {
  using C = Combine<T, OP>;
#pragma unroll
  for (int s = 1; s < N; s *= 2) {
#pragma unroll
    for (int i = 0; i + s < N; i += 2 * s) v[i] = C::apply(v[i], v[i + s]);
  }
  return v[0];
}

// Reduces one tile of BLOCK*ROWS vectors per block into partials[blockIdx.x]
// with a fixed tree: thread (pairwise over ROWS*V) -> wave (DPP) -> block
// (pairwise over waves). With `apply_init`, block 0 writes op(init, total)
// (used when one block covers the whole input). NT: non-temporal loads, for
// an input read once (reduce); the float scans' partials pass keeps ordinary
// loads, since their prefixed pass reads the same input again right after.
template <typename T, int OP, int BLOCK, int ROWS, bool NT = false>
__global__ __launch_bounds__(BLOCK) void reduce_tile_kernel(const T* __restrict__ in, uint32_t n,
                                                            bool aligned, T* __restrict__ partials,
                                                            bool apply_init, T init) {
  using C = Combine<T, OP>;
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * ROWS * V;
  __shared__ T s_wave[NW];
  const uint32_t tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * TILE;
  const bool vec_ok = aligned && (base + TILE <= n) && src_tile_ok(in, base, base + TILE);
  T x[ROWS][V];
  load_tile<T, BLOCK, ROWS, NT>(in, base, n, vec_ok, x, C::identity());
  T flat[ROWS * V];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int k = 0; k < V; ++k) flat[r * V + k] = x[r][k];
  T t = tree_fold<T, OP, ROWS * V>(flat);
  t = wave_reduce<T, OP>(t);
  if ((tid & (kWave - 1)) == 0) s_wave[tid / kWave] = t;
  __syncthreads();
  if (tid == 0) {
    T w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = s_wave[i];
    T total = tree_fold<T, OP, NW>(w);
    partials[blockIdx.x] = apply_init ? C::apply(init, total) : total;
  }
}

// Order-preserving tile reduction, for user operators that need not be
// commutative: row by row, each row in index order (the thread's vector, then
// lanes, then waves), and then the rows in order. (reduce_tile_kernel folds a
// thread's rows first, which interleaves rows that lie far apart.) NT as for
// reduce_tile_kernel (pointer sources only).
template <typename T, int OP, int BLOCK, int ROWS, bool NT = false,
          typename Src = const T* __restrict__>
__global__ __launch_bounds__(BLOCK) void reduce_tile_ordered_kernel(Src in,
                                                                    uint32_t n, bool aligned,
                                                                    T* __restrict__ partials,
                                                                    bool apply_init, T init) {
  using C = Combine<T, OP>;
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * ROWS * V;
  __shared__ T s_wave[ROWS][NW];
  const uint32_t tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * TILE;
  const bool vec_ok = aligned && (base + TILE <= n) && src_tile_ok(in, base, base + TILE);
  T x[ROWS][V];
  load_tile<T, BLOCK, ROWS, NT>(in, base, n, vec_ok, x, C::identity());
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    T v = x[r][0];
#pragma unroll
    for (int k = 1; k < V; ++k) v = C::apply(v, x[r][k]);
    v = wave_reduce<T, OP>(v);
    if ((tid & (kWave - 1)) == 0) s_wave[r][tid / kWave] = v;
  }
  __syncthreads();
  if (tid == 0) {
    T acc = s_wave[0][0];
    for (int r = 0; r < ROWS; ++r)
      for (int w = (r == 0 ? 1 : 0); w < NW; ++w) acc = C::apply(acc, s_wave[r][w]);
    partials[blockIdx.x] = apply_init ? C::apply(init, acc) : acc;
  }
}

}  // namespace msv
