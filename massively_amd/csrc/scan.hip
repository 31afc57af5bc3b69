// vector::inclusive_scan / exclusive_scan / reduce for gfx950.
//
// Integer scans: one pass, decoupled lookback. A 256-thread block owns a
// tile of BLOCK*ROWS 16-byte vectors (4096 u32). Each thread holds ROWS
// vectors loaded row-striped (row r of thread t starts at (r*BLOCK+t)*V), so
// every load and store instruction moves 1 KiB contiguous per wave. The scan
// order inside the tile is row-major: a thread scans its vector, the wave
// scans the vector totals with DPP, and the ROWS x NW wave totals are folded
// from LDS. Wave 0 publishes the tile aggregate, looks back over up to 64
// predecessor tiles per step, then publishes the inclusive prefix.
//
// Float scans and all reductions: fixed-shape trees (tile partials, then an
// exclusive scan of partials, then the tile scan with its prefix). The
// association depends only on (n, tile shape), so results are bitwise
// reproducible run to run; a decoupled lookback would make float results
// depend on timing.
//
// Reference behaviour: massively/src/core/scan.rs:1125-1234 (inclusive,
// exclusive with init placed before the first element), core/reduce.rs:992-1009
// (reduce returns op(init, total)).
#include "device_common.hpp"
#include "msv_internal.hpp"

namespace msv {

constexpr int kScanBlock = 256;
constexpr int kScanRows = 4;
constexpr int kReduceRows = 16;

template <typename T>
struct alignas(16) Vec16 {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
};

template <typename T, int BLOCK, int ROWS>
__device__ inline void load_tile(const T* __restrict__ in, size_t base, uint32_t n, bool vec_ok,
                                 T (&x)[ROWS][16 / sizeof(T)], T fill) {
  constexpr int V = 16 / sizeof(T);
  const uint32_t tid = threadIdx.x;
  if (vec_ok) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const Vec16<T> q =
          *reinterpret_cast<const Vec16<T>*>(in + base + ((size_t)r * BLOCK + tid) * V);
#pragma unroll
      for (int k = 0; k < V; ++k) x[r][k] = q.v[k];
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

template <typename T, int BLOCK, int ROWS>
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
      *reinterpret_cast<Vec16<T>*>(out + base + ((size_t)r * BLOCK + tid) * V) = q;
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

enum ScanMode : int { SCAN_LOOKBACK = 0, SCAN_PREFIXED = 1 };

// One tile per block. SCAN_LOOKBACK draws tile ids from `counter` and
// resolves the tile prefix by decoupled lookback; SCAN_PREFIXED uses
// blockIdx.x and `prefixes[tile]` (or the identity when prefixes is null).
template <typename T, int OP, bool EXCL, int MODE, int BLOCK, int ROWS>
__global__ __launch_bounds__(BLOCK) void scan_tile_kernel(const T* __restrict__ in,
                                                          T* __restrict__ out, uint32_t n, T init,
                                                          bool aligned, uint64_t* counter,
                                                          uint64_t* status, uint32_t ntiles,
                                                          const T* __restrict__ prefixes) {
  using C = Combine<T, OP>;
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t TILE = (uint32_t)BLOCK * ROWS * V;
  __shared__ T s_wave[ROWS][NW];
  __shared__ T s_prefix;
  __shared__ TileTicket s_ticket;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (kWave - 1);
  const uint32_t wid = tid / kWave;

  uint32_t tile;
  uint32_t gen = 0;
  if constexpr (MODE == SCAN_LOOKBACK) {
    if (tid == 0) s_ticket = draw_tile(counter, ntiles);
    __syncthreads();
    tile = s_ticket.tile;
    gen = s_ticket.gen;
  } else {
    tile = blockIdx.x;
  }
  const size_t base = (size_t)tile * TILE;
  const bool vec_ok = aligned && (base + TILE <= n);
  const T id = C::identity();

  T x[ROWS][V];
  load_tile<T, BLOCK, ROWS>(in, base, n, vec_ok, x, id);

  // Vector-local inclusive scans and wave scans of the vector totals.
  T ws[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
#pragma unroll
    for (int k = 1; k < V; ++k) x[r][k] = C::apply(x[r][k - 1], x[r][k]);
    ws[r] = wave_inclusive_scan<T, OP>(x[r][V - 1]);
    if (lane == kWave - 1) s_wave[r][wid] = ws[r];
  }
  __syncthreads();

  // Exclusive prefix of (row r, wave wid) inside the tile; `agg` = tile total.
  T pre[ROWS];
  T agg = id;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if (w == (int)wid) pre[r] = agg;
      agg = C::apply(agg, s_wave[r][w]);
    }
  }

  if constexpr (MODE == SCAN_LOOKBACK) {
    if (wid == 0) {
      TileStatus<T> st{status, gen};
      T excl = id;
      if (tile == 0) {
        if (lane == 0) st.publish(0, kFlagInc, 0, agg);
      } else {
        if (lane == 0) st.publish(tile, kFlagAgg, 0, agg);
        excl = wave_lookback<T, OP>(st, tile);
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

  T y[ROWS][V];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const T wex = shift_up1<T>(ws[r], id);
    const T p = C::apply(C::apply(tp, pre[r]), wex);
    if constexpr (EXCL) {
      y[r][0] = p;
#pragma unroll
      for (int k = 1; k < V; ++k) y[r][k] = C::apply(p, x[r][k - 1]);
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) y[r][k] = C::apply(p, x[r][k]);
    }
  }
  store_tile<T, BLOCK, ROWS>(out, base, n, vec_ok, y);
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
// (used when one block covers the whole input).
template <typename T, int OP, int BLOCK, int ROWS>
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
  const bool vec_ok = aligned && (base + TILE <= n);
  T x[ROWS][V];
  load_tile<T, BLOCK, ROWS>(in, base, n, vec_ok, x, C::identity());
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

// ---------------------------------------------------------------------------
// Host dispatch.
// ---------------------------------------------------------------------------
template <typename T>
static bool is_aligned16(const void* p) {
  return ((uintptr_t)p & 15u) == 0;
}

template <typename T, int OP>
static msv_status reduce_impl(Ctx& c, const T* in, uint32_t n, T init, T* d_result) {
  constexpr uint32_t TILE = (uint32_t)kScanBlock * kReduceRows * (16 / sizeof(T));
  const T* src = in;
  uint32_t m = n;
  bool aligned = is_aligned16<T>(in);
  for (;;) {
    const uint32_t tiles = m == 0 ? 1 : (m + TILE - 1) / TILE;
    const bool last = tiles == 1;
    T* dst = last ? d_result : (T*)c.scratch((size_t)tiles * sizeof(T));
    if (!dst) return MSV_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL((reduce_tile_kernel<T, OP, kScanBlock, kReduceRows>), dim3(tiles),
                       dim3(kScanBlock), 0, c.stream(), src, m, aligned, dst, last, init);
    MSV_HIP(hipGetLastError());
    if (last) return MSV_OK;
    src = dst;
    m = tiles;
    aligned = true;
  }
}

template <typename T, int OP, bool EXCL>
static msv_status float_scan_impl(Ctx& c, const T* in, T* out, uint32_t n, T init) {
  constexpr uint32_t TILE = (uint32_t)kScanBlock * kScanRows * (16 / sizeof(T));
  const bool aligned = is_aligned16<T>(in) && is_aligned16<T>(out);
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  const T* prefixes = nullptr;
  if (ntiles > 1) {
    T* partials = (T*)c.scratch((size_t)ntiles * sizeof(T));
    T* pre = (T*)c.scratch((size_t)ntiles * sizeof(T));
    if (!partials || !pre) return MSV_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL((reduce_tile_kernel<T, OP, kScanBlock, kScanRows>), dim3(ntiles),
                       dim3(kScanBlock), 0, c.stream(), in, n, aligned, partials, false,
                       Combine<T, OP>::identity());
    MSV_HIP(hipGetLastError());
    MSV_TRY((float_scan_impl<T, OP, true>(c, partials, pre, ntiles, Combine<T, OP>::identity())));
    prefixes = pre;
  }
  hipLaunchKernelGGL((scan_tile_kernel<T, OP, EXCL, SCAN_PREFIXED, kScanBlock, kScanRows>),
                     dim3(ntiles), dim3(kScanBlock), 0, c.stream(), in, out, n, init, aligned,
                     (uint64_t*)nullptr, (uint64_t*)nullptr, ntiles, prefixes);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename T, int OP, bool EXCL>
static msv_status int_scan_impl(Ctx& c, const T* in, T* out, uint32_t n, T init) {
  constexpr uint32_t TILE = (uint32_t)kScanBlock * kScanRows * (16 / sizeof(T));
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  uint64_t* counter;
  uint64_t* status;
  MSV_TRY(c.lookback((size_t)ntiles * TileStatus<T>::kWords, &counter, &status));
  const bool aligned = is_aligned16<T>(in) && is_aligned16<T>(out);
  hipLaunchKernelGGL((scan_tile_kernel<T, OP, EXCL, SCAN_LOOKBACK, kScanBlock, kScanRows>),
                     dim3(ntiles), dim3(kScanBlock), 0, c.stream(), in, out, n, init, aligned,
                     counter, status, ntiles, (const T*)nullptr);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename T, int OP>
static msv_status scan_typed(Ctx& c, bool exclusive, const void* in, void* out, uint32_t n,
                             const void* init_host) {
  T init = Combine<T, OP>::identity();
  if (exclusive) memcpy(&init, init_host, sizeof(T));
  if (n == 0) return MSV_OK;
  if constexpr (IsFloat<T>::value) {
    return exclusive ? float_scan_impl<T, OP, true>(c, (const T*)in, (T*)out, n, init)
                     : float_scan_impl<T, OP, false>(c, (const T*)in, (T*)out, n, init);
  } else {
    return exclusive ? int_scan_impl<T, OP, true>(c, (const T*)in, (T*)out, n, init)
                     : int_scan_impl<T, OP, false>(c, (const T*)in, (T*)out, n, init);
  }
}

template <typename T, int OP>
static msv_status reduce_typed(Ctx& c, const void* in, uint32_t n, const void* init_host,
                               void* out_host) {
  T init;
  memcpy(&init, init_host, sizeof(T));
  T* d_result = (T*)c.scratch(sizeof(T));
  if (!d_result) return MSV_ERR_OUT_OF_MEMORY;
  MSV_TRY((reduce_impl<T, OP>(c, (const T*)in, n, init, d_result)));
  MSV_HIP(hipMemcpyAsync(out_host, d_result, sizeof(T), hipMemcpyDeviceToHost, c.stream()));
  MSV_HIP(hipStreamSynchronize(c.stream()));
  return MSV_OK;
}

#define MSV_DISPATCH_OP(T, FN, ...)                                     \
  switch (op) {                                                        \
    case MSV_OP_ADD: return FN<T, OP_ADD>(__VA_ARGS__);                \
    case MSV_OP_MIN: return FN<T, OP_MIN>(__VA_ARGS__);                \
    case MSV_OP_MAX: return FN<T, OP_MAX>(__VA_ARGS__);                \
    case MSV_OP_XOR:                                                   \
      if constexpr (!IsFloat<T>::value) return FN<T, OP_XOR>(__VA_ARGS__); \
      break;                                                           \
  }

template <typename T>
static msv_status scan_by_op(Ctx& c, msv_op op, bool exclusive, const void* in, void* out,
                             uint32_t n, const void* init) {
  MSV_DISPATCH_OP(T, scan_typed, c, exclusive, in, out, n, init)
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for this dtype", (int)op);
  return MSV_ERR_INVALID_ARGUMENT;
}

template <typename T>
static msv_status reduce_by_op(Ctx& c, msv_op op, const void* in, uint32_t n, const void* init,
                               void* out) {
  MSV_DISPATCH_OP(T, reduce_typed, c, in, n, init, out)
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for this dtype", (int)op);
  return MSV_ERR_INVALID_ARGUMENT;
}

#define MSV_DISPATCH_DTYPE(FN, ...)                         \
  switch (dt) {                                             \
    case MSV_U32: return FN<uint32_t>(__VA_ARGS__);         \
    case MSV_I32: return FN<int32_t>(__VA_ARGS__);          \
    case MSV_F32: return FN<float>(__VA_ARGS__);            \
    case MSV_U64: return FN<uint64_t>(__VA_ARGS__);         \
    case MSV_I64: return FN<int64_t>(__VA_ARGS__);          \
    case MSV_F64: return FN<double>(__VA_ARGS__);           \
  }

msv_status scan_run(Ctx& c, msv_dtype dt, msv_op op, bool exclusive, const void* in, void* out,
                    uint32_t n, const void* init_host) {
  MSV_DISPATCH_DTYPE(scan_by_op, c, op, exclusive, in, out, n, init_host)
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status reduce_run(Ctx& c, msv_dtype dt, msv_op op, const void* in, uint32_t n,
                      const void* init_host, void* out_host) {
  MSV_DISPATCH_DTYPE(reduce_by_op, c, op, in, n, init_host, out_host)
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

static bool op_valid(msv_dtype dt, int op) {
  if (op < MSV_OP_ADD || op > MSV_OP_XOR) return false;
  return !(op == MSV_OP_XOR && dtype_is_float(dt));
}

}  // namespace msv

using namespace msv;

extern "C" {

msv_status msv_inclusive_scan(msv_executor* ex, msv_slice input, msv_op op, msv_vec** out) {
  clear_error();
  View in;
  MSV_TRY(resolve(ex, input, &in));
  if (!op_valid(in.dtype, op)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for dtype %d", (int)op,
              (int)in.dtype);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, in.dtype, in.len, false, out));
  Ctx c(ex);
  msv_status s = scan_run(c, in.dtype, op, false, in.ptr, (*out)->ptr, in.len, nullptr);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

msv_status msv_exclusive_scan(msv_executor* ex, msv_slice input, const void* init, msv_op op,
                              msv_vec** out) {
  clear_error();
  View in;
  MSV_TRY(resolve(ex, input, &in));
  if (!init || !op_valid(in.dtype, op)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null init or operator %d undefined for dtype %d",
              (int)op, (int)in.dtype);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, in.dtype, in.len, false, out));
  Ctx c(ex);
  msv_status s = scan_run(c, in.dtype, op, true, in.ptr, (*out)->ptr, in.len, init);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

msv_status msv_reduce(msv_executor* ex, msv_slice input, const void* init, msv_op op,
                      void* out_host) {
  clear_error();
  View in;
  MSV_TRY(resolve(ex, input, &in));
  if (!init || !out_host || !op_valid(in.dtype, op)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null init/output or operator %d undefined",
              (int)op);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  return reduce_run(c, in.dtype, op, in.ptr, in.len, init, out_host);
}

}  // extern "C"
