// vector::inclusive_scan / exclusive_scan / reduce for gfx950.
//
// Integer scans: one pass, decoupled lookback. A 256-thread block owns tile
// blockIdx.x of BLOCK*ROWS 16-byte vectors (64 KiB for inputs of 2^22 or more
// elements, 16 KiB below). Each thread holds ROWS vectors loaded row-striped
// (row r of thread t starts at (r*BLOCK+t)*V), so every load and store
// instruction moves 1 KiB contiguous per wave, non-temporally. The scan order
// inside the tile is row-major: a thread scans its vector, the wave scans the
// vector totals with DPP, and the ROWS x NW wave totals are folded from LDS.
// Wave 0 publishes the tile aggregate, looks back over up to 64 predecessor
// tiles per step (recomputing a predecessor's aggregate from the input if its
// status does not appear), then publishes the inclusive prefix. Tile ids are
// block indices: no atomic counter serialises the launch.
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
#include "msv_internal.hpp"
#include "scan_kernels.hpp"

namespace msv {

// ---------------------------------------------------------------------------
// Host dispatch.
// ---------------------------------------------------------------------------
template <typename T>
static bool is_aligned16(const void* p) {
  return ((uintptr_t)p & 15u) == 0;
}

template <typename T, int OP>
static msv_status reduce_impl(Ctx& c, const T* in, uint32_t n, T init, T* d_result) {
  constexpr uint32_t TILE = (uint32_t)kReduceBlock * kReduceRows * (16 / sizeof(T));
  const T* src = in;
  uint32_t m = n;
  bool aligned = is_aligned16<T>(in);
  for (;;) {
    const uint32_t tiles = m == 0 ? 1 : div_up(m, TILE);
    const bool last = tiles == 1;
    T* dst = last ? d_result : (T*)c.scratch((size_t)tiles * sizeof(T));
    if (!dst) return MSV_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL((reduce_tile_kernel<T, OP, kReduceBlock, kReduceRows, true>), dim3(tiles),
                       dim3(kReduceBlock), 0, c.stream(), src, m, aligned, dst, last, init);
    MSV_HIP(hipGetLastError());
    if (last) return MSV_OK;
    src = dst;
    m = tiles;
    aligned = true;
  }
}

template <typename T, int OP, bool EXCL>
static msv_status float_scan_impl(Ctx& c, const T* in, T* out, uint32_t n, T init) {
  constexpr int kScanRows = kScanRowsSmall;
  constexpr uint32_t TILE = (uint32_t)kScanBlock * kScanRows * (16 / sizeof(T));
  const bool aligned = is_aligned16<T>(in) && is_aligned16<T>(out);
  const uint32_t ntiles = div_up(n, TILE);
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
                     (uint64_t*)nullptr, 0u, prefixes);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename T, int OP, bool EXCL, int ROWS>
static msv_status int_scan_rows(Ctx& c, const T* in, T* out, uint32_t n, T init) {
  constexpr uint32_t TILE = (uint32_t)kScanBlock * ROWS * (16 / sizeof(T));
  const uint32_t ntiles = div_up(n, TILE);
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback((size_t)ntiles * TileStatus<T>::kWords, &counter, &status, &gen));
  const bool aligned = is_aligned16<T>(in) && is_aligned16<T>(out);
  hipLaunchKernelGGL((scan_tile_kernel<T, OP, EXCL, SCAN_LOOKBACK, kScanBlock, ROWS>),
                     dim3(ntiles), dim3(kScanBlock), 0, c.stream(), in, out, n, init, aligned,
                     status, gen, (const T*)nullptr);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename T, int OP, bool EXCL>
static msv_status int_scan_impl(Ctx& c, const T* in, T* out, uint32_t n, T init) {
  if (n >= kScanLargeMin)
    return int_scan_rows<T, OP, EXCL, scan_rows_large<T>()>(c, in, out, n, init);
  return int_scan_rows<T, OP, EXCL, kScanRowsSmall>(c, in, out, n, init);
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
  MSV_TRY(no_capture(c.ex(), "reduce (host result)"));
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

template <typename T, int OP>
static msv_status reduce_typed_device(Ctx& c, const void* in, uint32_t n, const void* init_host,
                                      void* d_out) {
  T init;
  memcpy(&init, init_host, sizeof(T));
  return reduce_impl<T, OP>(c, (const T*)in, n, init, (T*)d_out);
}

template <typename T>
static msv_status reduce_device_by_op(Ctx& c, msv_op op, const void* in, uint32_t n,
                                      const void* init, void* d_out) {
  MSV_DISPATCH_OP(T, reduce_typed_device, c, in, n, init, d_out)
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

msv_status reduce_run_device(Ctx& c, msv_dtype dt, msv_op op, const void* in, uint32_t n,
                             const void* init_host, void* d_out) {
  MSV_DISPATCH_DTYPE(reduce_device_by_op, c, op, in, n, init_host, d_out)
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
