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

// ---------------------------------------------------------------------------
// Inclusive scan without allocation. Its look-back status words (integer
// scans) and tile partials (float scans) live in `scratch`, a static array of
// kNoAllocScratchBytes per device; the scan runs the same host code and
// kernels as msv_inclusive_scan through a Ctx in arena mode.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchBytes / 8];

struct NoAllocDevice {
  std::mutex mu;
  uint64_t* ptr = nullptr;  // device address of `scratch` on this device
  // Generation of the next uncaptured call. Generation 1 belongs to captured
  // calls, which clear their words first, so uncaptured ones start at 2.
  uint32_t gen = 2;
};
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];

static __global__ void noalloc_clear_kernel(uint64_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

static msv_status noalloc_clear(uint64_t* p, hipStream_t s) {
  hipLaunchKernelGGL(noalloc_clear_kernel, dim3(256), dim3(256), 0, s, p,
                     kNoAllocScratchBytes / 8);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

msv_status noalloc_setup(int device, hipStream_t stream) {
  if (device < 0 || device >= kNoAllocMaxDevices) return MSV_OK;  // entry point reports it
  NoAllocDevice& d = g_noalloc[device];
  std::lock_guard<std::mutex> lk(d.mu);
  if (d.ptr) return MSV_OK;
  void* p = nullptr;
  MSV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  MSV_TRY(noalloc_clear((uint64_t*)p, stream));
  MSV_HIP(hipStreamSynchronize(stream));
  d.ptr = (uint64_t*)p;
  return MSV_OK;
}

// Arena bytes the inclusive scan of n elements of T takes (the pieces
// float_scan_impl and int_scan_rows request, in the order they request them).
template <typename T>
static size_t scan_arena_bytes(uint32_t n) {
  if (n == 0) return 0;
  if constexpr (IsFloat<T>::value) {
    constexpr uint32_t TILE = (uint32_t)kScanBlock * kScanRowsSmall * (16 / sizeof(T));
    size_t total = 0;
    for (uint32_t m = n;;) {
      const uint32_t tiles = div_up(m, TILE);
      if (tiles <= 1) return total;
      total += 2 * arena_round((size_t)tiles * sizeof(T));
      m = tiles;
    }
  } else {
    const uint32_t rows = n >= kScanLargeMin ? scan_rows_large<T>() : kScanRowsSmall;
    const uint32_t tile = (uint32_t)kScanBlock * rows * (16 / sizeof(T));
    return arena_round(((size_t)div_up(n, tile) * TileStatus<T>::kWords + 1) * 8);
  }
}

static size_t scan_arena_bytes(msv_dtype dt, uint32_t n) {
  switch (dt) {
    case MSV_U32: return scan_arena_bytes<uint32_t>(n);
    case MSV_I32: return scan_arena_bytes<int32_t>(n);
    case MSV_F32: return scan_arena_bytes<float>(n);
    case MSV_U64: return scan_arena_bytes<uint64_t>(n);
    case MSV_I64: return scan_arena_bytes<int64_t>(n);
    case MSV_F64: return scan_arena_bytes<double>(n);
  }
  return 0;
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

msv_status msv_inclusive_scan_noalloc(msv_executor* ex, msv_slice input, msv_op op,
                                      msv_slice output) {
  clear_error();
  View in, out;
  MSV_TRY(resolve(ex, input, &in));
  MSV_TRY(resolve(ex, output, &out));
  if (!op_valid(in.dtype, op)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for dtype %d", (int)op,
              (int)in.dtype);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (out.dtype != in.dtype) {
    set_error(MSV_ERR_INVALID_ARGUMENT, (uint64_t)in.dtype, (uint64_t)out.dtype,
              "output dtype %d differs from input dtype %d", (int)out.dtype, (int)in.dtype);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (out.len < in.len) {
    set_error(MSV_ERR_OUTPUT_TOO_SHORT, in.len, out.len, "output of %u elements for %u inputs",
              out.len, in.len);
    return MSV_ERR_OUTPUT_TOO_SHORT;
  }
  const size_t es = dtype_size(in.dtype);
  const uintptr_t ib = (uintptr_t)in.ptr, ob = (uintptr_t)out.ptr;
  if (in.len && ib < ob + (size_t)in.len * es && ob < ib + (size_t)in.len * es) {
    // A look-back tile may re-read a predecessor's input after it was scanned.
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "input and output ranges overlap");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  const size_t need = scan_arena_bytes(in.dtype, in.len);
  if (need > kNoAllocScratchBytes) {
    set_error(MSV_ERR_LENGTH_TOO_LARGE, in.len, 0,
              "a scan of %u elements of dtype %d needs %zu bytes of static scratch; the maximum "
              "is %zu", in.len, (int)in.dtype, need, kNoAllocScratchBytes);
    return MSV_ERR_LENGTH_TOO_LARGE;
  }
  if (ex->device < 0 || ex->device >= kNoAllocMaxDevices || !g_noalloc[ex->device].ptr) {
    set_error(MSV_ERR_UNSUPPORTED, (uint64_t)ex->device, 0,
              "no static scratch on device %d", ex->device);
    return MSV_ERR_UNSUPPORTED;
  }
  if (in.len == 0) return MSV_OK;
  MSV_HIP(hipSetDevice(ex->device));
  NoAllocDevice& d = g_noalloc[ex->device];
  uint32_t gen = 1;
  if (!ex->capturing) {
    std::lock_guard<std::mutex> lk(d.mu);
    // Generations are 29 bits; before one would repeat, the array is cleared.
    if (d.gen >= (1u << 29) - 1) {
      MSV_TRY(noalloc_clear(d.ptr, ex->stream));
      d.gen = 2;
    }
    gen = d.gen++;
  }
  Ctx c(ex);
  c.use_arena(d.ptr, kNoAllocScratchBytes, gen);
  return scan_run(c, in.dtype, op, false, in.ptr, out.ptr, in.len, nullptr);
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
