// LSD radix sort (8-bit digits, reduce-then-scan passes) for gfx950: vector::radix_sort_by_key, sort_by_key
// and sort with the built-in comparators.
//
// Reference: massively/src/core/radix.rs:369-546 (4-bit digits, 4 launches
// per pass, every pass reading keys through a permutation, a 1-thread bucket
// scan, then a final gather of the values). Order semantics follow
// core/radix.rs:13-118: unsigned order, signed via the flipped sign bit, and
// IEEE total order for floats. For the comparator forms (`Less`/`Greater`)
// -0.0 and +0.0 compare equal, so they are ordered as equal keys (stably).
//
// Pipeline (keys K of 4 or 8 bytes, optional 4/8-byte values):
//  1. pass 0's upsweep reads the caller's keys and also records, per tile,
//     the OR and the AND of the order words;
//  2. plan (one block): a digit position on which OR and AND agree over all
//     keys holds one value and is skipped; the ping-pong source/destination
//     of every remaining pass is written into a device-side table, so no
//     host round trip is needed;
//  3. per remaining digit, reduce-then-scan: an upsweep counts the digits of
//     every tile (512 threads x 16 or 8 keys) into a digit-major
//     count matrix, one exclusive scan of that matrix (the library's lookback
//     scan) gives every (digit, tile) run its global position, and the
//     downsweep ranks each tile's keys stably inside each wave by matching
//     8-bit digits with ballots (wave-level multi-split), reorders the
//     elements through LDS into digit-sorted runs and writes the runs out.
//     Keys and values travel together as one interleaved element between
//     passes (one reorder and one stream of >= 128-byte runs per pass); the
//     first pass reads the caller's arrays and the last writes them. No tile
//     waits on another: a per-digit decoupled lookback measured ~2 ms per
//     2^28-key pass on MI355X, because the per-XCD L2s send every status poll
//     to memory while ~1000 tiles are in flight.
#include "msv_internal.hpp"
#include "radix_kernels.hpp"

namespace msv {

// ---- host -------------------------------------------------------------------
// Tile shape per element width, from a sweep of (threads, keys per thread) at
// 2^28 u32 and 2^27 u64 keys on MI355X: 512 x 16 for 4- and 8-byte elements
// (8192-key tiles, 31-39 Gkeys/s), 512 x 8 for 12- and 16-byte elements (LDS
// holds the whole tile, so wider elements take fewer keys per thread).
template <int EW>
constexpr int radix_block() { return 512; }
template <int EW>
constexpr int radix_ipt() { return EW <= 2 ? 16 : 8; }

template <typename K, int MODE, int VS>
static msv_status radix_sort_impl(Ctx& c, const void* keys, const void* vals, uint32_t n,
                                  void* out_k, void* out_v) {
  constexpr int P = sizeof(K);
  constexpr int EW = (int)(sizeof(K) + VS) / 4;
  constexpr int BLOCK = radix_block<EW>();
  constexpr int IPT = radix_ipt<EW>();
  constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  if (n == 0) return MSV_OK;
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  const size_t elem_bytes = (size_t)n * EW * 4;
  void* aos0 = c.scratch(elem_bytes);
  void* aos1 = c.scratch(elem_bytes);
  uint64_t* orand = (uint64_t*)c.scratch(sizeof(uint64_t) * 2 * ntiles);
  RadixPass* table = (RadixPass*)c.scratch(sizeof(RadixPass) * P);
  uint32_t* counts = (uint32_t*)c.scratch(sizeof(uint32_t) * kRadixBins * (size_t)ntiles);
  uint32_t* offs = (uint32_t*)c.scratch(sizeof(uint32_t) * kRadixBins * (size_t)ntiles);
  if (!aos0 || !aos1 || !orand || !table || !counts || !offs) return MSV_ERR_OUT_OF_MEMORY;

  // Pass 0's counts and the per-tile OR / AND of the order words, then the
  // plan (which passes run, and where each reads and writes).
  hipLaunchKernelGGL((radix_upsweep0_kernel<K, MODE, BLOCK, IPT>), dim3(ntiles), dim3(BLOCK), 0,
                     c.stream(), (const K*)keys, n, ntiles, counts, orand);
  hipLaunchKernelGGL((radix_plan_kernel<P>), dim3(1), dim3(kRadixBins), 0, c.stream(), orand,
                     ntiles, n, table, keys, vals, out_k, out_v, aos0, aos1);
  MSV_HIP(hipGetLastError());
  const uint32_t zero = 0;
  for (int p = 0; p < P; ++p) {
    if (p > 0) {
      hipLaunchKernelGGL((radix_upsweep_kernel<K, MODE, EW, BLOCK, IPT>), dim3(ntiles), dim3(BLOCK),
                         0, c.stream(), table, p, n, ntiles, counts);
      MSV_HIP(hipGetLastError());
    }
    MSV_TRY(scan_run(c, MSV_U32, MSV_OP_ADD, true, counts, offs, kRadixBins * ntiles, &zero));
    hipLaunchKernelGGL((radix_downsweep_kernel<K, MODE, EW, BLOCK, IPT>), dim3(ntiles), dim3(BLOCK), 0,
                       c.stream(), table, p, offs, n, ntiles);
    MSV_HIP(hipGetLastError());
  }
  return MSV_OK;
}

template <typename K, int MODE>
static msv_status radix_by_vsize(Ctx& c, size_t vs, const void* keys, const void* vals,
                                 uint32_t n, void* out_k, void* out_v) {
  if (vs == 0) return radix_sort_impl<K, MODE, 0>(c, keys, nullptr, n, out_k, nullptr);
  if (vs == 4) return radix_sort_impl<K, MODE, 4>(c, keys, vals, n, out_k, out_v);
  return radix_sort_impl<K, MODE, 8>(c, keys, vals, n, out_k, out_v);
}

template <typename K>
static msv_status radix_by_mode(Ctx& c, int mode, size_t vs, const void* keys, const void* vals,
                                uint32_t n, void* out_k, void* out_v) {
  if (mode == ORD_GREATER) return radix_by_vsize<K, ORD_GREATER>(c, vs, keys, vals, n, out_k, out_v);
  if constexpr (IsFloat<K>::value) {
    if (mode == ORD_LESS) return radix_by_vsize<K, ORD_LESS>(c, vs, keys, vals, n, out_k, out_v);
  }
  return radix_by_vsize<K, ORD_TOTAL>(c, vs, keys, vals, n, out_k, out_v);
}

msv_status radix_sort_run(Ctx& c, msv_dtype kdt, int mode, size_t vsize, const void* keys,
                          const void* vals, uint32_t n, void* out_k, void* out_v) {
  switch (kdt) {
    case MSV_U32: return radix_by_mode<uint32_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_I32: return radix_by_mode<int32_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_F32: return radix_by_mode<float>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_U64: return radix_by_mode<uint64_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_I64: return radix_by_mode<int64_t>(c, mode, vsize, keys, vals, n, out_k, out_v);
    case MSV_F64: return radix_by_mode<double>(c, mode, vsize, keys, vals, n, out_k, out_v);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad key dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

}  // namespace msv

using namespace msv;

// Shared front end of the three sorting entry points.
static msv_status sort_api(msv_executor* ex, msv_slice keys, const msv_slice* values, int mode,
                           msv_vec** out_values, msv_vec** out_keys) {
  clear_error();
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  if (values) {
    MSV_TRY(resolve(ex, *values, &v));
    if (k.len != v.len) {
      set_error(MSV_ERR_LENGTH_MISMATCH, k.len, v.len, "input lengths differ: %u != %u", k.len,
                v.len);
      return MSV_ERR_LENGTH_MISMATCH;
    }
    if (!out_values) {
      set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
      return MSV_ERR_INVALID_ARGUMENT;
    }
  }
  msv_vec* ov = nullptr;
  msv_vec* ok = nullptr;
  if (values) MSV_TRY(new_vec(ex, v.dtype, v.len, false, &ov));
  Ctx c(ex);
  void* kdst = nullptr;
  if (out_keys) {
    msv_status s = new_vec(ex, k.dtype, k.len, false, &ok);
    if (s != MSV_OK) {
      if (ov) msv_vec_release(ov);
      return s;
    }
    kdst = ok->ptr;
  } else {
    kdst = c.scratch((size_t)k.len * dtype_size(k.dtype) + 16);
    if (!kdst) {
      if (ov) msv_vec_release(ov);
      return MSV_ERR_OUT_OF_MEMORY;
    }
  }
  msv_status s = radix_sort_run(c, k.dtype, mode, values ? dtype_size(v.dtype) : 0, k.ptr,
                                values ? v.ptr : nullptr, k.len, kdst, ov ? ov->ptr : nullptr);
  if (s != MSV_OK) {
    if (ov) msv_vec_release(ov);
    if (ok) msv_vec_release(ok);
    return s;
  }
  if (out_values) *out_values = ov;
  if (out_keys) *out_keys = ok;
  return MSV_OK;
}

extern "C" {

msv_status msv_radix_sort_by_key(msv_executor* ex, msv_slice keys, msv_slice values,
                                 msv_vec** out_values, msv_vec** out_keys) {
  return sort_api(ex, keys, &values, ORD_TOTAL, out_values, out_keys);
}

msv_status msv_sort_by_key(msv_executor* ex, msv_slice keys, msv_slice values, msv_cmp cmp,
                           msv_vec** out_values, msv_vec** out_keys) {
  if (cmp != MSV_CMP_LESS && cmp != MSV_CMP_GREATER) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown comparator %d", (int)cmp);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return sort_api(ex, keys, &values, cmp == MSV_CMP_LESS ? ORD_LESS : ORD_GREATER, out_values,
                  out_keys);
}

msv_status msv_sort(msv_executor* ex, msv_slice input, msv_cmp cmp, msv_vec** out) {
  if (cmp != MSV_CMP_LESS && cmp != MSV_CMP_GREATER) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown comparator %d", (int)cmp);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (!out) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return sort_api(ex, input, nullptr, cmp == MSV_CMP_LESS ? ORD_LESS : ORD_GREATER, nullptr, out);
}

}  // extern "C"
