// LSD radix sort (8-bit digits) for gfx950: vector::radix_sort_by_key,
// sort_by_key and sort with the built-in comparators.
//
// Reference: massively/src/core/radix.rs:369-546 (4-bit digits, 4 launches
// per pass, every pass reading keys through a permutation, a 1-thread bucket
// scan, then a final gather of the values). Order semantics follow
// core/radix.rs:13-118: unsigned order, signed via the flipped sign bit, and
// IEEE total order for floats. For the comparator forms (`Less`/`Greater`)
// -0.0 and +0.0 compare equal, so they are ordered as equal keys (stably).
//
// Keys and values travel together as one interleaved element between passes
// (one reorder and one stream of runs per pass); the first pass reads the
// caller's arrays and the last writes them. Digit positions on which every key
// has the same digit are skipped (no host round trip: a device-side table
// gives every remaining pass its source and destination). Every pass ranks a
// tile's keys stably inside each wave by matching 8-bit digits with ballots
// (wave-level multi-split) and reorders the tile through LDS into digit runs.
// Two schedules:
//  * onesweep (the default): one read of the keys builds the
//    histograms of every digit position (per-wave LDS histograms), a one-block
//    plan scans them into digit starts and the pass table, then each pass
//    claims tiles in order, publishes a tile's digit counts before ranking it,
//    and finds each digit's global position by a decoupled look-back over the
//    earlier tiles (4 status words per step); elements of 12 or 16 bytes go
//    through LDS as keys and then values, and stay in separate key and value
//    arrays between passes; every pass runs 1024-thread tiles at 8 waves
//    per SIMD (2^27 u64 keys + u32 values: uniform 10.28 ms, Zipf 1.1
//    10.10 ms, constant top byte 9.01 ms);
//  * reduce-then-scan (11.86 / 10.84 / 10.43 ms there, 8.22 ms against 7.80
//    for 2^28 u32 + u32; kept behind the option): pass 0's upsweep also records per
//    tile the OR and AND of the order words (the skip test), and per digit an
//    upsweep counts the digits of every tile into a digit-major matrix, the
//    library's scan turns it into run starts, and a downsweep ranks and writes;
//    no tile waits on another.
// MSV_OPT_RADIX_ONESWEEP selects either.
#include "msv_internal.hpp"
#include "radix_kernels.hpp"

namespace msv {

// ---- host -------------------------------------------------------------------
// Tile shape of the reduce-then-scan passes per element width, from a sweep
// of (threads, keys per thread) at 2^28 u32 and 2^27 u64 keys on MI355X:
// 512 x 16 for 4- and 8-byte elements
// (8192-key tiles, 31-39 Gkeys/s), 512 x 8 for 12- and 16-byte elements (LDS
// holds the whole tile, so wider elements take fewer keys per thread).
template <int EW>
constexpr int radix_block() { return 512; }
template <int EW>
constexpr int radix_ipt() { return EW <= 2 ? 16 : 8; }
// Onesweep passes run 1024-thread blocks: 1024 x 8 for 4-byte elements,
// 1024 x 6 for 8-byte and 1024 x 5 for 12- and 16-byte elements. Each shape
// fits two blocks' LDS per CU (52-68 KB each) at 54-63 VGPRs, so 8 waves per
// SIMD, where the 512-thread tiles (512 x 16, and 512 x 8 at 78 VGPRs for 12
// bytes) held 4-6. Measured against those: 2^28 u32 keys 6.53 -> 5.58 ms,
// 2^28 u32 + u32 7.80 -> 7.35 ms, 2^27 u64 keys 7.61 -> 6.96 ms, 2^27 u64 +
// u32 10.18 -> 9.33 ms, 2^27 u64 + u64 12.46 -> 10.53 ms. Slower: any shape
// above 64 VGPRs (one block per CU: 1024 x 10 or 12 for 4-byte, 1024 x 7 for
// 8-byte elements), smaller tiles at the same occupancy (1024 x 4 for 8-byte
// elements: 9.05 ms for u32 + u32; 1024 x 4 for 12-byte: 10.18 ms), and
// 512 x 4 / 512 x 6 for 12-byte elements.
template <int EW>
constexpr int radix_onesweep_block() { return 1024; }
template <int EW>
constexpr int radix_onesweep_ipt() { return EW >= 3 ? 5 : (EW == 2 ? 6 : 8); }

template <typename K, int MODE, int VS>
static msv_status radix_sort_impl(Ctx& c, const void* keys, const void* vals, uint32_t n,
                                  void* out_k, void* out_v) {
  constexpr int P = sizeof(K);
  constexpr int EW = (int)(sizeof(K) + VS) / 4;
  constexpr int BLOCK = radix_block<EW>();
  constexpr int IPT = radix_ipt<EW>();
  constexpr uint32_t TILE = (uint32_t)BLOCK * IPT;
  if (n == 0) return MSV_OK;
  const uint32_t ntiles = div_up(n, TILE);
  const size_t elem_bytes = (size_t)n * EW * 4;
  // Two-phase onesweep passes keep keys and values apart between passes: the
  // values start at a 256-byte boundary after the keys.
  const size_t vofs = radix_split(EW) ? (((size_t)n * sizeof(K) + 255) & ~(size_t)255) : 0;
  void* aos0 = c.scratch(elem_bytes + 256);
  void* aos1 = c.scratch(elem_bytes + 256);
  RadixPass* table = (RadixPass*)c.scratch(sizeof(RadixPass) * P);
  if (!aos0 || !aos1 || !table) return MSV_ERR_OUT_OF_MEMORY;
  // Schedule: onesweep unless the executor option forces reduce-then-scan.
  const int64_t sched = c.ex()->opt_radix_onesweep;
  if (sched != 0) {
    // Onesweep: histograms of every digit position in one read of the keys,
    // then one look-back pass per digit position that needs one.
    uint32_t* hist = (uint32_t*)c.scratch(sizeof(uint32_t) * (kRadixBins * P + P));
    uint32_t* gbase = (uint32_t*)c.scratch(sizeof(uint32_t) * kRadixBins * P);
    if (!hist || !gbase) return MSV_ERR_OUT_OF_MEMORY;
    uint32_t* tile_counters = hist + kRadixBins * P;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(div_up(kRadixBins * P, 256)), dim3(256), 0,
                       c.stream(), hist, (uint32_t)(kRadixBins * P), 0u);
    const uint32_t hist_blocks = min(div_up(n, 256u * 16u), (uint32_t)c.num_cus() * 8u);
    hipLaunchKernelGGL((radix_hist_kernel<K, MODE, 256>), dim3(hist_blocks), dim3(256), 0,
                       c.stream(), (const K*)keys, n, hist);
    hipLaunchKernelGGL((radix_onesweep_plan_kernel<P>), dim3(1), dim3(kRadixBins), 0, c.stream(),
                       hist, n, table, gbase, tile_counters, keys, vals, out_k, out_v, aos0, aos1,
                       vofs);
    MSV_HIP(hipGetLastError());
    constexpr int OBLOCK = radix_onesweep_block<EW>();
    constexpr int OIPT = radix_onesweep_ipt<EW>();
    const uint32_t otiles = div_up(n, (uint32_t)OBLOCK * OIPT);
    for (int p = 0; p < P; ++p) {
      uint64_t *counter, *status;
      uint32_t gen;
      MSV_TRY(c.lookback((size_t)otiles * kRadixBins, &counter, &status, &gen));
      hipLaunchKernelGGL((radix_onesweep_kernel<K, MODE, EW, OBLOCK, OIPT>), dim3(otiles),
                         dim3(OBLOCK), 0, c.stream(), table, p, gbase, tile_counters, status, gen,
                         n);
      MSV_HIP(hipGetLastError());
    }
    return MSV_OK;
  }
  uint64_t* orand = (uint64_t*)c.scratch(sizeof(uint64_t) * 2 * ntiles);
  uint32_t* counts = (uint32_t*)c.scratch(sizeof(uint32_t) * kRadixBins * (size_t)ntiles);
  uint32_t* offs = (uint32_t*)c.scratch(sizeof(uint32_t) * kRadixBins * (size_t)ntiles);
  if (!orand || !counts || !offs) return MSV_ERR_OUT_OF_MEMORY;

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


// ---- compound keys (up to 3 columns, lexicographic) ------------------------------
// The reference sorts a zip of key columns with one stable radix pass set per
// column, rightmost column first (core/radix.rs:499-531; api/iter.rs:380-388:
// up to three columns, leftmost primary). Here the order-preserving words of
// adjacent columns are packed into one u64 whenever they fit (two u32
// columns: one 64-bit sort instead of two 32-bit ones), so a key needs one
// stable sort per 64-bit group, rightmost group first, each carrying the row
// permutation; value columns are then gathered once.
struct KeyColumn {
  const void* p;
  int dt;    // msv_dtype
  int bits;  // 32 or 64
};

template <typename K>
__device__ inline uint64_t order_bits(const void* p, uint32_t i) {
  return (uint64_t)radix_encode<K>(((const K*)p)[i]);
}

__device__ inline uint64_t column_order_word(const KeyColumn& c, uint32_t i) {
  switch (c.dt) {
    case MSV_U32: return order_bits<uint32_t>(c.p, i);
    case MSV_I32: return order_bits<int32_t>(c.p, i);
    case MSV_F32: return order_bits<float>(c.p, i);
    case MSV_U64: return order_bits<uint64_t>(c.p, i);
    case MSV_I64: return order_bits<int64_t>(c.p, i);
    default: return order_bits<double>(c.p, i);
  }
}

// out[i] = the packed order words of row perm[i] (or i) over columns [0, ncols).
__global__ __launch_bounds__(256) void pack_key_group_kernel(KeyColumn c0, KeyColumn c1, int ncols,
                                                             const uint32_t* __restrict__ perm,
                                                             uint32_t n, uint64_t* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t r = perm ? perm[i] : i;
    uint64_t w = column_order_word(c0, r);
    if (ncols == 2) w = (w << c1.bits) | column_order_word(c1, r);
    out[i] = w;
  }
}

// Row permutation of a stable lexicographic sort by up to 3 key columns.
static msv_status compound_key_perm(Ctx& c, const KeyColumn* cols, int ncols, uint32_t n,
                                    uint32_t* perm) {
  // Groups of adjacent columns of at most 64 bits, formed from the right.
  int gstart[3], gcount[3], ng = 0;
  for (int j = ncols - 1; j >= 0;) {
    int bits = cols[j].bits, first = j;
    while (first > 0 && bits + cols[first - 1].bits <= 64) bits += cols[--first].bits;
    gstart[ng] = first;
    gcount[ng] = j - first + 1;
    ++ng;
    j = first - 1;
  }
  uint64_t* w = (uint64_t*)c.scratch((size_t)n * 8);
  uint64_t* wsorted = (uint64_t*)c.scratch((size_t)n * 8);
  uint32_t* tmp = (uint32_t*)c.scratch((size_t)n * 4);
  if (!w || !wsorted || !tmp) return MSV_ERR_OUT_OF_MEMORY;
  const uint32_t blocks = min((n + 255) / 256, 16384u);
  MSV_TRY(iota_run(c, tmp, n, 0));
  uint32_t* cur = tmp;   // permutation so far (rows in the order sorted by the groups done)
  uint32_t* next = perm;
  for (int g = 0; g < ng; ++g) {
    const KeyColumn a = cols[gstart[g]];
    const KeyColumn b = gcount[g] == 2 ? cols[gstart[g] + 1] : KeyColumn{nullptr, MSV_U32, 32};
    hipLaunchKernelGGL(pack_key_group_kernel, dim3(blocks), dim3(256), 0, c.stream(), a, b,
                       gcount[g], g == 0 ? (const uint32_t*)nullptr : cur, n, w);
    MSV_HIP(hipGetLastError());
    MSV_TRY(radix_sort_run(c, MSV_U64, ORD_TOTAL, 4, w, cur, n, wsorted, next));
    uint32_t* t = cur;
    cur = next;
    next = t;
  }
  if (cur != perm)
    MSV_HIP(hipMemcpyAsync(perm, cur, (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream()));
  return MSV_OK;
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

msv_status msv_radix_sort_by_key_columns(msv_executor* ex, const msv_slice* keys, uint32_t nkeys,
                                         const msv_slice* values, uint32_t nvalues,
                                         msv_vec** out_values) {
  clear_error();
  if (!keys || !values || !out_values || nkeys < 1 || nkeys > 3 || nvalues < 1 || nvalues > 12) {
    set_error(MSV_ERR_INVALID_ARGUMENT, nkeys, nvalues,
              "need 1..3 key columns and 1..12 value columns (got %u, %u)", nkeys, nvalues);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  View kv[3], vv[12];
  for (uint32_t j = 0; j < nkeys; ++j) MSV_TRY(resolve(ex, keys[j], &kv[j]));
  for (uint32_t j = 0; j < nvalues; ++j) MSV_TRY(resolve(ex, values[j], &vv[j]));
  const uint32_t n = kv[0].len;
  for (uint32_t j = 1; j < nkeys; ++j)
    if (kv[j].len != n) {
      set_error(MSV_ERR_LENGTH_MISMATCH, n, kv[j].len, "input lengths differ: %u != %u", n, kv[j].len);
      return MSV_ERR_LENGTH_MISMATCH;
    }
  for (uint32_t j = 0; j < nvalues; ++j)
    if (vv[j].len != n) {
      set_error(MSV_ERR_LENGTH_MISMATCH, n, vv[j].len, "input lengths differ: %u != %u", n, vv[j].len);
      return MSV_ERR_LENGTH_MISMATCH;
    }
  for (uint32_t j = 0; j < nvalues; ++j) out_values[j] = nullptr;
  msv_status s = MSV_OK;
  for (uint32_t j = 0; j < nvalues && s == MSV_OK; ++j)
    s = new_vec(ex, vv[j].dtype, n, false, &out_values[j]);
  Ctx c(ex);
  if (s == MSV_OK && n > 0) {
    KeyColumn cols[3];
    for (uint32_t j = 0; j < nkeys; ++j)
      cols[j] = KeyColumn{kv[j].ptr, (int)kv[j].dtype, (int)dtype_size(kv[j].dtype) * 8};
    uint32_t* perm = (uint32_t*)c.scratch((size_t)n * 4);
    s = perm ? compound_key_perm(c, cols, (int)nkeys, n, perm) : MSV_ERR_OUT_OF_MEMORY;
    for (uint32_t j = 0; j < nvalues && s == MSV_OK; ++j)
      s = gather_run(c, vv[j].dtype, vv[j].ptr, n, perm, out_values[j]->ptr, n);
  }
  if (s != MSV_OK) {
    for (uint32_t j = 0; j < nvalues; ++j)
      if (out_values[j]) {
        msv_vec_release(out_values[j]);
        out_values[j] = nullptr;
      }
  }
  return s;
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
