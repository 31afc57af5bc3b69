// seg::ForEachSegment({Filter, Unique, Take}) for gfx950: the compacting
// segment operations, each one pass over the values.
//
// Reference: massively/src/seg/mod.rs:597-640 (Unique), :810-846 (Filter),
// :847-880 (Take), :1245-1310 (owned results: the kept count is read back,
// the values are an exact-length vector and the offsets a new Segmentation);
// seg/control.rs:555-618 (there: keep flags, an inclusive scan of positions,
// a selected-offsets kernel and a gather). The oracle's loops are
// verification/oracle/src/seg.rs:109-153.
//
// A 512-thread block owns 8192 elements laid out as in copy_where (row r of
// thread t holds elements (r*512 + t)*4 .. +3). It marks the segment starts
// inside its tile in an LDS bitmap (a 64-way wave search finds the tile's
// first segment) and computes each element's keep flag:
//   Filter: the predicate;
//   Unique: a segment start, or a value that differs (numeric ==, as the
//           built-in Equal) from the previous element;
//   Take:   fewer than k elements after the start of its segment (the last
//           start at or before it: the bitmap, a per-word running maximum of
//           start positions, or the start of the segment holding the tile's
//           first element).
// The tile's output offset comes from a decoupled look-back on keep counts
// (blockIdx tiles; a stalled predecessor's count is recomputed by one wave
// from the values and offsets). Kept elements are staged in LDS in input
// order and written contiguously. The new offset of a segment that starts at
// position p is the number of kept elements before p: the tile holding p
// reads it from a kept-bitmap and per-word prefix counts; the last tile
// writes the offsets of segments that start at n and the kept count.
#include "device_common.hpp"
#include "msv_internal.hpp"
#include "predicate.hpp"

namespace msv {

// 512 threads x 4 rows of 4-element groups (8192-element tiles); 512-thread
// blocks ran Filter and Take 8-11% faster than 256 x 8 on MI355X.
constexpr int kCmpBlock = 512;
constexpr int kCmpRows = 4;
constexpr uint32_t kCmpTile = (uint32_t)kCmpBlock * kCmpRows * 4;
constexpr int kCmpWords = (int)(kCmpTile / 32);
static_assert(kCmpWords == 4 * kWave, "one wave covers the bitmap with 4 words per lane");

enum CmpForm : int { CMP_FILTER = 0, CMP_UNIQUE = 1, CMP_TAKE = 2 };

// Kept elements of tile q, counted by one wave from the input (self-help).
template <typename T, int FORM>
__device__ uint32_t wave_count_kept(const T* __restrict__ in, uint32_t n,
                                    const uint32_t* __restrict__ offsets, uint32_t nseg,
                                    const Pred<T>& pred, uint32_t take, uint32_t q) {
  const uint32_t lane = __lane_id();
  const size_t qb = (size_t)q * kCmpTile;
  const uint32_t qe = (uint32_t)min((size_t)n, qb + kCmpTile);
  uint32_t c = 0;
  if constexpr (FORM == CMP_FILTER) {
    for (size_t i = qb + lane; i < qe; i += kWave) c += pred(in[i]) ? 1u : 0u;
  } else {
    const uint32_t lb = wave_lower_bound(offsets, 0, nseg + 1, (uint32_t)qb);
    if constexpr (FORM == CMP_UNIQUE) {
      // Value changes, plus the distinct segment starts whose value equals
      // the previous element's.
      for (size_t i = qb + lane; i < qe; i += kWave)
        c += (i == 0 || !key_equal(in[i - 1], in[i])) ? 1u : 0u;
      for (uint32_t s = lb + lane; s < nseg; s += kWave) {
        const uint32_t p = offsets[s];
        if (p >= qe) break;
        if (p > 0 && offsets[s - 1] != p && key_equal(in[p - 1], in[p])) ++c;
      }
    } else {
      // Each segment meeting the tile keeps [start, start + k) of it.
      for (uint32_t s = (lb > 0 ? lb - 1 : 0) + lane; s < nseg; s += kWave) {
        const uint32_t st = offsets[s];
        if (st >= qe) break;
        const uint64_t lo = max((uint64_t)st, (uint64_t)qb);
        const uint64_t hi = min(min((uint64_t)offsets[s + 1], (uint64_t)st + take), (uint64_t)qe);
        if (hi > lo) c += (uint32_t)(hi - lo);
      }
    }
  }
  return wave_reduce<uint32_t, OP_ADD>(c);
}

template <typename T, int FORM>
__global__ __launch_bounds__(kCmpBlock) void seg_compact_kernel(
    const T* __restrict__ in, uint32_t n, const uint32_t* __restrict__ offsets, uint32_t nseg,
    Pred<T> pred, uint32_t take, bool aligned, T* __restrict__ out,
    uint32_t* __restrict__ out_offsets, uint32_t* __restrict__ out_len, uint64_t* status,
    uint32_t gen, uint32_t ntiles) {
  constexpr int NW = kCmpBlock / kWave;
  __shared__ T s_stage[kCmpTile];
  __shared__ uint32_t s_heads[kCmpWords];
  __shared__ int32_t s_lastpos[FORM == CMP_TAKE ? kCmpWords : 1];
  __shared__ uint32_t s_keep[kCmpWords];
  __shared__ uint32_t s_wpre[kCmpWords];
  __shared__ uint32_t s_wave[kCmpRows][NW];
  __shared__ uint32_t s_first, s_prefix;
  __shared__ int64_t s_cstart;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * kCmpTile;
  const uint32_t tile_end = (uint32_t)min((size_t)n, base + kCmpTile);
  const bool full = aligned && base + kCmpTile <= n;

  for (uint32_t i = tid; i < (uint32_t)kCmpWords; i += kCmpBlock) s_heads[i] = 0;
  if (wid == 0) {
    const uint32_t lb = wave_lower_bound(offsets, 0, nseg + 1, (uint32_t)base);
    if (lane == 0) {
      s_first = lb;
      // Start of the segment holding element `base` (offsets[0] = 0 <= base,
      // and offsets[nseg] = n > base, so lb is in [0, nseg]).
      s_cstart = offsets[lb] == (uint32_t)base ? (int64_t)base : (int64_t)offsets[lb - 1];
    }
  }
  __syncthreads();
  const uint32_t s0 = s_first;
  // Mark the segment starts inside the tile (segment nseg starts at n).
  for (uint32_t c = 0;; c += kCmpBlock) {
    const uint32_t s = s0 + c + tid;
    bool more = false;
    if (s < nseg) {
      const uint32_t st = offsets[s];
      if (st < tile_end) {
        more = tid == kCmpBlock - 1;  // (sorted starts)
        atomicOr(&s_heads[(st - (uint32_t)base) >> 5], 1u << ((st - (uint32_t)base) & 31u));
      }
    }
    if (!__syncthreads_or(more)) break;
  }
  if constexpr (FORM == CMP_TAKE) {
    // s_lastpos[w]: the last start position (in the tile) at or before the
    // end of word w, -1 if none.
    if (wid == 0) {
      int32_t lp[4];
      int32_t run = -1;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t w = lane * 4 + q;
        const uint32_t h = s_heads[w];
        const int32_t last = h ? (int32_t)(w * 32 + 31 - (uint32_t)__builtin_clz(h)) : -1;
        run = max(run, last);
        lp[q] = run;
      }
      const int32_t before = shift_up1<int32_t>(wave_inclusive_scan<int32_t, OP_MAX>(run), -1);
#pragma unroll
      for (int q = 0; q < 4; ++q) s_lastpos[lane * 4 + q] = max(before, lp[q]);
    }
    __syncthreads();
  }

  T x[kCmpRows][4];
  uint32_t fb[kCmpRows];  // keep bits of the thread's group in row r
#pragma unroll
  for (int r = 0; r < kCmpRows; ++r) {
    const uint32_t li = ((uint32_t)r * kCmpBlock + tid) * 4;
    const size_t i = base + li;
    load4<T, true>(in, i, n, full, x[r]);
    uint32_t b = 0;
    if constexpr (FORM == CMP_FILTER) {
#pragma unroll
      for (int k = 0; k < 4; ++k) b |= ((i + k < n) && pred(x[r][k]) ? 1u : 0u) << k;
    } else if constexpr (FORM == CMP_UNIQUE) {
      // The value before the group: the previous lane's last value (lane 0
      // reads it; consecutive lanes hold consecutive groups).
      const T before = shift_up1<T>(x[r][3], (lane == 0 && i > 0 && i <= n) ? in[i - 1] : T());
      const uint32_t hw = s_heads[li >> 5] >> (li & 31u);  // the group's 4 bits share a word
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const T pk = k == 0 ? before : x[r][k - 1];
        const bool keep = ((hw >> k) & 1u) || !key_equal(pk, x[r][k]);
        b |= ((i + k < n) && keep ? 1u : 0u) << k;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t j = li + k;
        const uint32_t w = j >> 5;
        const uint32_t bits = s_heads[w] & (0xffffffffu >> (31u - (j & 31u)));
        int64_t start;
        if (bits) start = (int64_t)base + (int64_t)(w * 32 + 31 - (uint32_t)__builtin_clz(bits));
        else if (w > 0 && s_lastpos[w - 1] >= 0) start = (int64_t)base + s_lastpos[w - 1];
        else start = s_cstart;
        const bool keep = (int64_t)(i + k) - start < (int64_t)take;
        b |= ((i + k < n) && keep ? 1u : 0u) << k;
      }
    }
    fb[r] = b;
  }

  uint32_t ws[kCmpRows];
#pragma unroll
  for (int r = 0; r < kCmpRows; ++r) {
    ws[r] = wave_inclusive_scan<uint32_t, OP_ADD>((uint32_t)__builtin_popcount(fb[r]));
    if (lane == 63) s_wave[r][wid] = ws[r];
  }
  __syncthreads();
  // Exclusive count before (row r, wave wid) and the tile total, scanned by
  // every wave over the row x wave partials and taken by readlane.
  constexpr int NP = kCmpRows * NW;
  uint32_t q = lane < (uint32_t)NP ? s_wave[lane / NW][lane % NW] : 0u;
  q = wave_inclusive_scan<uint32_t, OP_ADD>(q);
  const uint32_t agg = readlane(q, NP - 1);
  const uint32_t xq = shift_up1<uint32_t>(q, 0u);
  const int w0 = __builtin_amdgcn_readfirstlane((int)wid);
  uint32_t pre[kCmpRows];
#pragma unroll
  for (int r = 0; r < kCmpRows; ++r) pre[r] = readlane(xq, r * NW + w0);

  if (wid == 0) {
    TileStatus<uint32_t> st{status, gen};
    uint32_t excl = 0;
    if (tile == 0) {
      if (lane == 0) st.publish(0, kFlagInc, 0, agg);
    } else {
      if (lane == 0) st.publish(tile, kFlagAgg, 0, agg);
      auto count_of = [&](uint32_t qt) -> uint32_t {
        return wave_count_kept<T, FORM>(in, n, offsets, nseg, pred, take, qt);
      };
      excl = wave_lookback_sh<uint32_t, OP_ADD>(st, tile, count_of);
      if (lane == 0) st.publish(tile, kFlagInc, 0, excl + agg);
    }
    if (lane == 0) s_prefix = excl;
  }

  // Stage kept elements at their in-tile rank; the keep bits of 8
  // consecutive groups (lanes 8j .. 8j+7 of a row) form one bitmap word.
#pragma unroll
  for (int r = 0; r < kCmpRows; ++r) {
    uint32_t pos = pre[r] + shift_up1<uint32_t>(ws[r], 0u);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t f = (fb[r] >> k) & 1u;
      if (f) s_stage[pos] = x[r][k];
      pos += f;
    }
    uint32_t word = fb[r] << (4u * (lane & 7u));
    word |= (uint32_t)__shfl_xor((int)word, 1);
    word |= (uint32_t)__shfl_xor((int)word, 2);
    word |= (uint32_t)__shfl_xor((int)word, 4);
    if ((lane & 7u) == 0) s_keep[((uint32_t)r * kCmpBlock + tid) >> 3] = word;
  }
  __syncthreads();
  const uint32_t prefix = s_prefix;
  if (wid == 0) {
    // Exclusive prefix counts of the bitmap words.
    uint32_t c4[4], tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      c4[k] = (uint32_t)__builtin_popcount(s_keep[lane * 4 + k]);
      tot += c4[k];
    }
    uint32_t run = wave_inclusive_scan<uint32_t, OP_ADD>(tot) - tot;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      s_wpre[lane * 4 + k] = run;
      run += c4[k];
    }
  }
  for (uint32_t i = tid; i < agg; i += kCmpBlock) out[(size_t)prefix + i] = s_stage[i];
  __syncthreads();
  // New offsets of the segments that start in the tile.
  for (uint32_t c = 0;; c += kCmpBlock) {
    const uint32_t s = s0 + c + tid;
    bool more = false;
    if (s < nseg) {
      const uint32_t st = offsets[s];
      if (st < tile_end) {
        more = tid == kCmpBlock - 1;  // (sorted starts)
        const uint32_t j = st - (uint32_t)base;
        out_offsets[s] = prefix + s_wpre[j >> 5] +
                         (uint32_t)__builtin_popcount(s_keep[j >> 5] & ((1u << (j & 31u)) - 1u));
      }
    }
    if (!__syncthreads_or(more)) break;
  }
  if (tile == ntiles - 1) {
    // Segments that start at n (trailing empty ones and the end offset).
    const uint32_t total = prefix + agg;
    if (wid == 0) {
      const uint32_t lbn = wave_lower_bound(offsets, 0, nseg + 1, n);
      if (lane == 0) {
        s_first = lbn;
        *out_len = total;
      }
    }
    __syncthreads();
    for (uint32_t s = s_first + tid; s <= nseg; s += kCmpBlock) out_offsets[s] = total;
  }
}

static __global__ void cmp_zero_kernel(uint32_t* __restrict__ p, uint32_t m, uint32_t* len) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) p[i] = 0;
  if (i == 0) *len = 0;
}

template <typename T, int FORM>
static msv_status seg_compact_launch(Ctx& c, const void* in, uint32_t n, const uint32_t* offsets,
                                     uint32_t nseg, const msv_predicate* p, uint32_t take,
                                     void* out, uint32_t* out_offsets, uint32_t* out_len) {
  if (n == 0) {
    hipLaunchKernelGGL(cmp_zero_kernel, dim3((nseg + 1 + 255) / 256), dim3(256), 0, c.stream(),
                       out_offsets, nseg + 1, out_len);
    MSV_HIP(hipGetLastError());
    return MSV_OK;
  }
  const Pred<T> pred = make_pred<T>(p);
  const uint32_t ntiles = div_up(n, kCmpTile);
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
  const bool aligned = ((uintptr_t)in & 15u) == 0;
  hipLaunchKernelGGL((seg_compact_kernel<T, FORM>), dim3(ntiles), dim3(kCmpBlock), 0, c.stream(),
                     (const T*)in, n, offsets, nseg, pred, take, aligned, (T*)out, out_offsets,
                     out_len, status, gen, ntiles);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <int FORM>
static msv_status seg_compact_run(Ctx& c, msv_dtype dt, const void* in, uint32_t n,
                                  const uint32_t* offsets, uint32_t nseg, const msv_predicate* p,
                                  uint32_t take, void* out, uint32_t* out_offsets,
                                  uint32_t* out_len) {
#define MSV_CMP(T) \
  return seg_compact_launch<T, FORM>(c, in, n, offsets, nseg, p, take, out, out_offsets, out_len)
  if constexpr (FORM == CMP_TAKE) {
    // Take moves elements without looking at them.
    if (dtype_size(dt) == 4) MSV_CMP(uint32_t);
    MSV_CMP(uint64_t);
  } else {
    switch (dt) {
      case MSV_U32: MSV_CMP(uint32_t);
      case MSV_I32: MSV_CMP(int32_t);
      case MSV_F32: MSV_CMP(float);
      case MSV_U64: MSV_CMP(uint64_t);
      case MSV_I64: MSV_CMP(int64_t);
      case MSV_F64: MSV_CMP(double);
    }
  }
#undef MSV_CMP
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

}  // namespace msv

using namespace msv;

// Shared by the three entry points: checks, outputs, one launch, and the
// read-back of the kept count (the reference reads it too, to size the
// exact-length result and build the Segmentation).
static msv_status seg_compact_api(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                                  int form, const msv_predicate* pred, uint32_t take,
                                  msv_vec** out_values, msv_segmentation** out_seg) {
  clear_error();
  MSV_TRY(no_capture(ex, "a compacting segment operation"));
  View v;
  MSV_TRY(resolve(ex, values, &v));
  if (!seg || !seg->offsets || !out_values || !out_seg) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null segmentation or output");
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
  if (pred && pred->kind > PRED_BITS_EQ) {
    set_error(MSV_ERR_INVALID_ARGUMENT, pred->kind, 0, "unknown predicate kind %u", pred->kind);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  const uint32_t nseg = seg->segment_count;
  msv_vec *vals = nullptr, *offs = nullptr;
  MSV_TRY(new_vec(ex, v.dtype, v.len, false, &vals));
  msv_status s = new_vec(ex, MSV_U32, nseg + 1, false, &offs);
  Ctx c(ex);
  uint32_t* d_count = s == MSV_OK ? (uint32_t*)c.scratch(sizeof(uint32_t)) : nullptr;
  if (s == MSV_OK && !d_count) s = MSV_ERR_OUT_OF_MEMORY;
  const uint32_t* offsets = (const uint32_t*)seg->offsets->ptr;
  if (s == MSV_OK) {
    uint32_t* oo = (uint32_t*)offs->ptr;
    if (form == CMP_FILTER)
      s = seg_compact_run<CMP_FILTER>(c, v.dtype, v.ptr, v.len, offsets, nseg, pred, 0, vals->ptr, oo, d_count);
    else if (form == CMP_UNIQUE)
      s = seg_compact_run<CMP_UNIQUE>(c, v.dtype, v.ptr, v.len, offsets, nseg, nullptr, 0, vals->ptr, oo, d_count);
    else
      s = seg_compact_run<CMP_TAKE>(c, v.dtype, v.ptr, v.len, offsets, nseg, nullptr, take, vals->ptr, oo, d_count);
  }
  uint32_t kept = 0;
  if (s == MSV_OK) {
    hipError_t e = hipMemcpyAsync(&kept, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, ex->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ex->stream);
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "segment compaction failed: %s", hipGetErrorString(e));
      s = MSV_ERR_LAUNCH;
    }
  }
  if (s != MSV_OK) {
    if (vals) msv_vec_release(vals);
    if (offs) msv_vec_release(offs);
    return s;
  }
  vals->host_len = kept;
  msv_segmentation* sg = new msv_segmentation;
  sg->offsets = offs;
  sg->segment_count = nseg;
  sg->value_count = kept;
  *out_values = vals;
  *out_seg = sg;
  return MSV_OK;
}

extern "C" {

msv_status msv_seg_filter(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                          const msv_predicate* pred, msv_vec** out_values,
                          msv_segmentation** out_seg) {
  if (!pred) {
    clear_error();
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null predicate");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return seg_compact_api(ex, values, seg, CMP_FILTER, pred, 0, out_values, out_seg);
}

msv_status msv_seg_unique(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                          msv_vec** out_values, msv_segmentation** out_seg) {
  return seg_compact_api(ex, values, seg, CMP_UNIQUE, nullptr, 0, out_values, out_seg);
}

msv_status msv_seg_take(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                        uint32_t count, msv_vec** out_values, msv_segmentation** out_seg) {
  return seg_compact_api(ex, values, seg, CMP_TAKE, nullptr, count, out_values, out_seg);
}

}  // extern "C"
