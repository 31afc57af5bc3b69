// The one-pass compaction kernel of copy_where / remove_where, partition and
// unique_by_key (see select.hip for the algorithm and the reference it
// follows). A header so that kernels for user-defined predicates can be
// instantiated from it at run time.
#pragma once
#include "device_common.hpp"
#include "predicate.hpp"

namespace msv {

constexpr int kSelBlock = 512;
// Rows of 4-element groups per thread (512-thread blocks, which ran the
// predicate forms 7-9% faster than 256-thread blocks with the same tiles): 2
// (4096-element tiles) for short inputs, 4 (8192) from 2^22 elements on, so
// that fewer tiles take part in the look-back.
constexpr int kSelRowsSmall = 2;
constexpr int kSelRowsLarge = 4;
// Long inputs of 4-byte elements: 8 rows (16384-element tiles, half as many
// look-back steps; the staging LDS is 64 KiB). 8-byte elements keep 4.
template <typename T>
constexpr int sel_rows_large() { return sizeof(T) == 4 ? 8 : kSelRowsLarge; }
constexpr uint32_t kSelLargeMin = 1u << 22;
// Long inputs of 4-byte elements under the stencil and run-head forms (FORM 0
// and 3): 1024 threads x 4 rows, the same 16384-element tiles at 51-63 VGPRs
// and 8 waves per SIMD instead of 4 (124 VGPRs at 512 x 8); 2^28 u32 stencil
// 10% / 50% 0.520 / 0.602 -> 0.493 / 0.553 ms, unique_by_key 0.513 -> 0.482
// ms. The predicate forms keep 512 x 8 (10-15% slower at 1024 x 4).
template <typename T, int FORM>
constexpr int sel_block_large() {
  return sizeof(T) == 4 && (FORM == 0 || FORM == 3) ? 1024 : kSelBlock;
}
template <typename T, int FORM>
constexpr int sel_rows_large_of() {
  return sel_block_large<T, FORM>() == 1024 ? 4 : sel_rows_large<T>();
}

// FORM 0: u32 stencil array (flag = stencil != 0).
// FORM 1: predicate on a separate stencil column of type S.
// FORM 2: predicate on the input itself (S == T, one read).
// FORM 3: run heads of a key column of type S (unique_by_key): an element is
//         kept when it is the first or its key differs from the previous key.
// PART (partition): failing elements are written too, after all passing
// ones: `part_total` holds the number of passing elements, and a failing
// element goes to part_total + (failing elements before it).
// PredT: the predicate's type (the built-in Pred<S>, or a user-defined
// functor compiled at run time). Src: the input source (a pointer, or a lazy
// source compiled at run time).
template <typename T, typename S, int FORM, int ROWS, bool PART = false, typename PredT = Pred<S>,
          typename Src = const T* __restrict__, int BLOCK = kSelBlock>
__global__ __launch_bounds__(BLOCK) void select_kernel(Src in,
                                                           const S* __restrict__ stencil,
                                                           PredT pred, bool invert,
                                                           uint32_t n, bool aligned,
                                                           T* __restrict__ out,
                                                           uint32_t* __restrict__ out_len,
                                                           uint64_t* status, uint32_t gen,
                                                           uint32_t ntiles,
                                                           const uint32_t* __restrict__ part_total) {
  constexpr int NW = BLOCK / kWave;
  constexpr uint32_t kSelTile = (uint32_t)BLOCK * ROWS * 4;
  __shared__ T s_stage[kSelTile];
  __shared__ uint32_t s_wave[ROWS][NW];
  __shared__ uint32_t s_prefix;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * kSelTile;
  const bool full = aligned && base + kSelTile <= n && src_tile_ok(in, base, base + kSelTile);

  T x[ROWS][4];
  uint32_t f[ROWS][4];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const size_t i = base + ((size_t)r * BLOCK + tid) * 4;
    load4<T, FORM == 2>(in, i, n, full, x[r]);  // non-temporal when the input is the only stream
    if constexpr (FORM == 0) {
      S s[4];
      load4<S>(stencil, i, n, full, s);
#pragma unroll
      for (int k = 0; k < 4; ++k) f[r][k] = ((s[k] != 0) != invert) && (i + k < n);
    } else if constexpr (FORM == 1) {
      S s[4];
      load4<S>(stencil, i, n, full, s);
#pragma unroll
      for (int k = 0; k < 4; ++k) f[r][k] = (pred(s[k]) != invert) && (i + k < n);
    } else if constexpr (FORM == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k) f[r][k] = (pred(bit_cast<S>(x[r][k])) != invert) && (i + k < n);
    } else {
      S s[4];
      load4<S>(stencil, i, n, full, s);
      // The key before the group: the previous lane's last key (lane 0
      // reads it; consecutive lanes hold consecutive groups).
      const S before = shift_up1<S>(s[3], (lane == 0 && i > 0 && i <= n) ? stencil[i - 1] : S());
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const S pk = k == 0 ? before : s[k - 1];
        f[r][k] = (i + k < n) && (i + k == 0 || !key_equal(pk, s[k]));
      }
    }
  }

  uint32_t ws[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const uint32_t c = f[r][0] + f[r][1] + f[r][2] + f[r][3];
    ws[r] = wave_inclusive_scan<uint32_t, OP_ADD>(c);
    if (lane == 63) s_wave[r][wid] = ws[r];
  }
  __syncthreads();
  // Exclusive count before (row r, wave wid) and the tile total: every wave
  // scans the ROWS x NW partials in (row, wave) order, one per lane, and takes
  // its own with readlane (wave-uniform values, no per-thread copies).
  static_assert(ROWS * NW <= kWave, "one partial per lane");
  constexpr int NP = ROWS * NW;
  uint32_t q = lane < (uint32_t)NP ? s_wave[lane / NW][lane % NW] : 0u;
  q = wave_inclusive_scan<uint32_t, OP_ADD>(q);
  const uint32_t agg = readlane(q, NP - 1);
  const uint32_t xq = shift_up1<uint32_t>(q, 0u);
  const int w0 = __builtin_amdgcn_readfirstlane((int)wid);
  uint32_t pre[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) pre[r] = readlane(xq, r * NW + w0);

  if (wid == 0) {
    TileStatus<uint32_t> st{status, gen};
    uint32_t excl = 0;
    if (tile == 0) {
      if (lane == 0) st.publish(0, kFlagInc, 0, agg);
    } else {
      if (lane == 0) st.publish(tile, kFlagAgg, 0, agg);
      // Self-help: count the selected elements of tile q from the input.
      auto count_of = [&](uint32_t q) -> uint32_t {
        const size_t b = (size_t)q * kSelTile;
        const size_t e = min((size_t)n, b + kSelTile);
        uint32_t c = 0;
        for (size_t i = b + lane; i < e; i += kWave) {
          bool f;
          if constexpr (FORM == 0) f = stencil[i] != 0;
          else if constexpr (FORM == 1) f = pred(stencil[i]);
          else if constexpr (FORM == 2) f = pred(bit_cast<S>((T)src_at(in, i)));
          else f = i == 0 || !key_equal(stencil[i - 1], stencil[i]);
          c += (f != invert) ? 1u : 0u;
        }
        return wave_reduce<uint32_t, OP_ADD>(c);
      };
      excl = wave_lookback_sh<uint32_t, OP_ADD>(st, tile, count_of);
      if (lane == 0) st.publish(tile, kFlagInc, 0, excl + agg);
    }
    if (lane == 0) {
      s_prefix = excl;
      if (out_len && tile == ntiles - 1) *out_len = excl + agg;
    }
  }

  // Stage selected elements at their in-tile rank (input order).
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    uint32_t pos = pre[r] + shift_up1<uint32_t>(ws[r], 0u);
    const uint32_t j0 = ((uint32_t)r * BLOCK + tid) * 4;  // in-tile index of the group
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (f[r][k]) s_stage[pos] = x[r][k];
      // A failing element's rank among the tile's failing ones: the elements
      // before it minus the passing ones.
      if (PART && !f[r][k] && base + j0 + k < n) s_stage[agg + (j0 + k - pos)] = x[r][k];
      pos += f[r][k];
    }
  }
  __syncthreads();
  const size_t o = s_prefix;
  for (uint32_t i = tid; i < agg; i += BLOCK) out[o + i] = s_stage[i];
  if constexpr (PART) {
    const uint32_t cnt = (uint32_t)min((size_t)kSelTile, (size_t)n - base);
    const size_t fo = (size_t)*part_total + (base - o);  // failing elements before the tile
    for (uint32_t i = agg + tid; i < cnt; i += BLOCK)
      if (fo + (i - agg) < n) out[fo + (i - agg)] = s_stage[i];  // always true when counts agree
  }
}

}  // namespace msv
