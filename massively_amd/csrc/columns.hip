// Zipped rows (up to 12 columns of equal length, the reference's zip2 ..
// zip12 leaves) moved by one launch that reads the shared input once:
//
//  * gather (massively/src/core/indexed.rs:16-196, indexed_copy over the
//    leaves of a zipped read): each thread loads its 16 indices once, then for
//    every column issues its 16 reads and stores 16 results;
//  * scatter (api/algorithm/scatter.rs:248-265 over zipped values): the same,
//    storing to the columns' outputs (outputs below the binned-scatter size;
//    larger ones are scattered column by column in the binned form, where the
//    random writes, not the index reads, set the cost);
//  * copy_where / remove_where (core/selection.rs:151-411, the selection copy
//    over zipped leaves): one pass per tile computes the keep flags from the
//    stencil, their in-tile ranks and, by decoupled look-back, the tile's
//    output offset; then every column's tile is loaded (the next column's
//    loads in flight while the current one is staged), staged in LDS at the
//    ranks and stored contiguously. The output length is written to every
//    column's device length.
//
// Columns of 4-byte and 8-byte elements mix freely: a column is moved as
// words of its width.
#include "device_common.hpp"
#include "msv_internal.hpp"
#include "predicate.hpp"

namespace msv {

constexpr uint32_t kMaxColumns = 12;

struct ColumnSet {
  const void* in[kMaxColumns];
  void* out[kMaxColumns];
  uint32_t* out_len[kMaxColumns];
  uint32_t words[kMaxColumns];  // 1: 4-byte elements, 2: 8-byte elements
  uint32_t ncols;
};

constexpr int kColBlock = 256;
constexpr int kColRows = 4;  // 4 groups of 4 consecutive elements per thread
constexpr uint32_t kColTile = (uint32_t)kColBlock * kColRows * 4;

// The thread's 16 indices: rows of 4 consecutive positions; positions past n
// read as 0xffffffff, past every column's end, so one test per element
// covers both.
__device__ __forceinline__ void load_indices(const uint32_t* __restrict__ idx, size_t base,
                                             uint32_t n, bool full, uint32_t (&ix)[kColRows][4]) {
  if (full) {
#pragma unroll
    for (int r = 0; r < kColRows; ++r)
      load4<uint32_t>(idx, base + ((size_t)r * kColBlock + threadIdx.x) * 4, n, true, ix[r]);
  } else {
#pragma unroll
    for (int r = 0; r < kColRows; ++r)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t i = base + ((size_t)r * kColBlock + threadIdx.x) * 4 + k;
        ix[r][k] = i < n ? idx[i] : 0xffffffffu;
      }
  }
}

template <typename W>
__device__ __forceinline__ void gather_column(const W* __restrict__ in, W* __restrict__ out,
                                              const uint32_t (&ix)[kColRows][4], size_t base,
                                              uint32_t n, uint32_t vlen, bool full) {
  W v[kColRows][4];
#pragma unroll
  for (int r = 0; r < kColRows; ++r)
#pragma unroll
    for (int k = 0; k < 4; ++k) v[r][k] = ix[r][k] < vlen ? in[ix[r][k]] : W(0);
#pragma unroll
  for (int r = 0; r < kColRows; ++r) {
    const size_t i = base + ((size_t)r * kColBlock + threadIdx.x) * 4;
    if (full) {
      if constexpr (sizeof(W) == 4) {
        *reinterpret_cast<uint4*>(out + i) = make_uint4(v[r][0], v[r][1], v[r][2], v[r][3]);
      } else {
        reinterpret_cast<uint4*>(out + i)[0] =
            make_uint4((uint32_t)v[r][0], (uint32_t)(v[r][0] >> 32), (uint32_t)v[r][1],
                       (uint32_t)(v[r][1] >> 32));
        reinterpret_cast<uint4*>(out + i)[1] =
            make_uint4((uint32_t)v[r][2], (uint32_t)(v[r][2] >> 32), (uint32_t)v[r][3],
                       (uint32_t)(v[r][3] >> 32));
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (i + k < n) out[i + k] = v[r][k];
    }
  }
}

// out_c[i] = in_c[idx[i]] for every column (0 past a column's end). WIDE:
// some column has 8-byte elements (an instance for 4-byte columns only keeps
// the register count, and so the random reads in flight, of the one-column
// gather).
template <bool WIDE>
__global__ __launch_bounds__(kColBlock) void gather_columns_kernel(ColumnSet cs,
                                                                   const uint32_t* __restrict__ idx,
                                                                   uint32_t n, uint32_t vlen,
                                                                   bool aligned) {
  const size_t base = (size_t)blockIdx.x * kColTile;
  const bool full = aligned && base + kColTile <= n;
  uint32_t ix[kColRows][4];
  load_indices(idx, base, n, full, ix);
#pragma unroll 1
  for (uint32_t c = 0; c < cs.ncols; ++c) {
    if (!WIDE || cs.words[c] == 1)
      gather_column<uint32_t>((const uint32_t*)cs.in[c], (uint32_t*)cs.out[c], ix, base, n, vlen,
                              full);
    else
      gather_column<uint64_t>((const uint64_t*)cs.in[c], (uint64_t*)cs.out[c], ix, base, n, vlen,
                              full);
  }
}

template <typename W>
__device__ __forceinline__ void scatter_column(const W* __restrict__ in, W* __restrict__ out,
                                               const uint32_t (&ix)[kColRows][4], size_t base,
                                               uint32_t n, uint32_t olen, bool full) {
  W v[kColRows][4];
#pragma unroll
  for (int r = 0; r < kColRows; ++r)
    load4<W>(in, base + ((size_t)r * kColBlock + threadIdx.x) * 4, n, full, v[r]);
#pragma unroll
  for (int r = 0; r < kColRows; ++r)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (ix[r][k] < olen) out[ix[r][k]] = v[r][k];
}

// out_c[idx[i]] = in_c[i] for every column (indices past the outputs' end are
// dropped).
__global__ __launch_bounds__(kColBlock) void scatter_columns_kernel(ColumnSet cs,
                                                                    const uint32_t* __restrict__ idx,
                                                                    uint32_t n, uint32_t olen,
                                                                    bool aligned) {
  const size_t base = (size_t)blockIdx.x * kColTile;
  const bool full = aligned && base + kColTile <= n;
  uint32_t ix[kColRows][4];
  load_indices(idx, base, n, full, ix);
  for (uint32_t c = 0; c < cs.ncols; ++c) {
    if (cs.words[c] == 1)
      scatter_column<uint32_t>((const uint32_t*)cs.in[c], (uint32_t*)cs.out[c], ix, base, n, olen,
                               full);
    else
      scatter_column<uint64_t>((const uint64_t*)cs.in[c], (uint64_t*)cs.out[c], ix, base, n, olen,
                               full);
  }
}

// Copy_where over columns. 512 threads, 2 rows of 4-element groups per thread
// (4096-element tiles: the staging area holds a tile of 8-byte words).
constexpr int kCselBlock = 512;
constexpr int kCselRows = 2;
constexpr uint32_t kCselTile = (uint32_t)kCselBlock * kCselRows * 4;

__device__ __forceinline__ void load_column_words(const void* p, uint32_t words, size_t i,
                                                  uint32_t n, bool full, uint64_t (&v)[4]) {
  if (words == 1) {
    uint32_t w[4];
    load4<uint32_t>((const uint32_t*)p, i, n, full, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = w[k];
  } else {
    load4<uint64_t>((const uint64_t*)p, i, n, full, v);
  }
}

// FORM 0: u32 stencil (flag = stencil != 0); FORM 1: predicate on a stencil
// column of type S.
template <typename S, int FORM>
__global__ __launch_bounds__(kCselBlock) void select_columns_kernel(
    ColumnSet cs, const S* __restrict__ stencil, Pred<S> pred, bool invert, uint32_t n,
    bool aligned, uint64_t* status, uint32_t gen, uint32_t ntiles) {
  constexpr int NW = kCselBlock / kWave;
  constexpr int ROWS = kCselRows;
  __shared__ uint64_t s_stage[kCselTile];
  __shared__ uint32_t s_wave[ROWS][NW];
  __shared__ uint32_t s_prefix;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid / 64u;
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * kCselTile;
  const bool full = aligned && base + kCselTile <= n;

  // The first column's loads go out with the stencil's.
  uint64_t x[ROWS][4];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
    load_column_words(cs.in[0], cs.words[0], base + ((size_t)r * kCselBlock + tid) * 4, n, full,
                      x[r]);
  uint32_t f[ROWS][4];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const size_t i = base + ((size_t)r * kCselBlock + tid) * 4;
    S s[4];
    load4<S>(stencil, i, n, full, s);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool keep = FORM == 0 ? s[k] != (S)0 : pred(s[k]);
      f[r][k] = (keep != invert) && (i + k < n);
    }
  }
  uint32_t ws[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    ws[r] = wave_inclusive_scan<uint32_t, OP_ADD>(f[r][0] + f[r][1] + f[r][2] + f[r][3]);
    if (lane == 63) s_wave[r][wid] = ws[r];
  }
  __syncthreads();
  static_assert(ROWS * NW <= kWave, "one partial per lane");
  constexpr int NP = ROWS * NW;
  uint32_t q = lane < (uint32_t)NP ? s_wave[lane / NW][lane % NW] : 0u;
  q = wave_inclusive_scan<uint32_t, OP_ADD>(q);
  const uint32_t agg = readlane(q, NP - 1);
  const uint32_t xq = shift_up1<uint32_t>(q, 0u);
  const int w0 = __builtin_amdgcn_readfirstlane((int)wid);
  uint32_t pre[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) pre[r] = readlane(xq, r * NW + w0) + shift_up1<uint32_t>(ws[r], 0u);

  if (wid == 0) {
    TileStatus<uint32_t> st{status, gen};
    uint32_t excl = 0;
    if (tile == 0) {
      if (lane == 0) st.publish(0, kFlagInc, 0, agg);
    } else {
      if (lane == 0) st.publish(tile, kFlagAgg, 0, agg);
      // Self-help: count the kept elements of tile q from the stencil.
      auto count_of = [&](uint32_t qt) -> uint32_t {
        const size_t b = (size_t)qt * kCselTile;
        const size_t e = min((size_t)n, b + kCselTile);
        uint32_t c = 0;
        for (size_t i = b + lane; i < e; i += kWave) {
          const bool keep = FORM == 0 ? stencil[i] != (S)0 : pred(stencil[i]);
          c += (keep != invert) ? 1u : 0u;
        }
        return wave_reduce<uint32_t, OP_ADD>(c);
      };
      excl = wave_lookback_sh<uint32_t, OP_ADD>(st, tile, count_of);
      if (lane == 0) st.publish(tile, kFlagInc, 0, excl + agg);
    }
    if (lane == 0) s_prefix = excl;
  }

  for (uint32_t c = 0; c < cs.ncols; ++c) {
    uint64_t y[ROWS][4];
    if (c + 1 < cs.ncols) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        load_column_words(cs.in[c + 1], cs.words[c + 1],
                          base + ((size_t)r * kCselBlock + tid) * 4, n, full, y[r]);
    }
    // Stage the kept elements at their in-tile rank (input order).
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      uint32_t pos = pre[r];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (f[r][k]) s_stage[pos] = x[r][k];
        pos += f[r][k];
      }
    }
    __syncthreads();
    const size_t o = s_prefix;
    if (cs.words[c] == 1) {
      uint32_t* out = (uint32_t*)cs.out[c];
      for (uint32_t i = tid; i < agg; i += kCselBlock) out[o + i] = (uint32_t)s_stage[i];
    } else {
      uint64_t* out = (uint64_t*)cs.out[c];
      for (uint32_t i = tid; i < agg; i += kCselBlock) out[o + i] = s_stage[i];
    }
    if (tid == 0 && tile == ntiles - 1) *cs.out_len[c] = (uint32_t)o + agg;
    __syncthreads();
    if (c + 1 < cs.ncols) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) x[r][k] = y[r][k];
    }
  }
}

__global__ void column_lengths_zero_kernel(ColumnSet cs) {
  if (threadIdx.x < cs.ncols) *cs.out_len[threadIdx.x] = 0;
}

// ---- host ---------------------------------------------------------------------

// Resolves `ncols` column slices of one length into `views`.
static msv_status resolve_columns(msv_executor* ex, const msv_slice* cols, uint32_t ncols,
                                  View* views) {
  if (!cols || ncols == 0 || ncols > kMaxColumns) {
    set_error(MSV_ERR_INVALID_ARGUMENT, ncols, kMaxColumns, "1..%u columns expected, got %u",
              kMaxColumns, ncols);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t c = 0; c < ncols; ++c) {
    MSV_TRY(resolve(ex, cols[c], &views[c]));
    if (views[c].len != views[0].len) {
      set_error(MSV_ERR_LENGTH_MISMATCH, views[0].len, views[c].len,
                "input lengths differ: %u != %u", views[0].len, views[c].len);
      return MSV_ERR_LENGTH_MISMATCH;
    }
  }
  return MSV_OK;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

static void release_all(msv_vec** out, uint32_t k) {
  for (uint32_t c = 0; c < k; ++c) {
    if (out[c]) msv_vec_release(out[c]);
    out[c] = nullptr;
  }
}

template <typename S>
static msv_status select_columns_launch(Ctx& c, ColumnSet& cs, const void* st,
                                        const msv_predicate* p, bool invert, uint32_t n,
                                        bool aligned) {
  const uint32_t ntiles = div_up(n, kCselTile);
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback((size_t)ntiles * TileStatus<uint32_t>::kWords, &counter, &status, &gen));
  const Pred<S> pred = make_pred<S>(p);
  if (p)
    hipLaunchKernelGGL((select_columns_kernel<S, 1>), dim3(ntiles), dim3(kCselBlock), 0,
                       c.stream(), cs, (const S*)st, pred, invert, n, aligned, status, gen, ntiles);
  else
    hipLaunchKernelGGL((select_columns_kernel<S, 0>), dim3(ntiles), dim3(kCselBlock), 0,
                       c.stream(), cs, (const S*)st, pred, invert, n, aligned, status, gen, ntiles);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

}  // namespace msv

using namespace msv;

extern "C" {

msv_status msv_gather_columns(msv_executor* ex, const msv_slice* values, uint32_t ncols,
                              msv_slice indices, msv_vec** out) {
  clear_error();
  View v[kMaxColumns], ix;
  MSV_TRY(resolve_columns(ex, values, ncols, v));
  MSV_TRY(resolve(ex, indices, &ix));
  if (ix.dtype != MSV_U32 || !out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "indices must be u32 (MIndex), outputs non-null");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t c = 0; c < ncols; ++c) out[c] = nullptr;
  ColumnSet cs{};
  cs.ncols = ncols;
  for (uint32_t c = 0; c < ncols; ++c) {
    msv_status s = new_vec(ex, v[c].dtype, ix.len, false, &out[c]);
    if (s != MSV_OK) {
      release_all(out, c);
      return s;
    }
    cs.in[c] = v[c].ptr;
    cs.out[c] = out[c]->ptr;
    cs.words[c] = (uint32_t)(dtype_size(v[c].dtype) / 4);
  }
  if (ix.len == 0) return MSV_OK;
  Ctx c(ex);
  bool wide = false;
  for (uint32_t k = 0; k < ncols; ++k) wide = wide || cs.words[k] == 2;
  if (wide)
    hipLaunchKernelGGL(gather_columns_kernel<true>, dim3(div_up(ix.len, kColTile)), dim3(kColBlock),
                       0, c.stream(), cs, (const uint32_t*)ix.ptr, ix.len, v[0].len,
                       aligned16(ix.ptr));
  else
    hipLaunchKernelGGL(gather_columns_kernel<false>, dim3(div_up(ix.len, kColTile)),
                       dim3(kColBlock), 0, c.stream(), cs, (const uint32_t*)ix.ptr, ix.len,
                       v[0].len, aligned16(ix.ptr));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    release_all(out, ncols);
    set_error(MSV_ERR_LAUNCH, 0, 0, "gather_columns launch failed: %s", hipGetErrorString(e));
    return MSV_ERR_LAUNCH;
  }
  return MSV_OK;
}

msv_status msv_scatter_columns(msv_executor* ex, const msv_slice* values, uint32_t ncols,
                               msv_slice indices, const msv_slice* outputs) {
  clear_error();
  View v[kMaxColumns], o[kMaxColumns], ix;
  MSV_TRY(resolve_columns(ex, values, ncols, v));
  MSV_TRY(resolve_columns(ex, outputs, ncols, o));
  MSV_TRY(resolve(ex, indices, &ix));
  if (ix.dtype != MSV_U32) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "indices must be u32 (MIndex)");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (v[0].len != ix.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, v[0].len, ix.len, "input lengths differ: %u != %u",
              v[0].len, ix.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  ColumnSet cs{};
  cs.ncols = ncols;
  bool aligned = aligned16(ix.ptr);
  bool binned = false;
  for (uint32_t c = 0; c < ncols; ++c) {
    if (o[c].dtype != v[c].dtype) {
      set_error(MSV_ERR_INVALID_ARGUMENT, c, 0, "output column %u dtype differs from values", c);
      return MSV_ERR_INVALID_ARGUMENT;
    }
    cs.in[c] = v[c].ptr;
    cs.out[c] = o[c].ptr;
    cs.words[c] = (uint32_t)(dtype_size(v[c].dtype) / 4);
    aligned = aligned && aligned16(v[c].ptr);
    binned = binned || scatter_binned(v[c].dtype, ix.len, o[c].len);
  }
  if (ix.len == 0) return MSV_OK;
  Ctx c(ex);
  if (binned) {
    // Outputs too large for the caches: column by column in the binned form.
    for (uint32_t k = 0; k < ncols; ++k)
      MSV_TRY(scatter_run(c, v[k].dtype, v[k].ptr, (const uint32_t*)ix.ptr, ix.len, o[k].ptr,
                          o[k].len));
    return MSV_OK;
  }
  hipLaunchKernelGGL(scatter_columns_kernel, dim3(div_up(ix.len, kColTile)), dim3(kColBlock), 0,
                     c.stream(), cs, (const uint32_t*)ix.ptr, ix.len, o[0].len, aligned);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

msv_status msv_copy_where_columns(msv_executor* ex, const msv_slice* input, uint32_t ncols,
                                  msv_slice stencil, const msv_predicate* pred, int32_t invert,
                                  msv_vec** out) {
  clear_error();
  View v[kMaxColumns], st;
  MSV_TRY(resolve_columns(ex, input, ncols, v));
  MSV_TRY(resolve(ex, stencil, &st));
  if (!out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null outputs");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (v[0].len != st.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, v[0].len, st.len, "input lengths differ: %u != %u",
              v[0].len, st.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  if (!pred && st.dtype != MSV_U32) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "stencil must be u32 flags (MFlag)");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (pred && pred->kind > PRED_BITS_EQ) {
    set_error(MSV_ERR_INVALID_ARGUMENT, pred->kind, 0, "unknown predicate kind %u", pred->kind);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t c = 0; c < ncols; ++c) out[c] = nullptr;
  ColumnSet cs{};
  cs.ncols = ncols;
  bool aligned = aligned16(st.ptr);
  for (uint32_t c = 0; c < ncols; ++c) {
    msv_status s = new_vec(ex, v[c].dtype, v[c].len, true, &out[c]);
    if (s != MSV_OK) {
      release_all(out, c);
      return s;
    }
    cs.in[c] = v[c].ptr;
    cs.out[c] = out[c]->ptr;
    cs.out_len[c] = out[c]->d_len;
    cs.words[c] = (uint32_t)(dtype_size(v[c].dtype) / 4);
    aligned = aligned && aligned16(v[c].ptr);
  }
  const uint32_t n = v[0].len;
  msv_status s = MSV_OK;
  {
    Ctx c(ex);
    if (n == 0) {
      // Empty input: no tile writes the lengths (a kernel, not a memset, so
      // that the call can be captured in a graph).
      hipLaunchKernelGGL(column_lengths_zero_kernel, dim3(1), dim3(kMaxColumns), 0, c.stream(), cs);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        set_error(MSV_ERR_LAUNCH, 0, 0, "%s", hipGetErrorString(e));
        s = MSV_ERR_LAUNCH;
      }
    } else {
      const bool inv = invert != 0;
      switch (pred ? st.dtype : MSV_U32) {
        case MSV_U32: s = select_columns_launch<uint32_t>(c, cs, st.ptr, pred, inv, n, aligned); break;
        case MSV_I32: s = select_columns_launch<int32_t>(c, cs, st.ptr, pred, inv, n, aligned); break;
        case MSV_F32: s = select_columns_launch<float>(c, cs, st.ptr, pred, inv, n, aligned); break;
        case MSV_U64: s = select_columns_launch<uint64_t>(c, cs, st.ptr, pred, inv, n, aligned); break;
        case MSV_I64: s = select_columns_launch<int64_t>(c, cs, st.ptr, pred, inv, n, aligned); break;
        default: s = select_columns_launch<double>(c, cs, st.ptr, pred, inv, n, aligned); break;
      }
    }
  }
  if (s != MSV_OK) release_all(out, ncols);
  return s;
}

}  // extern "C"
