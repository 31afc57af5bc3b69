// vector::gather / scatter and seeded synthetic inputs for gfx950.
//
// gather: out[i] = values[idx[i]]   (massively/src/core/indexed.rs:16-196)
// scatter: out[idx[i]] = values[i]  (massively/src/core/scatter.rs:8-21)
// Each thread owns ROWS 16-byte index vectors laid out row-striped like the
// scan tiles, so index loads and the sequential-side stores are full-wave
// dwordx4 accesses and each thread keeps 16 independent random accesses in
// flight. The reference leaves out-of-range indices undefined; here a gather
// reads zero for them and a scatter drops them, so no index can fault.
#include <algorithm>

#include "device_common.hpp"
#include "msv_internal.hpp"

namespace msv {

constexpr int kMoveBlock = 256;
constexpr int kMoveRows = 4;  // 4 x uint4 of indices = 16 indices per thread
constexpr uint32_t kMoveTile = kMoveBlock * kMoveRows * 4;

template <typename T>
__global__ __launch_bounds__(kMoveBlock) void gather_kernel(const T* __restrict__ values,
                                                            uint32_t vlen,
                                                            const uint32_t* __restrict__ idx,
                                                            T* __restrict__ out, uint32_t n,
                                                            bool aligned) {
  const uint32_t tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * kMoveTile;
  const bool full = aligned && base + kMoveTile <= n;
  uint32_t ix[kMoveRows][4];
  if (full) {
#pragma unroll
    for (int r = 0; r < kMoveRows; ++r) {
      const uint4 q = *reinterpret_cast<const uint4*>(idx + base + ((size_t)r * kMoveBlock + tid) * 4);
      ix[r][0] = q.x; ix[r][1] = q.y; ix[r][2] = q.z; ix[r][3] = q.w;
    }
  } else {
#pragma unroll
    for (int r = 0; r < kMoveRows; ++r)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t i = base + ((size_t)r * kMoveBlock + tid) * 4 + k;
        ix[r][k] = i < n ? idx[i] : 0xffffffffu;
      }
  }
  T v[kMoveRows][4];
#pragma unroll
  for (int r = 0; r < kMoveRows; ++r)
#pragma unroll
    for (int k = 0; k < 4; ++k) v[r][k] = ix[r][k] < vlen ? values[ix[r][k]] : (T)0;
#pragma unroll
  for (int r = 0; r < kMoveRows; ++r) {
    const size_t o = base + ((size_t)r * kMoveBlock + tid) * 4;
    if (full) {
      if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<uint4*>(out + o) =
            make_uint4(bit_cast<uint32_t>(v[r][0]), bit_cast<uint32_t>(v[r][1]),
                       bit_cast<uint32_t>(v[r][2]), bit_cast<uint32_t>(v[r][3]));
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[o + k] = v[r][k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (o + k < n) out[o + k] = v[r][k];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kMoveBlock) void scatter_kernel(const T* __restrict__ values,
                                                             const uint32_t* __restrict__ idx,
                                                             uint32_t n, T* __restrict__ out,
                                                             uint32_t olen) {
  const uint32_t tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * kMoveTile;
#pragma unroll
  for (int r = 0; r < kMoveRows; ++r)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t i = base + ((size_t)r * kMoveBlock + tid) * 4 + k;
      if (i < n) {
        const uint32_t d = idx[i];
        if (d < olen) out[d] = values[i];
      }
    }
}

// ---- binned scatter ----------------------------------------------------------
// A scatter into an output far larger than the caches writes one partial
// cache line per element (a 4-byte write costs a whole line of HBM traffic).
// The binned form partitions the (index, value) pairs by destination window
// (256 windows over the output; one look-back pass like a radix digit pass,
// unordered inside a window: ranks come from LDS atomics), then writes them
// window by window. Workgroups go to the 8 XCDs round-robin, and each XCD has
// its own L2, so the partition orders the windows XCD by XCD (window w goes to
// group w % 8) and the write pass keeps every block on its own group's pairs
// (block b writes group b % 8): each L2 then merges the writes of the one or
// two windows its blocks have in flight into whole lines before they reach
// HBM.
constexpr int kBinBlock = 512;
constexpr uint32_t kBinBuckets = 256;  // one look-back thread per window
constexpr uint32_t kXcds = 8;
// Pairs per thread: 16 for 4-byte values, 8 for 8-byte values (the tile's
// pairs are staged in LDS: 64 or 48 KiB).
template <typename T>
constexpr int bin_ipt() { return sizeof(T) == 4 ? 16 : 8; }

// Partition position of window w: XCD group w % 8 first, then w / 8.
__device__ __forceinline__ uint32_t bin_slot(uint32_t w) {
  return (w % kXcds) * (kBinBuckets / kXcds) + w / kXcds;
}

__global__ void bin_fill_kernel(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// hist[slot] = number of indices < olen whose window (index >> shift) has
// partition position `slot`. Four 16-byte index loads per thread per step.
__global__ __launch_bounds__(1024) void bin_hist_kernel(const uint32_t* __restrict__ idx,
                                                        uint32_t n, uint32_t olen, uint32_t shift,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[kBinBuckets];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < kBinBuckets; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  const bool aligned = ((uintptr_t)idx & 15u) == 0;
  const size_t step = (size_t)gridDim.x * blockDim.x * 16;
  for (size_t b0 = (size_t)blockIdx.x * blockDim.x * 16; b0 < n; b0 += step) {
    uint32_t d[16];
    if (aligned && b0 + (size_t)blockDim.x * 16 <= n) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint4 q = *reinterpret_cast<const uint4*>(idx + b0 + ((size_t)r * blockDim.x + tid) * 4);
        d[4 * r] = q.x; d[4 * r + 1] = q.y; d[4 * r + 2] = q.z; d[4 * r + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const size_t i = b0 + ((size_t)r * blockDim.x + tid) * 4 + k;
          d[4 * r + k] = i < n ? idx[i] : 0xffffffffu;
        }
    }
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (d[j] < olen) atomicAdd(&s_h[bin_slot(d[j] >> shift)], 1u);
  }
  __syncthreads();
  for (uint32_t i = tid; i < kBinBuckets; i += blockDim.x)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// base[s] = exclusive scan of hist, base[512] = total; zeroes the tile counter.
__global__ __launch_bounds__(kBinBuckets) void bin_plan_kernel(const uint32_t* __restrict__ hist,
                                                               uint32_t* __restrict__ base,
                                                               uint32_t* __restrict__ counter) {
  __shared__ uint32_t s_w[kBinBuckets / kWave];
  const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
  const uint32_t c = hist[t];
  const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(c);
  if (lane == 63) s_w[wid] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (uint32_t w = 0; w < kBinBuckets / kWave; ++w) {
    off += w < wid ? s_w[w] : 0u;
    tot += s_w[w];
  }
  base[t] = off + incl - c;
  if (t == 0) {
    base[kBinBuckets] = tot;
    *counter = 0;
  }
}

// Partitions the pairs of one tile by window: a block claims the next tile
// id, ranks its pairs inside their windows with LDS atomics, publishes its
// window counts at once, looks back over earlier tiles for each window's
// prefix (one thread per window, four predecessors per step), and writes each
// window's run of pairs, compacted through LDS, to base[slot] + prefix.
template <typename T>
__global__ __launch_bounds__(kBinBlock) void bin_partition_kernel(
    const uint32_t* __restrict__ idx, const T* __restrict__ vals, uint32_t n, uint32_t olen,
    uint32_t shift, const uint32_t* __restrict__ base, uint32_t* __restrict__ counter,
    uint64_t* __restrict__ status, uint32_t gen, uint32_t* __restrict__ out_idx,
    T* __restrict__ out_vals) {
  constexpr int IPT = bin_ipt<T>();
  constexpr uint32_t TILE = (uint32_t)kBinBlock * IPT;
  __shared__ uint32_t s_idx[TILE];
  __shared__ T s_val[TILE];
  __shared__ uint32_t s_cnt[kBinBuckets];
  __shared__ uint32_t s_start[kBinBuckets];
  __shared__ uint32_t s_goff[kBinBuckets];
  __shared__ uint32_t s_w[kBinBuckets / kWave];
  __shared__ uint32_t s_tile;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const bool window_thread = tid < kBinBuckets;
  if (window_thread) s_cnt[tid] = 0;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t t0 = (size_t)tile * TILE;
  uint32_t d[IPT], slot[IPT], rank[IPT];
  T v[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const size_t i = t0 + (size_t)j * kBinBlock + tid;
    d[j] = i < n ? idx[i] : 0xffffffffu;
    v[j] = i < n ? vals[i] : T();
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    slot[j] = d[j] < olen ? bin_slot(d[j] >> shift) : kBinBuckets;
    rank[j] = slot[j] < kBinBuckets ? atomicAdd(&s_cnt[slot[j]], 1u) : 0u;
  }
  __syncthreads();
  uint64_t* my = status + (size_t)tile * kBinBuckets + tid;
  uint32_t cnt = 0, incl = 0;
  if (window_thread) {
    cnt = s_cnt[tid];
    st_agent(my, pack_status(gen, tile == 0 ? kFlagInc : kFlagAgg, 0, cnt));
    incl = wave_inclusive_scan<uint32_t, OP_ADD>(cnt);
    if (lane == 63) s_w[wid] = incl;
  }
  __syncthreads();
  if (window_thread) {
    uint32_t off = 0;
    for (uint32_t w = 0; w < wid; ++w) off += s_w[w];
    const uint32_t start = off + incl - cnt;
    s_start[tid] = start;
    uint32_t prefix = 0;
    if (tile > 0) {
      int64_t q = (int64_t)tile - 1;
      for (;;) {
        uint64_t w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          w[k] = q - k >= 0 ? ld_agent(status + (size_t)(q - k) * kBinBuckets + tid) : 0ull;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (done || used < k) continue;
          const uint32_t f = status_flag(w[k], gen);
          if (f == 0) continue;
          prefix += (uint32_t)w[k];
          used = k + 1;
          done = f == kFlagInc;
        }
        if (done) break;
        q -= used;
        if (used == 0) __builtin_amdgcn_s_sleep(1);
      }
      st_agent(my, pack_status(gen, kFlagInc, 0, prefix + cnt));
    }
    s_goff[tid] = base[tid] + prefix - start;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if (slot[j] < kBinBuckets) {
      const uint32_t pos = s_start[slot[j]] + rank[j];
      s_idx[pos] = d[j];
      s_val[pos] = v[j];
    }
  }
  __syncthreads();
  uint32_t total = 0;
  for (uint32_t w = 0; w < kBinBuckets / kWave; ++w) total += s_w[w];
  for (uint32_t i = tid; i < total; i += kBinBlock) {
    const uint32_t dst = s_idx[i];
    const uint32_t o = s_goff[bin_slot(dst >> shift)] + i;
    out_idx[o] = dst;
    out_vals[o] = s_val[i];
  }
}

// Writes the partitioned pairs: block b walks the tiles of XCD group b % 8
// (the pairs of windows w % 8 == b % 8, in window order), grid-strided over
// the blocks of that group.
template <typename T>
__global__ __launch_bounds__(kMoveBlock) void bin_scatter_kernel(const uint32_t* __restrict__ pidx,
                                                                 const T* __restrict__ pvals,
                                                                 const uint32_t* __restrict__ base,
                                                                 T* __restrict__ out) {
  const uint32_t g = blockIdx.x % kXcds;
  const uint32_t per_group = gridDim.x / kXcds;
  const size_t g0 = base[g * (kBinBuckets / kXcds)];
  const size_t g1 = base[(g + 1) * (kBinBuckets / kXcds)];
  for (size_t t0 = g0 + (size_t)(blockIdx.x / kXcds) * kMoveTile; t0 < g1;
       t0 += (size_t)per_group * kMoveTile) {
    uint32_t dd[kMoveRows * 4];
    T vv[kMoveRows * 4];
#pragma unroll
    for (int r = 0; r < kMoveRows * 4; ++r) {
      // The pair stream is read once: non-temporal, so the L2 keeps the
      // output lines being filled.
      const size_t i = t0 + (size_t)r * kMoveBlock + threadIdx.x;
      dd[r] = i < g1 ? __builtin_nontemporal_load(pidx + i) : 0xffffffffu;
      vv[r] = i < g1 ? __builtin_nontemporal_load(pvals + i) : T();
    }
#pragma unroll
    for (int r = 0; r < kMoveRows * 4; ++r)
      if (dd[r] != 0xffffffffu) out[dd[r]] = vv[r];
  }
}

// ---- seeded streams (massively/src/api/util/random.rs:12-22,331-396) ------
// pcg_hash32 and seed_key: msv_internal.hpp.

__global__ void uniform_u32_kernel(uint32_t* __restrict__ out, uint32_t n, uint32_t start,
                                   uint32_t key, uint32_t lo, uint32_t span) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = pcg_hash32((start + i) ^ key);
  out[i] = (span == 0xffffffffu) ? v : lo + v % (span + 1u);
}

__global__ void uniform_u64_kernel(uint64_t* __restrict__ out, uint32_t n, uint32_t start,
                                   uint32_t key, uint64_t lo, uint64_t span) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = start + i;
  const uint64_t v = (uint64_t)pcg_hash32(p ^ key) | ((uint64_t)pcg_hash32(p ^ key ^ 0x9e3779b9u) << 32);
  out[i] = (span == ~0ull) ? v : lo + v % (span + 1ull);
}

__global__ void iota_u32_kernel(uint32_t* __restrict__ out, uint32_t n, uint32_t start) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = start + i;
}

// ---- host ------------------------------------------------------------------
msv_status gather_run(Ctx& c, msv_dtype dt, const void* values, uint32_t vlen,
                      const uint32_t* idx, void* out, uint32_t n) {
  if (n == 0) return MSV_OK;
  const uint32_t blocks = div_up(n, kMoveTile);
  const bool aligned = (((uintptr_t)idx | (uintptr_t)out) & 15u) == 0;
  if (dtype_size(dt) == 4)
    hipLaunchKernelGGL(gather_kernel<uint32_t>, dim3(blocks), dim3(kMoveBlock), 0, c.stream(),
                       (const uint32_t*)values, vlen, idx, (uint32_t*)out, n, aligned);
  else
    hipLaunchKernelGGL(gather_kernel<uint64_t>, dim3(blocks), dim3(kMoveBlock), 0, c.stream(),
                       (const uint64_t*)values, vlen, idx, (uint64_t*)out, n, aligned);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

template <typename T>
static msv_status binned_scatter(Ctx& c, const T* values, const uint32_t* idx, uint32_t n, T* out,
                                 uint32_t olen) {
  uint32_t shift = 0;
  while (((uint64_t)(olen - 1) >> shift) >= kBinBuckets) ++shift;
  const uint32_t ntiles = div_up(n, (uint32_t)kBinBlock * bin_ipt<T>());
  uint32_t* hist = (uint32_t*)c.scratch(sizeof(uint32_t) * (2 * kBinBuckets + 2));
  uint32_t* pidx = (uint32_t*)c.scratch(sizeof(uint32_t) * (size_t)n);
  T* pvals = (T*)c.scratch(sizeof(T) * (size_t)n);
  if (!hist || !pidx || !pvals) return MSV_ERR_OUT_OF_MEMORY;
  uint32_t* base = hist + kBinBuckets;       // kBinBuckets + 1 words
  uint32_t* counter = base + kBinBuckets + 1;
  uint64_t *lb_counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback((size_t)ntiles * kBinBuckets, &lb_counter, &status, &gen));
  hipLaunchKernelGGL(bin_fill_kernel, dim3(div_up(kBinBuckets, 256)), dim3(256), 0, c.stream(),
                     hist, kBinBuckets);
  const uint32_t hist_blocks =
      std::min<uint32_t>(div_up(n, 1024u * 16u), (uint32_t)c.num_cus() * 4u);
  hipLaunchKernelGGL(bin_hist_kernel, dim3(hist_blocks), dim3(1024), 0, c.stream(), idx, n, olen,
                     shift, hist);
  hipLaunchKernelGGL(bin_plan_kernel, dim3(1), dim3(kBinBuckets), 0, c.stream(), hist, base,
                     counter);
  hipLaunchKernelGGL(bin_partition_kernel<T>, dim3(ntiles), dim3(kBinBlock), 0, c.stream(), idx,
                     values, n, olen, shift, base, counter, status, gen, pidx, pvals);
  // A multiple of 8 blocks, a few per CU: blocks stay on their XCD's group.
  const uint32_t grid = kXcds * std::max<uint32_t>(1u, std::min<uint32_t>(
                                    div_up(n, kMoveTile * kXcds), (uint32_t)c.num_cus() * 4u / kXcds));
  hipLaunchKernelGGL(bin_scatter_kernel<T>, dim3(grid), dim3(kMoveBlock), 0, c.stream(), pidx,
                     pvals, base, out);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

// Outputs from this size up take the binned form (the direct scatter's
// partial-line writes stay in the caches below it).
constexpr uint64_t kBinScatterMinBytes = 64ull << 20;

bool scatter_binned(msv_dtype dt, uint32_t n, uint32_t olen) {
  return (uint64_t)olen * dtype_size(dt) >= kBinScatterMinBytes && n >= (1u << 20);
}

msv_status scatter_run(Ctx& c, msv_dtype dt, const void* values, const uint32_t* idx,
                       uint32_t n, void* out, uint32_t olen) {
  if (n == 0) return MSV_OK;
  if (scatter_binned(dt, n, olen)) {
    if (dtype_size(dt) == 4)
      return binned_scatter<uint32_t>(c, (const uint32_t*)values, idx, n, (uint32_t*)out, olen);
    return binned_scatter<uint64_t>(c, (const uint64_t*)values, idx, n, (uint64_t*)out, olen);
  }
  const uint32_t blocks = div_up(n, kMoveTile);
  if (dtype_size(dt) == 4)
    hipLaunchKernelGGL(scatter_kernel<uint32_t>, dim3(blocks), dim3(kMoveBlock), 0, c.stream(),
                       (const uint32_t*)values, idx, n, (uint32_t*)out, olen);
  else
    hipLaunchKernelGGL(scatter_kernel<uint64_t>, dim3(blocks), dim3(kMoveBlock), 0, c.stream(),
                       (const uint64_t*)values, idx, n, (uint64_t*)out, olen);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

msv_status uniform_run(Ctx& c, msv_dtype dt, void* out, uint32_t n, uint64_t seed) {
  if (n == 0) return MSV_OK;
  const dim3 grid((n + 255) / 256), block(256);
  switch (dt) {
    case MSV_U32:
    case MSV_I32:
      hipLaunchKernelGGL(uniform_u32_kernel, grid, block, 0, c.stream(), (uint32_t*)out, n, 0u,
                         seed_key(seed), 0u, 0xffffffffu);
      break;
    case MSV_F32:  // bit patterns of [0.5, 1)
      hipLaunchKernelGGL(uniform_u32_kernel, grid, block, 0, c.stream(), (uint32_t*)out, n, 0u,
                         seed_key(seed), 0x3F000000u, 0x007FFFFFu);
      break;
    case MSV_U64:
    case MSV_I64:
      hipLaunchKernelGGL(uniform_u64_kernel, grid, block, 0, c.stream(), (uint64_t*)out, n, 0u,
                         seed_key(seed), 0ull, ~0ull);
      break;
    case MSV_F64:  // bit patterns of [0.5, 1)
      hipLaunchKernelGGL(uniform_u64_kernel, grid, block, 0, c.stream(), (uint64_t*)out, n, 0u,
                         seed_key(seed), 0x3FE0000000000000ull, 0x000FFFFFFFFFFFFFull);
      break;
  }
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

msv_status iota_run(Ctx& c, uint32_t* out, uint32_t n, uint32_t start) {
  if (n == 0) return MSV_OK;
  hipLaunchKernelGGL(iota_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream(), out, n,
                     start);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}


// ---- vector::scatter_reduce ----------------------------------------------------
// out[d] = op(out[d], init (+) the values whose index is d), for destinations
// present in `indices` (api/algorithm/scatter.rs:337-420; the reference sorts
// the proposals by index, reduces each run and combines each result with its
// destination, core/scatter_reduce.rs:19-150). `init` must be op's identity,
// as the reference requires. Integers: one pass of device-scope atomics, exact
// whatever the order. Floats: the reference's shape (stable radix sort of
// (index, value), reduce_by_key, one combine per destination), so results are
// reproducible bit for bit.
template <typename T, int OP>
__device__ inline void atomic_combine(T* p, T v) {
  if constexpr (OP == OP_ADD) {
    if constexpr (sizeof(T) == 4) atomicAdd((unsigned int*)p, (unsigned int)v);
    else atomicAdd((unsigned long long*)p, (unsigned long long)v);
  } else if constexpr (OP == OP_XOR) {
    if constexpr (sizeof(T) == 4) atomicXor((unsigned int*)p, (unsigned int)v);
    else atomicXor((unsigned long long*)p, (unsigned long long)v);
  } else if constexpr (OP == OP_MIN) {
    if constexpr (std::is_same<T, int64_t>::value) atomicMin((long long*)p, (long long)v);
    else if constexpr (std::is_same<T, uint64_t>::value) atomicMin((unsigned long long*)p, (unsigned long long)v);
    else atomicMin(p, v);
  } else {
    if constexpr (std::is_same<T, int64_t>::value) atomicMax((long long*)p, (long long)v);
    else if constexpr (std::is_same<T, uint64_t>::value) atomicMax((unsigned long long*)p, (unsigned long long)v);
    else atomicMax(p, v);
  }
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void scatter_reduce_atomic_kernel(const T* __restrict__ vals,
                                                                    const uint32_t* __restrict__ idx,
                                                                    uint32_t n, T* out,
                                                                    uint32_t out_len, bool aligned) {
  const size_t groups = ((size_t)n + 3) / 4;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups;
       g += (size_t)gridDim.x * blockDim.x) {
    const bool full = aligned && g * 4 + 4 <= n;
    uint32_t d[4];
    T v[4];
    load4<uint32_t>(idx, g * 4, n, full, d);
    load4<T>(vals, g * 4, n, full, v);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (g * 4 + k < n && d[k] < out_len) atomic_combine<T, OP>(out + d[k], v[k]);
  }
}

// One proposal per destination (keys distinct): out[d] = op(out[d], proposal).
template <typename T, int OP>
__global__ void scatter_combine_kernel(const uint32_t* __restrict__ keys, const T* __restrict__ vals,
                                       const uint32_t* __restrict__ count, T* __restrict__ out,
                                       uint32_t out_len) {
  const uint32_t m = *count;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
    const uint32_t d = keys[j];
    if (d < out_len) out[d] = Combine<T, OP>::apply(out[d], vals[j]);
  }
}

template <typename T, int OP>
static msv_status scatter_reduce_typed(Ctx& c, msv_dtype dt, msv_op op, const void* vals,
                                       const uint32_t* idx, uint32_t n, const void* init,
                                       void* out, uint32_t out_len) {
  if (n == 0 || out_len == 0) return MSV_OK;
  if constexpr (!IsFloat<T>::value) {
    (void)dt;
    (void)op;
    (void)init;
    const bool aligned = (((uintptr_t)vals | (uintptr_t)idx) & 15u) == 0;
    const uint32_t groups = (n + 3) / 4;
    const uint32_t blocks = min((groups + 255) / 256, 8192u);
    hipLaunchKernelGGL((scatter_reduce_atomic_kernel<T, OP>), dim3(blocks), dim3(256), 0,
                       c.stream(), (const T*)vals, idx, n, (T*)out, out_len, aligned);
    MSV_HIP(hipGetLastError());
    return MSV_OK;
  } else {
    uint32_t* sk = (uint32_t*)c.scratch((size_t)n * 4);
    T* sv = (T*)c.scratch((size_t)n * sizeof(T));
    uint32_t* rk = (uint32_t*)c.scratch((size_t)n * 4);
    T* rv = (T*)c.scratch((size_t)n * sizeof(T));
    uint32_t* cnt = (uint32_t*)c.scratch(8);
    if (!sk || !sv || !rk || !rv || !cnt) return MSV_ERR_OUT_OF_MEMORY;
    MSV_TRY(radix_sort_run(c, MSV_U32, 0, sizeof(T), idx, vals, n, sk, sv));
    MSV_TRY(runs_scan_run(c, MSV_U32, sk, dt, op, RUNS_REDUCE, sv, n, rv, init, rk, cnt, cnt + 1));
    const uint32_t blocks = min((n + 255) / 256, 8192u);
    hipLaunchKernelGGL((scatter_combine_kernel<T, OP>), dim3(blocks), dim3(256), 0, c.stream(), rk,
                       rv, cnt, (T*)out, out_len);
    MSV_HIP(hipGetLastError());
    return MSV_OK;
  }
}

template <typename T>
static msv_status scatter_reduce_by_op(Ctx& c, msv_dtype dt, msv_op op, const void* vals,
                                       const uint32_t* idx, uint32_t n, const void* init,
                                       void* out, uint32_t out_len) {
  switch (op) {
    case MSV_OP_ADD: return scatter_reduce_typed<T, OP_ADD>(c, dt, op, vals, idx, n, init, out, out_len);
    case MSV_OP_MIN: return scatter_reduce_typed<T, OP_MIN>(c, dt, op, vals, idx, n, init, out, out_len);
    case MSV_OP_MAX: return scatter_reduce_typed<T, OP_MAX>(c, dt, op, vals, idx, n, init, out, out_len);
    case MSV_OP_XOR:
      if constexpr (!IsFloat<T>::value)
        return scatter_reduce_typed<T, OP_XOR>(c, dt, op, vals, idx, n, init, out, out_len);
      break;
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator %d is not defined for this dtype", (int)op);
  return MSV_ERR_INVALID_ARGUMENT;
}

}  // namespace msv

using namespace msv;

static msv_status check_index_slice(const View& v) {
  if (v.dtype != MSV_U32) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "indices must be u32 (MIndex)");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return MSV_OK;
}

extern "C" {

msv_status msv_gather(msv_executor* ex, msv_slice values, msv_slice indices, msv_vec** out) {
  clear_error();
  View v, ix;
  MSV_TRY(resolve(ex, values, &v));
  MSV_TRY(resolve(ex, indices, &ix));
  MSV_TRY(check_index_slice(ix));
  MSV_TRY(new_vec(ex, v.dtype, ix.len, false, out));
  Ctx c(ex);
  msv_status s = gather_run(c, v.dtype, v.ptr, v.len, (const uint32_t*)ix.ptr, (*out)->ptr, ix.len);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

msv_status msv_scatter(msv_executor* ex, msv_slice values, msv_slice indices, msv_slice output) {
  clear_error();
  View v, ix, o;
  MSV_TRY(resolve(ex, values, &v));
  MSV_TRY(resolve(ex, indices, &ix));
  MSV_TRY(resolve(ex, output, &o));
  MSV_TRY(check_index_slice(ix));
  if (v.len != ix.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, v.len, ix.len, "input lengths differ: %u != %u", v.len,
              ix.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  if (o.dtype != v.dtype) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "output dtype differs from values");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  return scatter_run(c, v.dtype, v.ptr, (const uint32_t*)ix.ptr, v.len, o.ptr, o.len);
}

msv_status msv_scatter_reduce(msv_executor* ex, msv_slice values, msv_slice indices,
                              const void* init, msv_op op, msv_slice output) {
  clear_error();
  View v, ix, o;
  MSV_TRY(resolve(ex, values, &v));
  MSV_TRY(resolve(ex, indices, &ix));
  MSV_TRY(resolve(ex, output, &o));
  MSV_TRY(check_index_slice(ix));
  if (v.len != ix.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, v.len, ix.len, "input lengths differ: %u != %u", v.len,
              ix.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  if (o.dtype != v.dtype || !init) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "output dtype differs from values, or null init");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  const uint32_t* idx = (const uint32_t*)ix.ptr;
  switch (v.dtype) {
    case MSV_U32: return scatter_reduce_by_op<uint32_t>(c, v.dtype, op, v.ptr, idx, v.len, init, o.ptr, o.len);
    case MSV_I32: return scatter_reduce_by_op<int32_t>(c, v.dtype, op, v.ptr, idx, v.len, init, o.ptr, o.len);
    case MSV_F32: return scatter_reduce_by_op<float>(c, v.dtype, op, v.ptr, idx, v.len, init, o.ptr, o.len);
    case MSV_U64: return scatter_reduce_by_op<uint64_t>(c, v.dtype, op, v.ptr, idx, v.len, init, o.ptr, o.len);
    case MSV_I64: return scatter_reduce_by_op<int64_t>(c, v.dtype, op, v.ptr, idx, v.len, init, o.ptr, o.len);
    case MSV_F64: return scatter_reduce_by_op<double>(c, v.dtype, op, v.ptr, idx, v.len, init, o.ptr, o.len);
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad dtype");
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status msv_random_uniform_u32(msv_executor* ex, uint32_t min, uint32_t max, uint64_t seed,
                                  msv_slice out) {
  clear_error();
  View o;
  MSV_TRY(resolve(ex, out, &o));
  if (o.dtype != MSV_U32 || min > max) {
    set_error(MSV_ERR_LAUNCH, 0, 0, "random distribution range must satisfy min <= max (u32 out)");
    return MSV_ERR_LAUNCH;
  }
  if (o.len == 0) return MSV_OK;
  Ctx c(ex);
  hipLaunchKernelGGL(uniform_u32_kernel, dim3((o.len + 255) / 256), dim3(256), 0, c.stream(),
                     (uint32_t*)o.ptr, o.len, 0u, seed_key(seed), min, max - min);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

msv_status msv_random_uniform_u64(msv_executor* ex, uint64_t min, uint64_t max, uint64_t seed,
                                  msv_slice out) {
  clear_error();
  View o;
  MSV_TRY(resolve(ex, out, &o));
  if (o.dtype != MSV_U64 || min > max) {
    set_error(MSV_ERR_LAUNCH, 0, 0, "random distribution range must satisfy min <= max (u64 out)");
    return MSV_ERR_LAUNCH;
  }
  if (o.len == 0) return MSV_OK;
  Ctx c(ex);
  hipLaunchKernelGGL(uniform_u64_kernel, dim3((o.len + 255) / 256), dim3(256), 0, c.stream(),
                     (uint64_t*)o.ptr, o.len, 0u, seed_key(seed), min, max - min);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

msv_status msv_iota_u32(msv_executor* ex, uint32_t start, msv_slice out) {
  clear_error();
  View o;
  MSV_TRY(resolve(ex, out, &o));
  if (o.dtype != MSV_U32) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "iota output must be u32");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  return iota_run(c, (uint32_t*)o.ptr, o.len, start);
}

}  // extern "C"
