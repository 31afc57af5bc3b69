// Timing proxy of the reference's GPU radix schedule (not part of libmsv; a
// measuring stick for the benchmark, since the reference's CubeCL-HIP build
// cannot be made offline). It follows the kernel schedule of
// massively/src/core/radix.rs:120-498 for sort_by_key of u32 keys + u32
// values: 4-bit LSD digits (8 passes), per pass
//   1. a 256-item block digit sort reading the keys through the current
//      permutation (four stable binary splits in LDS), writing the block's
//      permutation, per-item (digit, rank) metadata and a 16 x blocks
//      histogram (radix.rs:120-279);
//   2. a bucket-prefix launch of 16 workgroups, each summing and scanning one
//      bucket's row of the histogram (radix.rs:281-335);
//   3. a one-thread bucket-offset scan (radix.rs:337-346);
//   4. a one-item-per-thread scatter of the permutation (radix.rs:348-367);
// then keys and values are gathered through the final permutation. The
// splits use wave ballots instead of shuffle ladders, which only makes the
// proxy faster than the schedule it models.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kItems = 256;  // RADIX_BLOCK_ITEMS
constexpr int kBits = 4;     // RADIX_BITS
constexpr int kBuckets = 1 << kBits;
constexpr int kScatterBlock = 256;

__global__ __launch_bounds__(kItems) void block_digit_sort(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ perm, uint32_t n, uint32_t shift,
    uint32_t nblocks, uint32_t identity, uint32_t* __restrict__ block_perm,
    uint32_t* __restrict__ meta, uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_perm[2][kItems];
  __shared__ uint32_t s_dig[2][kItems];
  __shared__ uint32_t s_wave[kItems / 64 + 1];
  __shared__ uint32_t s_first[kBuckets];
  const uint32_t block = blockIdx.x, local = threadIdx.x, lane = local & 63u, wave = local >> 6;
  const uint32_t start = block * kItems;
  const uint32_t end = min(n, start + kItems);
  const uint32_t cnt = end > start ? end - start : 0u;
  const uint32_t global = start + local;
  const bool valid = global < end;
  uint32_t src = 0, dig = 0;
  if (valid) {
    src = identity ? global : perm[global];
    dig = (keys[src] >> shift) & (kBuckets - 1u);
  }
  s_perm[0][local] = src;
  s_dig[0][local] = dig;
  __syncthreads();
  int cur = 0;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  for (int b = 0; b < kBits; ++b) {
    const uint32_t d = s_dig[cur][local];
    const bool zero = valid && ((d >> b) & 1u) == 0u;
    const uint64_t z = __ballot(zero);
    if (lane == 0) s_wave[wave] = (uint32_t)__builtin_popcountll(z);
    __syncthreads();
    if (local == 0) {
      uint32_t run = 0;
      for (uint32_t w = 0; w < kItems / 64; ++w) {
        const uint32_t c = s_wave[w];
        s_wave[w] = run;
        run += c;
      }
      s_wave[kItems / 64] = run;
    }
    __syncthreads();
    if (valid) {
      const uint32_t ex = s_wave[wave] + (uint32_t)__builtin_popcountll(z & lt);
      const uint32_t zeros = s_wave[kItems / 64];
      const uint32_t dst = zero ? ex : zeros + local - ex;
      s_perm[cur ^ 1][dst] = s_perm[cur][local];
      s_dig[cur ^ 1][dst] = d;
    }
    __syncthreads();
    cur ^= 1;
  }
  const uint32_t sd = s_dig[cur][local];
  if (local < kBuckets) {
    s_first[local] = cnt;
    hist[local * nblocks + block] = 0u;
  }
  __syncthreads();
  if (valid && (local == 0 || s_dig[cur][local - 1] != sd)) s_first[sd] = local;
  __syncthreads();
  if (valid) {
    block_perm[global] = s_perm[cur][local];
    meta[global] = sd * kItems + local - s_first[sd];
    const uint32_t nd = local + 1 < cnt ? s_dig[cur][local + 1] : (uint32_t)kBuckets;
    if (nd != sd) hist[sd * nblocks + block] = local + 1 - s_first[sd];
  }
}

__global__ void bucket_prefix(const uint32_t* __restrict__ hist, uint32_t nblocks,
                              uint32_t* __restrict__ prefixes, uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_chunk[256];
  const uint32_t bucket = blockIdx.x, lane = threadIdx.x, lanes = blockDim.x;
  const uint32_t chunk = (nblocks + lanes - 1) / lanes;
  const uint32_t b0 = min(nblocks, lane * chunk), b1 = min(nblocks, b0 + chunk);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; ++b) sum += hist[bucket * nblocks + b];
  s_chunk[lane] = sum;
  __syncthreads();
  if (lane == 0) {
    uint32_t run = 0;
    for (uint32_t l = 0; l < lanes; ++l) {
      const uint32_t c = s_chunk[l];
      s_chunk[l] = run;
      run += c;
    }
    totals[bucket] = run;
  }
  __syncthreads();
  uint32_t run = s_chunk[lane];
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t i = bucket * nblocks + b;
    prefixes[i] = run;
    run += hist[i];
  }
}

__global__ void bucket_offsets(const uint32_t* __restrict__ totals, uint32_t* __restrict__ offsets) {
  uint32_t run = 0;
  for (int b = 0; b < kBuckets; ++b) {
    offsets[b] = run;
    run += totals[b];
  }
}

__global__ void digit_scatter(const uint32_t* __restrict__ block_perm, const uint32_t* __restrict__ meta,
                              const uint32_t* __restrict__ prefixes, const uint32_t* __restrict__ offsets,
                              uint32_t n, uint32_t nblocks, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = meta[i] / kItems, rank = meta[i] % kItems, block = i / kItems;
  out[offsets[d] + prefixes[d * nblocks + block] + rank] = block_perm[i];
}

__global__ void gather_pair(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                            const uint32_t* __restrict__ perm, uint32_t n, uint32_t* __restrict__ ok,
                            uint32_t* __restrict__ ov) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
  ok[i] = keys[p];
  ov[i] = vals[p];
}

struct Scratch {
  uint32_t cap = 0;
  uint32_t *perm0 = nullptr, *perm1 = nullptr, *bperm = nullptr, *meta = nullptr;
  uint32_t *hist = nullptr, *pre = nullptr, *totals = nullptr, *offs = nullptr;
};
Scratch g_s;

int ensure(uint32_t n) {
  if (n <= g_s.cap) return 0;
  hipFree(g_s.perm0); hipFree(g_s.perm1); hipFree(g_s.bperm); hipFree(g_s.meta);
  hipFree(g_s.hist); hipFree(g_s.pre); hipFree(g_s.totals); hipFree(g_s.offs);
  const uint32_t nb = (n + kItems - 1) / kItems;
  const size_t b = (size_t)n * 4, h = (size_t)kBuckets * nb * 4;
  if (hipMalloc(&g_s.perm0, b) || hipMalloc(&g_s.perm1, b) || hipMalloc(&g_s.bperm, b) ||
      hipMalloc(&g_s.meta, b) || hipMalloc(&g_s.hist, h) || hipMalloc(&g_s.pre, h) ||
      hipMalloc(&g_s.totals, 64) || hipMalloc(&g_s.offs, 64))
    return 1;
  g_s.cap = n;
  return 0;
}

int run(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t* ok, uint32_t* ov,
        hipStream_t st) {
  if (n == 0) return 0;
  if (ensure(n)) return 1;
  const uint32_t nb = (n + kItems - 1) / kItems;
  const uint32_t lanes = nb < 256 ? nb : 256;
  uint32_t *cur = g_s.perm0, *nxt = g_s.perm1;
  for (uint32_t shift = 0, first = 1; shift < 32; shift += kBits, first = 0) {
    hipLaunchKernelGGL(block_digit_sort, dim3(nb), dim3(kItems), 0, st, keys, cur, n, shift, nb,
                       first, g_s.bperm, g_s.meta, g_s.hist);
    hipLaunchKernelGGL(bucket_prefix, dim3(kBuckets), dim3(lanes), 0, st, g_s.hist, nb, g_s.pre,
                       g_s.totals);
    hipLaunchKernelGGL(bucket_offsets, dim3(1), dim3(1), 0, st, g_s.totals, g_s.offs);
    hipLaunchKernelGGL(digit_scatter, dim3((n + kScatterBlock - 1) / kScatterBlock),
                       dim3(kScatterBlock), 0, st, g_s.bperm, g_s.meta, g_s.pre, g_s.offs, n, nb, nxt);
    uint32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
  hipLaunchKernelGGL(gather_pair, dim3((n + 255) / 256), dim3(256), 0, st, keys, vals, cur, n, ok, ov);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

extern "C" {

// Stable sort_by_key of n u32 keys + u32 values (device pointers) by the
// reference's schedule; returns the mean milliseconds of `reps` timed runs
// after one warm-up run, or a negative value on failure.
double refproxy_sort_by_key_u32_ms(const uint32_t* keys, const uint32_t* vals, uint32_t n,
                                   uint32_t* out_keys, uint32_t* out_vals, int reps) {
  hipStream_t st = nullptr;
  if (run(keys, vals, n, out_keys, out_vals, st)) return -1.0;
  hipEvent_t a, b;
  if (hipEventCreate(&a) || hipEventCreate(&b)) return -1.0;
  hipEventRecord(a, st);
  for (int r = 0; r < reps; ++r)
    if (run(keys, vals, n, out_keys, out_vals, st)) return -1.0;
  hipEventRecord(b, st);
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  return reps > 0 ? ms / reps : 0.0;
}

}  // extern "C"
