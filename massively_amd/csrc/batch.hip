// Batched independent jobs (no counterpart in the reference): many small
// inclusive/exclusive scan, reduce, copy_where and sort_by_key jobs run as one
// persistent launch per job family. Blocks take jobs from a queue, longest
// first, and run each job from start to end inside the block:
//  * scan: chunks of 256 x 8 elements, thread-serial scan of 8, DPP wave
//    scan, wave totals through LDS, a running carry between chunks;
//  * reduce: per chunk a fixed tree (pairwise over 16, DPP wave, 4 waves),
//    chunk partials in LDS reduced by a second fixed tree, then init (+);
//  * copy_where: per chunk each thread counts its 8 flags, a block scan
//    gives the output positions, a running offset between chunks; the count
//    goes to the job's device word;
//  * sort_by_key: the LDS tile sort (<= 4096 elements) or the one-block
//    range sort (longer jobs, element scratch from the library).
// Jobs above the block limits run with the device-wide kernels on the same
// stream after the batched launches. Every result is the same whichever other
// jobs share the call; floating-point scan and reduce results follow the
// fixed block order above (deterministic; a different association than the
// device-wide kernels).
#include <algorithm>

#include "msv_internal.hpp"
#include "predicate.hpp"
#include "tile_sort.hpp"

namespace msv {

constexpr int kBatchBlock = 256;
constexpr int kBatchRows = 8;                          // scan / copy_where elements per thread
constexpr uint32_t kBatchChunk = kBatchBlock * kBatchRows;
constexpr uint32_t kBatchReduceChunk = kBatchBlock * 16;
constexpr uint32_t kBatchStreamMax = 1u << 17;       // scan / reduce / copy_where per block
constexpr uint32_t kBatchSortIpt = 16;
constexpr uint32_t kBatchSortTile = kBatchBlock * kBatchSortIpt;
constexpr uint32_t kBatchSortMax = 1u << 15;          // sort_by_key per block
constexpr uint32_t kBatchMaxPartials = kBatchStreamMax / kBatchReduceChunk;

struct BatchItem {
  uint32_t kind, dtype, op, n;
  uint32_t pkind, invert, vbytes, pad;
  uint64_t init, pa, pb;
  const void* in;
  const void* in2;
  void* out;
  void* out2;
  void* s0;
  void* s1;
};

// ---- scan ------------------------------------------------------------------
template <typename T, int OP>
__device__ void block_scan_job(char* smem, const BatchItem& it) {
  using C = Combine<T, OP>;
  using U = typename Bits<T>::U;
  constexpr int NW = kBatchBlock / kWave;
  T* s_w = (T*)smem;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const T* in = (const T*)it.in;
  T* out = (T*)it.out;
  const uint32_t n = it.n;
  const bool excl = it.kind == MSV_JOB_EXCLUSIVE_SCAN;
  T carry = excl ? bit_cast<T>((U)it.init) : C::identity();
  for (uint32_t c0 = 0; c0 < n; c0 += kBatchChunk) {
    const uint32_t base = c0 + tid * kBatchRows;
    T v[kBatchRows];
#pragma unroll
    for (int r = 0; r < kBatchRows; ++r) v[r] = base + r < n ? in[base + r] : C::identity();
#pragma unroll
    for (int r = 1; r < kBatchRows; ++r) v[r] = C::apply(v[r - 1], v[r]);
    const T incl = wave_inclusive_scan<T, OP>(v[kBatchRows - 1]);
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    T wpre = C::identity(), btot = C::identity();
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if ((uint32_t)w < wid) wpre = C::apply(wpre, s_w[w]);
      btot = C::apply(btot, s_w[w]);
    }
    const T pre = C::apply(C::apply(carry, wpre), shift_up1(incl, C::identity()));
#pragma unroll
    for (int r = 0; r < kBatchRows; ++r) {
      if (base + r < n) {
        const T y = excl ? (r == 0 ? pre : C::apply(pre, v[r - 1])) : C::apply(pre, v[r]);
        out[base + r] = y;
      }
    }
    carry = C::apply(carry, btot);
    __syncthreads();
  }
}

// ---- reduce ----------------------------------------------------------------
template <typename T, int OP>
__device__ void block_reduce_job(char* smem, const BatchItem& it) {
  using C = Combine<T, OP>;
  using U = typename Bits<T>::U;
  constexpr int NW = kBatchBlock / kWave;
  T* s_w = (T*)smem;
  T* s_part = s_w + NW;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const T* in = (const T*)it.in;
  const uint32_t n = it.n;
  const uint32_t nchunks = (n + kBatchReduceChunk - 1) / kBatchReduceChunk;
  for (uint32_t ch = 0; ch < nchunks; ++ch) {
    const uint32_t c0 = ch * kBatchReduceChunk;
    T v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const uint32_t i = c0 + (uint32_t)r * kBatchBlock + tid;
      v[r] = i < n ? in[i] : C::identity();
    }
#pragma unroll
    for (int s = 1; s < 16; s <<= 1)
#pragma unroll
      for (int r = 0; r < 16; r += 2 * s) v[r] = C::apply(v[r], v[r + s]);
    const T w = wave_reduce<T, OP>(v[0]);
    if (lane == 0) s_w[wid] = w;
    __syncthreads();
    if (tid == 0) s_part[ch] = C::apply(C::apply(s_w[0], s_w[1]), C::apply(s_w[2], s_w[3]));
    __syncthreads();
  }
  if (tid == 0) {
    // Pairwise tree over the chunk partials (fixed shape for a given n).
    uint32_t m = nchunks;
    while (m > 1) {
      const uint32_t h = (m + 1) / 2;
      for (uint32_t k = 0; k + h < m; ++k) s_part[k] = C::apply(s_part[k], s_part[k + h]);
      m = h;
    }
    const T init = bit_cast<T>((U)it.init);
    *(T*)it.out = nchunks ? C::apply(init, s_part[0]) : init;
  }
  __syncthreads();
}

// ---- copy_where (predicate on the input) ----------------------------------
template <typename T>
__device__ void block_select_job(char* smem, const BatchItem& it) {
  using U = typename Bits<T>::U;
  constexpr int NW = kBatchBlock / kWave;
  uint32_t* s_w = (uint32_t*)smem;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  Pred<T> pred{};
  pred.kind = it.pkind;
  pred.ua = (U)it.pa;
  pred.ub = (U)it.pb;
  pred.a = bit_cast<T>((U)it.pa);
  const T* in = (const T*)it.in;
  T* out = (T*)it.out;
  const uint32_t n = it.n;
  const bool inv = it.invert != 0;
  uint32_t written = 0;
  for (uint32_t c0 = 0; c0 < n; c0 += kBatchChunk) {
    const uint32_t base = c0 + tid * kBatchRows;
    T v[kBatchRows];
    uint32_t keep = 0, cnt = 0;
#pragma unroll
    for (int r = 0; r < kBatchRows; ++r) {
      const bool ok = base + r < n;
      v[r] = ok ? in[base + r] : (T)0;
      const bool k = ok && (pred(v[r]) != inv);
      keep |= (k ? 1u : 0u) << r;
      cnt += k ? 1u : 0u;
    }
    const uint32_t incl = wave_inclusive_scan<uint32_t, OP_ADD>(cnt);
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    uint32_t wpre = 0, btot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if ((uint32_t)w < wid) wpre += s_w[w];
      btot += s_w[w];
    }
    uint32_t pos = written + wpre + incl - cnt;
#pragma unroll
    for (int r = 0; r < kBatchRows; ++r) {
      if (keep & (1u << r)) out[pos++] = v[r];
    }
    written += btot;
    __syncthreads();
  }
  if (tid == 0) *(uint32_t*)it.out2 = written;
}

// ---- family kernels --------------------------------------------------------
#define MSV_BATCH_DTYPE(FN, ...)                              \
  switch (it.dtype) {                                         \
    case MSV_U32: FN<uint32_t>(__VA_ARGS__); break;           \
    case MSV_I32: FN<int32_t>(__VA_ARGS__); break;            \
    case MSV_F32: FN<float>(__VA_ARGS__); break;              \
    case MSV_U64: FN<uint64_t>(__VA_ARGS__); break;           \
    case MSV_I64: FN<int64_t>(__VA_ARGS__); break;            \
    default: FN<double>(__VA_ARGS__); break;                  \
  }

template <typename T>
__device__ void scan_or_reduce_job(char* smem, const BatchItem& it) {
  const bool red = it.kind == MSV_JOB_REDUCE;
  switch (it.op) {
    case MSV_OP_ADD:
      if (red) block_reduce_job<T, OP_ADD>(smem, it);
      else block_scan_job<T, OP_ADD>(smem, it);
      break;
    case MSV_OP_MIN:
      if (red) block_reduce_job<T, OP_MIN>(smem, it);
      else block_scan_job<T, OP_MIN>(smem, it);
      break;
    case MSV_OP_MAX:
      if (red) block_reduce_job<T, OP_MAX>(smem, it);
      else block_scan_job<T, OP_MAX>(smem, it);
      break;
    default:
      if constexpr (!IsFloat<T>::value) {
        if (red) block_reduce_job<T, OP_XOR>(smem, it);
        else block_scan_job<T, OP_XOR>(smem, it);
      }
      break;
  }
}

// Next job of this block from the family queue (uniform across the block).
__device__ inline uint32_t next_job(uint32_t* queue, uint32_t* s_item) {
  if (threadIdx.x == 0) *s_item = atomicAdd(queue, 1u);
  __syncthreads();
  const uint32_t j = *s_item;
  __syncthreads();
  return j;
}

__global__ __launch_bounds__(kBatchBlock) void batch_stream_kernel(const BatchItem* __restrict__ items,
                                                                   uint32_t count,
                                                                   uint32_t* queue) {
  __shared__ __align__(16) char smem[sizeof(uint64_t) * (kBatchBlock / kWave + kBatchMaxPartials)];
  __shared__ uint32_t s_item;
  for (;;) {
    const uint32_t j = next_job(queue, &s_item);
    if (j >= count) break;
    const BatchItem it = items[j];
    if (it.kind == MSV_JOB_COPY_WHERE) {
      MSV_BATCH_DTYPE(block_select_job, smem, it)
    } else {
      MSV_BATCH_DTYPE(scan_or_reduce_job, smem, it)
    }
  }
}

template <int KW, int EW>
__global__ __launch_bounds__(kBatchBlock) void batch_sort_kernel(const BatchItem* __restrict__ items,
                                                                 uint32_t count, uint32_t* queue) {
  using T = TileSmem<EW, kBatchBlock, kBatchSortIpt, 8>;
  using R = RangeSmem<EW, kBatchBlock, kBatchSortIpt>;
  __shared__ __align__(16) char smem[sizeof(T) > sizeof(R) ? sizeof(T) : sizeof(R)];
  __shared__ uint32_t s_item;
  for (;;) {
    const uint32_t j = next_job(queue, &s_item);
    if (j >= count) break;
    const BatchItem it = items[j];
    if (it.n == 0) continue;
    OrderMap om;
    om.kind = (it.dtype == MSV_F32 || it.dtype == MSV_F64) ? KEY_FLOAT
              : (it.dtype == MSV_I32 || it.dtype == MSV_I64) ? KEY_SIGNED
                                                             : KEY_UNSIGNED;
    om.mode = (int)it.op;
    if (it.n <= kBatchSortTile) {
      tile_sort_range<KW, EW, kBatchBlock, kBatchSortIpt, 8, false>(
          *(T*)smem, it.in, it.in2, it.out2, it.out, 0u, it.n, nullptr, 0, 0, om);
    } else {
      range_sort<KW, EW, kBatchBlock, kBatchSortIpt>(*(R*)smem, RangeIo{it.in, it.in2, 0u},
                                                     RangeIo{it.out2, it.out, 0u},
                                                     (Elem<EW>*)it.s0, (Elem<EW>*)it.s1, it.n, om);
    }
  }
}

// ---- host ------------------------------------------------------------------

static bool job_is_small(const msv_job& j) {
  if (j.kind == MSV_JOB_SORT_BY_KEY) return j.n <= kBatchSortMax;
  return j.n <= kBatchStreamMax;
}

static msv_status check_job(const msv_job& j, uint32_t i) {
  auto bad = [&](const char* why) {
    set_error(MSV_ERR_INVALID_ARGUMENT, i, 0, "job %u: %s", i, why);
    return MSV_ERR_INVALID_ARGUMENT;
  };
  if (j.kind < MSV_JOB_INCLUSIVE_SCAN || j.kind > MSV_JOB_SORT_BY_KEY) return bad("unknown kind");
  if (!dtype_valid((int)j.dtype)) return bad("bad dtype");
  const bool flt = dtype_is_float((msv_dtype)j.dtype);
  switch (j.kind) {
    case MSV_JOB_INCLUSIVE_SCAN:
    case MSV_JOB_EXCLUSIVE_SCAN:
    case MSV_JOB_REDUCE:
      if (j.op > MSV_OP_XOR || (j.op == MSV_OP_XOR && flt)) return bad("operator not defined");
      if (!j.out || (j.n && !j.in)) return bad("null pointer");
      break;
    case MSV_JOB_COPY_WHERE:
      if (j.pred.kind > MSV_PRED_BITS_EQ) return bad("unknown predicate");
      if (!j.out2 || (j.n && (!j.in || !j.out))) return bad("null pointer");
      break;
    case MSV_JOB_SORT_BY_KEY:
      if (j.op > MSV_ORDER_GREATER) return bad("unknown order");
      if (j.value_bytes != 0 && j.value_bytes != 4 && j.value_bytes != 8) return bad("value_bytes");
      if (j.n && (!j.in || !j.out2 || (j.value_bytes && (!j.in2 || !j.out)))) return bad("null pointer");
      break;
  }
  return MSV_OK;
}

static BatchItem to_item(const msv_job& j) {
  BatchItem it{};
  it.kind = j.kind;
  it.dtype = j.dtype;
  it.op = j.op;
  it.n = j.n;
  it.pkind = j.pred.kind;
  it.invert = j.invert ? 1u : 0u;
  it.vbytes = j.value_bytes;
  it.init = j.init;
  it.pa = j.pred.a;
  it.pb = j.pred.b;
  it.in = j.in;
  it.in2 = j.in2;
  it.out = j.out;
  it.out2 = j.out2;
  return it;
}

// One persistent launch over `items` (already on the device at d_items).
template <typename Kern>
static msv_status launch_family(Ctx& c, Kern kern, const BatchItem* d_items, uint32_t count,
                                uint32_t* queue, int* per_cu_cache) {
  if (count == 0) return MSV_OK;
  if (*per_cu_cache == 0) {
    int b = 0;
    MSV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, kBatchBlock, 0));
    *per_cu_cache = std::max(1, b);
  }
  const uint32_t grid = std::min<uint32_t>(count, (uint32_t)(*per_cu_cache * c.num_cus()));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBatchBlock), 0, c.stream(), d_items, count, queue);
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

// Element-width class of a sort job: 0..5 = (KW,EW) in (1,1) (1,2) (1,3) (2,2) (2,3) (2,4).
static int sort_class(const msv_job& j) {
  const int kw = (int)dtype_size((msv_dtype)j.dtype) / 4;
  const int vw = (int)j.value_bytes / 4;
  return kw == 1 ? vw : 3 + vw;
}

static msv_status large_job(Ctx& c, const msv_job& j) {
  const msv_dtype dt = (msv_dtype)j.dtype;
  switch (j.kind) {
    case MSV_JOB_INCLUSIVE_SCAN:
      return scan_run(c, dt, (msv_op)j.op, false, j.in, j.out, j.n, nullptr);
    case MSV_JOB_EXCLUSIVE_SCAN:
      return scan_run(c, dt, (msv_op)j.op, true, j.in, j.out, j.n, &j.init);
    case MSV_JOB_REDUCE:
      return reduce_run_device(c, dt, (msv_op)j.op, j.in, j.n, &j.init, j.out);
    case MSV_JOB_COPY_WHERE:
      return select_pred_run(c, dt, j.in, &j.pred, j.invert != 0, j.n, j.out,
                             (uint32_t*)j.out2);
    default:
      return radix_sort_run(c, dt, (int)j.op, j.value_bytes, j.in, j.value_bytes ? j.in2 : nullptr,
                            j.n, j.out2, j.value_bytes ? j.out : nullptr);
  }
}

msv_status run_jobs(Ctx& c, const msv_job* jobs, uint32_t count) {
  for (uint32_t i = 0; i < count; ++i) MSV_TRY(check_job(jobs[i], i));
  // Families: 0 = scan/reduce/copy_where, 1 + class = sort by element width.
  std::vector<uint32_t> fam[7];
  std::vector<uint32_t> large;
  for (uint32_t i = 0; i < count; ++i) {
    const msv_job& j = jobs[i];
    if (!job_is_small(j)) {
      large.push_back(i);
      continue;
    }
    fam[j.kind == MSV_JOB_SORT_BY_KEY ? 1 + sort_class(j) : 0].push_back(i);
  }
  std::vector<BatchItem> items;
  uint32_t start[8] = {0};
  size_t sort_scratch = 0;
  for (int f = 0; f < 7; ++f) {
    // Longest first: the queue hands them out before the short ones.
    std::stable_sort(fam[f].begin(), fam[f].end(),
                     [&](uint32_t a, uint32_t b) { return jobs[a].n > jobs[b].n; });
    start[f] = (uint32_t)items.size();
    for (uint32_t i : fam[f]) {
      items.push_back(to_item(jobs[i]));
      const msv_job& j = jobs[i];
      if (j.kind == MSV_JOB_SORT_BY_KEY && j.n > kBatchSortTile) {
        const size_t ew = dtype_size((msv_dtype)j.dtype) + j.value_bytes;
        sort_scratch += 2 * (((size_t)j.n * ew + 255) & ~(size_t)255);
      }
    }
  }
  start[7] = (uint32_t)items.size();
  if (!items.empty()) {
    BatchItem* d_items = (BatchItem*)c.scratch(sizeof(BatchItem) * items.size());
    uint32_t* queues = (uint32_t*)c.scratch(sizeof(uint32_t) * 8);
    char* scratch = sort_scratch ? (char*)c.scratch(sort_scratch) : nullptr;
    if (!d_items || !queues || (sort_scratch && !scratch)) return MSV_ERR_OUT_OF_MEMORY;
    size_t off = 0;
    for (BatchItem& it : items) {
      if (it.kind != MSV_JOB_SORT_BY_KEY || it.n <= kBatchSortTile) continue;
      const size_t bytes =
          ((size_t)it.n * (dtype_size((msv_dtype)it.dtype) + it.vbytes) + 255) & ~(size_t)255;
      it.s0 = scratch + off;
      it.s1 = scratch + off + bytes;
      off += 2 * bytes;
    }
    const void* src = c.staging(items.data(), sizeof(BatchItem) * items.size());
    if (!src) return MSV_ERR_OUT_OF_MEMORY;
    MSV_HIP(hipMemcpyAsync(d_items, src, sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice,
                           c.stream()));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, c.stream(), queues, 8u, 0u);
    int* per_cu = c.ex()->batch_per_cu;
    MSV_TRY(launch_family(c, batch_stream_kernel, d_items + start[0], start[1] - start[0],
                          queues + 0, &per_cu[0]));
    MSV_TRY(launch_family(c, batch_sort_kernel<1, 1>, d_items + start[1], start[2] - start[1], queues + 1, &per_cu[1]));
    MSV_TRY(launch_family(c, batch_sort_kernel<1, 2>, d_items + start[2], start[3] - start[2], queues + 2, &per_cu[2]));
    MSV_TRY(launch_family(c, batch_sort_kernel<1, 3>, d_items + start[3], start[4] - start[3], queues + 3, &per_cu[3]));
    MSV_TRY(launch_family(c, batch_sort_kernel<2, 2>, d_items + start[4], start[5] - start[4], queues + 4, &per_cu[4]));
    MSV_TRY(launch_family(c, batch_sort_kernel<2, 3>, d_items + start[5], start[6] - start[5], queues + 5, &per_cu[5]));
    MSV_TRY(launch_family(c, batch_sort_kernel<2, 4>, d_items + start[6], start[7] - start[6], queues + 6, &per_cu[6]));
  }
  for (uint32_t i : large) MSV_TRY(large_job(c, jobs[i]));
  return MSV_OK;
}

msv_status run_jobs_plain(Ctx& c, const msv_job* jobs, uint32_t count) {
  for (uint32_t i = 0; i < count; ++i) MSV_TRY(check_job(jobs[i], i));
  for (uint32_t i = 0; i < count; ++i) MSV_TRY(large_job(c, jobs[i]));
  return MSV_OK;
}

}  // namespace msv

using namespace msv;

extern "C" {

msv_status msv_run_jobs(msv_executor* ex, const msv_job* jobs, uint32_t count) {
  clear_error();
  if (!ex || (count && !jobs)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor or job array");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  return run_jobs(c, jobs, count);
}

}  // extern "C"
