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
#include <vector>

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

// A batch table on the device: this header, then the jobs grouped by family
// (family f at [start[f], start[f + 1]); 0 = scan / reduce / copy_where,
// 1..6 = sorts by element width), then the generation entries.
constexpr int kBatchFamilies = 7;
struct BatchHeader {
  uint32_t start[8];
  uint32_t queue[8];  // per-family job counters, uploaded as zero
  uint32_t ngen;      // generation entries
  uint32_t gen_chunks;
  uint32_t pad[6];
};
static_assert(sizeof(BatchHeader) % 16 == 0, "table alignment");
constexpr uint32_t kBatchGenChunk = 256 * 8;  // elements per generation chunk

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

// Family f's jobs are items[hdr->start[f], hdr->start[f + 1]); its queue is
// hdr->queue[f] (uploaded as zero with the table), so a launch's grid does
// not depend on the batch and a captured launch can be replayed for any.
__global__ __launch_bounds__(kBatchBlock) void batch_stream_kernel(const BatchItem* __restrict__ items,
                                                                   BatchHeader* hdr) {
  __shared__ __align__(16) char smem[sizeof(uint64_t) * (kBatchBlock / kWave + kBatchMaxPartials)];
  __shared__ uint32_t s_item;
  const uint32_t first = hdr->start[0], count = hdr->start[1] - first;
  for (;;) {
    const uint32_t j = next_job(&hdr->queue[0], &s_item);
    if (j >= count) break;
    const BatchItem it = items[first + j];
    if (it.kind == MSV_JOB_COPY_WHERE) {
      MSV_BATCH_DTYPE(block_select_job, smem, it)
    } else {
      MSV_BATCH_DTYPE(scan_or_reduce_job, smem, it)
    }
  }
}

// Family of the sort jobs with key width KW and element width EW (words).
template <int KW, int EW>
constexpr int sort_family() {
  return KW == 1 ? EW : 2 + EW;  // (1,1) (1,2) (1,3) (2,2) (2,3) (2,4) -> 1..6
}

template <int KW, int EW>
__global__ __launch_bounds__(kBatchBlock) void batch_sort_kernel(const BatchItem* __restrict__ items,
                                                                 BatchHeader* hdr) {
  using T = TileSmem<EW, kBatchBlock, kBatchSortIpt, 8>;
  using R = RangeSmem<EW, kBatchBlock, kBatchSortIpt>;
  constexpr int F = sort_family<KW, EW>();
  __shared__ __align__(16) char smem[sizeof(T) > sizeof(R) ? sizeof(T) : sizeof(R)];
  __shared__ uint32_t s_item;
  const uint32_t first = hdr->start[F], count = hdr->start[F + 1] - first;
  for (;;) {
    const uint32_t j = next_job(&hdr->queue[F], &s_item);
    if (j >= count) break;
    const BatchItem it = items[first + j];
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

// Input generation of a batch table: chunk c of all entries' elements, block
// by block (entry found by binary search over the entries' first chunks).
__global__ __launch_bounds__(256) void batch_gen_kernel(const BatchGen* __restrict__ gens,
                                                        const BatchHeader* __restrict__ hdr) {
  const uint32_t ngen = hdr->ngen, chunks = hdr->gen_chunks;
  for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
    uint32_t lo = 0, hi = ngen;  // last entry with chunk0 <= c
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (gens[mid].chunk0 <= c) lo = mid;
      else hi = mid;
    }
    const BatchGen g = gens[lo];
    const uint32_t base = (c - g.chunk0) * kBatchGenChunk;
    for (uint32_t i = base + threadIdx.x; i < min(base + kBatchGenChunk, g.n); i += 256) {
      if (g.in) {
        const uint32_t v = pcg_hash32(i ^ g.key);
        switch (g.dtype) {
          case MSV_U32:
          case MSV_I32: ((uint32_t*)g.in)[i] = v; break;
          case MSV_F32: ((uint32_t*)g.in)[i] = 0x3F000000u | (v % 0x00800000u); break;
          case MSV_U64:
          case MSV_I64:
          case MSV_F64: {
            const uint64_t w = (uint64_t)v | ((uint64_t)pcg_hash32(i ^ g.key ^ 0x9e3779b9u) << 32);
            ((uint64_t*)g.in)[i] =
                g.dtype == MSV_F64 ? 0x3FE0000000000000ull | (w % 0x0010000000000000ull) : w;
            break;
          }
          default: break;
        }
      }
      if (g.iota) g.iota[i] = i;
    }
  }
}

// Count gather of a batch table: counts[count_index] = *count_src.
__global__ __launch_bounds__(256) void batch_counts_kernel(const BatchGen* __restrict__ gens,
                                                           const BatchHeader* __restrict__ hdr,
                                                           uint32_t* __restrict__ counts) {
  for (uint32_t g = threadIdx.x; g < hdr->ngen; g += 256)
    if (gens[g].count_src) counts[gens[g].count_index] = *gens[g].count_src;
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

bool batch_job_small(const msv_job& j) { return job_is_small(j); }

size_t batch_table_bytes(uint32_t capacity) {
  return sizeof(BatchHeader) + (size_t)capacity * (sizeof(BatchItem) + sizeof(BatchGen));
}

msv_status batch_table_fill(Ctx& c, const msv_job* jobs, uint32_t count, const BatchGen* gens,
                            uint32_t ngen, uint32_t capacity, char* host, uint32_t* mask) {
  if (count > capacity || ngen > capacity) {
    set_error(MSV_ERR_LENGTH_TOO_LARGE, count, capacity, "batch larger than its table");
    return MSV_ERR_LENGTH_TOO_LARGE;
  }
  std::vector<uint32_t> fam[kBatchFamilies];
  for (uint32_t i = 0; i < count; ++i) {
    MSV_TRY(check_job(jobs[i], i));
    if (!job_is_small(jobs[i])) {
      set_error(MSV_ERR_INVALID_ARGUMENT, i, 0, "job %u is too large for a batch table", i);
      return MSV_ERR_INVALID_ARGUMENT;
    }
    fam[jobs[i].kind == MSV_JOB_SORT_BY_KEY ? 1 + sort_class(jobs[i]) : 0].push_back(i);
  }
  BatchHeader* hdr = (BatchHeader*)host;
  BatchItem* items = (BatchItem*)(host + sizeof(BatchHeader));
  BatchGen* tg = (BatchGen*)(host + sizeof(BatchHeader) + (size_t)capacity * sizeof(BatchItem));
  *hdr = BatchHeader{};
  *mask = 0;
  uint32_t k = 0;
  for (int f = 0; f < kBatchFamilies; ++f) {
    // Longest first: the queue hands them out before the short ones.
    std::stable_sort(fam[f].begin(), fam[f].end(),
                     [&](uint32_t a, uint32_t b) { return jobs[a].n > jobs[b].n; });
    hdr->start[f] = k;
    if (!fam[f].empty()) *mask |= 1u << f;
    for (uint32_t i : fam[f]) {
      BatchItem it = to_item(jobs[i]);
      if (it.kind == MSV_JOB_SORT_BY_KEY && it.n > kBatchSortTile) {
        const size_t bytes =
            ((size_t)it.n * (dtype_size((msv_dtype)it.dtype) + it.vbytes) + 255) & ~(size_t)255;
        char* sc = (char*)c.scratch(2 * bytes);
        if (!sc) return MSV_ERR_OUT_OF_MEMORY;
        it.s0 = sc;
        it.s1 = sc + bytes;
      }
      items[k++] = it;
    }
  }
  hdr->start[kBatchFamilies] = k;
  uint32_t chunks = 0;
  for (uint32_t g = 0; g < ngen; ++g) {
    tg[g] = gens[g];
    tg[g].chunk0 = chunks;
    chunks += (gens[g].n + kBatchGenChunk - 1) / kBatchGenChunk;
  }
  hdr->ngen = ngen;
  hdr->gen_chunks = chunks;
  return MSV_OK;
}

// Resident blocks per CU of the batched kernels (cached per executor).
static int family_per_cu(msv_executor* ex, int f) {
  std::atomic<int>& slot = ex->batch_per_cu[f];
  int pc = slot.load(std::memory_order_relaxed);
  if (pc == 0) {
    int b = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (f) {
      case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_stream_kernel, kBatchBlock, 0); break;
      case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_sort_kernel<1, 1>, kBatchBlock, 0); break;
      case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_sort_kernel<1, 2>, kBatchBlock, 0); break;
      case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_sort_kernel<1, 3>, kBatchBlock, 0); break;
      case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_sort_kernel<2, 2>, kBatchBlock, 0); break;
      case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_sort_kernel<2, 3>, kBatchBlock, 0); break;
      default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_sort_kernel<2, 4>, kBatchBlock, 0); break;
    }
    if (e != hipSuccess) (void)hipGetLastError();
    pc = std::max(1, b);
    slot.store(pc, std::memory_order_relaxed);
  }
  return pc;
}

void batch_occupancy_warm(msv_executor* ex) {
  for (int f = 0; f < kBatchFamilies; ++f) (void)family_per_cu(ex, f);
}

// ---- autotuned grids -------------------------------------------------------
// A family's grid is its resident blocks per CU shifted right by the chosen
// candidate c (0 = every block the CUs hold at once, 1 = half, 2 = a quarter,
// 3 = an eighth; at least one per CU), times the CUs. Fewer blocks mean fewer
// idle blocks to launch and retire and less contention on the job queue and
// the caches; more blocks hide the latency of short jobs. The choice is made
// per (family, job-count class, mean-length class) of a batch, by timing, and
// kept in the tuning table under "batch".
constexpr int kBatchGridCands = 4;

static uint32_t log2_floor(uint32_t x) { return x ? 31u - (uint32_t)__builtin_clz(x) : 0u; }
static uint32_t log2_ceil(uint32_t x) { return x <= 1 ? 0u : log2_floor(x - 1) + 1; }

uint32_t batch_tune_key(int f, uint32_t jobs, uint32_t mean_len) {
  return (uint32_t)f | (std::min(31u, log2_ceil(mean_len)) << 3) |
         (std::min(15u, log2_floor(jobs)) << 8);
}

// Jobs and mean length of family f in a host table.
static void family_shape(const char* host, int f, uint32_t* jobs, uint32_t* mean_len) {
  const BatchHeader* hdr = (const BatchHeader*)host;
  const BatchItem* items = (const BatchItem*)(host + sizeof(BatchHeader));
  const uint32_t a = hdr->start[f], b = hdr->start[f + 1];
  uint64_t total = 0;
  for (uint32_t i = a; i < b; ++i) total += items[i].n;
  *jobs = b - a;
  *mean_len = b > a ? (uint32_t)std::min<uint64_t>(total / (b - a), 0xffffffffu) : 0u;
}

static int family_choice(msv_executor* ex, const char* host, int f) {
  uint32_t jobs, mean;
  family_shape(host, f, &jobs, &mean);
  const int c = tune_get(ex, "batch", batch_tune_key(f, jobs, mean));
  return c >= 0 && c < kBatchGridCands ? c : 0;
}

static uint32_t family_grid(msv_executor* ex, int f, int choice) {
  return (uint32_t)std::max(1, family_per_cu(ex, f) >> choice) * (uint32_t)ex->num_cus;
}

uint32_t batch_shape_code(msv_executor* ex, const char* host, uint32_t mask) {
  uint32_t code = mask;
  for (int f = 0; f < kBatchFamilies; ++f)
    if (mask & (1u << f)) code |= (uint32_t)family_choice(ex, host, f) << (8 + 2 * f);
  return code;
}

// Byte ranges a job reads and writes (empty ranges for n == 0).
struct ByteRange {
  uintptr_t a, b;
};
static void job_ranges(const BatchItem& it, std::vector<ByteRange>* rd, std::vector<ByteRange>* wr) {
  if (it.n == 0) return;
  const uintptr_t es = dtype_size((msv_dtype)it.dtype), n = it.n;
  auto add = [](std::vector<ByteRange>* v, const void* p, uintptr_t bytes) {
    if (p && bytes) v->push_back({(uintptr_t)p, (uintptr_t)p + bytes});
  };
  add(rd, it.in, n * es);
  switch (it.kind) {
    case MSV_JOB_REDUCE: add(wr, it.out, es); break;
    case MSV_JOB_COPY_WHERE:
      add(wr, it.out, n * es);
      add(wr, it.out2, 4);
      break;
    case MSV_JOB_SORT_BY_KEY:
      add(rd, it.in2, n * it.vbytes);
      add(wr, it.out2, n * es);
      add(wr, it.out, n * it.vbytes);
      break;
    default: add(wr, it.out, n * es); break;
  }
}

// Whether no job of family f writes bytes that a job of the family reads, so
// that running the family again gives the same outputs.
static bool family_rerunnable(const char* host, int f) {
  const BatchHeader* hdr = (const BatchHeader*)host;
  const BatchItem* items = (const BatchItem*)(host + sizeof(BatchHeader));
  std::vector<ByteRange> rd, wr;
  for (uint32_t i = hdr->start[f]; i < hdr->start[f + 1]; ++i) job_ranges(items[i], &rd, &wr);
  std::sort(rd.begin(), rd.end(), [](const ByteRange& x, const ByteRange& y) { return x.a < y.a; });
  std::vector<uintptr_t> end_max(rd.size());
  for (size_t i = 0; i < rd.size(); ++i) end_max[i] = std::max(rd[i].b, i ? end_max[i - 1] : 0);
  for (const ByteRange& w : wr) {
    // Reads starting before w.b: the one reaching furthest must end at or before w.a.
    const size_t k = std::lower_bound(rd.begin(), rd.end(), w.b,
                                      [](const ByteRange& r, uintptr_t v) { return r.a < v; }) -
                     rd.begin();
    if (k && end_max[k - 1] > w.a) return false;
  }
  return true;
}

static msv_status launch_family(hipStream_t stream, char* dev, int f, uint32_t grid);

msv_status batch_autotune(msv_executor* ex, hipStream_t stream, const char* host, char* dev,
                          uint32_t capacity, uint32_t mask) {
  if (!ex->opt_autotune || ex->capturing) return MSV_OK;
  const BatchHeader* hdr = (const BatchHeader*)host;
  bool uploaded = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  msv_status st = MSV_OK;
  for (int f = 0; f < kBatchFamilies && st == MSV_OK; ++f) {
    if (!(mask & (1u << f))) continue;
    uint32_t jobs, mean;
    family_shape(host, f, &jobs, &mean);
    const uint32_t key = batch_tune_key(f, jobs, mean);
    if (jobs < 16 || tune_get(ex, "batch", key) >= 0 || !family_rerunnable(host, f)) continue;
    if (!uploaded) {
      uploaded = true;
      if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
          hipMemcpyAsync(dev, host, batch_table_bytes(capacity), hipMemcpyHostToDevice, stream) !=
              hipSuccess) {
        st = MSV_ERR_LAUNCH;
        break;
      }
      const BatchGen* gens =
          (const BatchGen*)(dev + sizeof(BatchHeader) + (size_t)capacity * sizeof(BatchItem));
      hipLaunchKernelGGL(batch_gen_kernel, dim3((uint32_t)ex->num_cus * 4u), dim3(256), 0, stream,
                         gens, (const BatchHeader*)dev);
    }
    int best = 0;
    float best_ms = 1e30f;
    uint32_t last_grid = 0;
    for (int cand = 0; cand < kBatchGridCands && st == MSV_OK; ++cand) {
      const uint32_t grid = family_grid(ex, f, cand);
      if (grid == last_grid) continue;
      last_grid = grid;
      float ms_min = 1e30f;
      // One untimed run, then the faster of two timed runs; each starts from
      // a queue counter of zero (the header is uploaded again).
      for (int rep = 0; rep < 3 && st == MSV_OK; ++rep) {
        if (hipMemcpyAsync(dev, host, sizeof(BatchHeader), hipMemcpyHostToDevice, stream) !=
                hipSuccess ||
            (rep && hipEventRecord(e0, stream) != hipSuccess)) {
          st = MSV_ERR_LAUNCH;
          break;
        }
        st = launch_family(stream, dev, f, grid);
        if (st != MSV_OK || !rep) continue;
        float ms = 0;
        if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          st = MSV_ERR_LAUNCH;
          break;
        }
        ms_min = std::min(ms_min, ms);
      }
      if (ms_min < best_ms) {
        best_ms = ms_min;
        best = cand;
      }
    }
    if (st == MSV_OK) tune_put(ex, "batch", key, best);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st != MSV_OK) {
    (void)hipGetLastError();
    set_error(st, 0, 0, "batch autotuning failed");
  }
  (void)hdr;
  return st;
}

// One persistent launch of family f of the table at `dev`.
static msv_status launch_family(hipStream_t stream, char* dev, int f, uint32_t grid) {
  const BatchItem* items = (const BatchItem*)(dev + sizeof(BatchHeader));
  BatchHeader* hdr = (BatchHeader*)dev;
  const dim3 g(grid), block(kBatchBlock);
  switch (f) {
    case 0: hipLaunchKernelGGL(batch_stream_kernel, g, block, 0, stream, items, hdr); break;
    case 1: hipLaunchKernelGGL((batch_sort_kernel<1, 1>), g, block, 0, stream, items, hdr); break;
    case 2: hipLaunchKernelGGL((batch_sort_kernel<1, 2>), g, block, 0, stream, items, hdr); break;
    case 3: hipLaunchKernelGGL((batch_sort_kernel<1, 3>), g, block, 0, stream, items, hdr); break;
    case 4: hipLaunchKernelGGL((batch_sort_kernel<2, 2>), g, block, 0, stream, items, hdr); break;
    case 5: hipLaunchKernelGGL((batch_sort_kernel<2, 3>), g, block, 0, stream, items, hdr); break;
    default: hipLaunchKernelGGL((batch_sort_kernel<2, 4>), g, block, 0, stream, items, hdr); break;
  }
  MSV_HIP(hipGetLastError());
  return MSV_OK;
}

// The family launches of a table at `dev`; grid_of(f) gives each grid.
template <typename GridOf>
static msv_status launch_families(hipStream_t stream, char* dev, uint32_t mask, GridOf grid_of) {
  for (int f = 0; f < kBatchFamilies; ++f)
    if (mask & (1u << f)) MSV_TRY(launch_family(stream, dev, f, grid_of(f)));
  return MSV_OK;
}

msv_status batch_table_enqueue(msv_executor* ex, hipStream_t stream, const char* host, char* dev,
                               uint32_t capacity, uint32_t mask, uint32_t* d_counts,
                               uint32_t* h_counts) {
  uint32_t grid[kBatchFamilies];
  for (int f = 0; f < kBatchFamilies; ++f)
    grid[f] = (mask & (1u << f)) ? family_grid(ex, f, family_choice(ex, host, f)) : 0u;
  MSV_HIP(hipMemcpyAsync(dev, host, batch_table_bytes(capacity), hipMemcpyHostToDevice, stream));
  const BatchHeader* hdr = (const BatchHeader*)dev;
  const BatchGen* gens = (const BatchGen*)(dev + sizeof(BatchHeader) + (size_t)capacity * sizeof(BatchItem));
  hipLaunchKernelGGL(batch_gen_kernel, dim3((uint32_t)ex->num_cus * 4u), dim3(256), 0, stream, gens,
                     hdr);
  MSV_HIP(hipGetLastError());
  MSV_TRY(launch_families(stream, dev, mask, [&](int f) { return grid[f]; }));
  hipLaunchKernelGGL(batch_counts_kernel, dim3(1), dim3(256), 0, stream, gens, hdr, d_counts);
  MSV_HIP(hipGetLastError());
  MSV_HIP(hipMemcpyAsync(h_counts, d_counts, sizeof(uint32_t) * capacity, hipMemcpyDeviceToHost,
                         stream));
  return MSV_OK;
}

msv_status run_jobs(Ctx& c, const msv_job* jobs, uint32_t count) {
  for (uint32_t i = 0; i < count; ++i) MSV_TRY(check_job(jobs[i], i));
  std::vector<msv_job> small;
  std::vector<uint32_t> large;
  for (uint32_t i = 0; i < count; ++i) {
    if (job_is_small(jobs[i])) small.push_back(jobs[i]);
    else large.push_back(i);
  }
  if (!small.empty()) {
    const uint32_t cap = (uint32_t)small.size();
    std::vector<char> host(batch_table_bytes(cap));
    uint32_t mask = 0;
    MSV_TRY(batch_table_fill(c, small.data(), cap, nullptr, 0, cap, host.data(), &mask));
    char* dev = (char*)c.scratch(host.size());
    if (!dev) return MSV_ERR_OUT_OF_MEMORY;
    msv_executor* ex = c.ex();
    // Families whose shape class has no tuned grid yet are timed first (a
    // one-time host synchronisation per class; skipped while capturing).
    MSV_TRY(batch_autotune(ex, c.stream(), host.data(), dev, cap, mask));
    const void* src = c.staging(host.data(), host.size());
    if (!src) return MSV_ERR_OUT_OF_MEMORY;
    MSV_HIP(hipMemcpyAsync(dev, src, host.size(), hipMemcpyHostToDevice, c.stream()));
    const BatchHeader* hh = (const BatchHeader*)host.data();
    MSV_TRY(launch_families(c.stream(), dev, mask, [&](int f) {
      const uint32_t n = hh->start[f + 1] - hh->start[f];
      return std::min<uint32_t>(n, family_grid(ex, f, family_choice(ex, host.data(), f)));
    }));
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

int32_t msv_batch_tuned_grid(uint32_t family, uint32_t jobs, uint32_t mean_len) {
  if (family >= (uint32_t)kBatchFamilies) return -1;
  return tune_get_current("batch", batch_tune_key((int)family, jobs, mean_len));
}

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
