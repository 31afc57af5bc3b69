// Host-side internals of libmsv: executor, vectors, caching allocator, error
// reporting, and the per-call launch context handed to kernel modules.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/msv.h"

namespace msv {

// ---- errors ---------------------------------------------------------------
void set_error(msv_status code, uint64_t left, uint64_t right, const char* fmt, ...);
void clear_error();

#define MSV_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      ::msv::set_error(MSV_ERR_LAUNCH, 0, 0, "%s: %s (%s:%d)", #expr,                   \
                       hipGetErrorString(_e), __FILE__, __LINE__);                     \
      return MSV_ERR_LAUNCH;                                                           \
    }                                                                                  \
  } while (0)

#define MSV_TRY(expr)                     \
  do {                                    \
    msv_status _s = (expr);               \
    if (_s != MSV_OK) return _s;          \
  } while (0)

inline size_t dtype_size(msv_dtype d) {
  switch (d) {
    case MSV_U32:
    case MSV_I32:
    case MSV_F32:
      return 4;
    case MSV_U64:
    case MSV_I64:
    case MSV_F64:
      return 8;
  }
  return 0;
}
inline bool dtype_valid(int d) { return d >= MSV_U32 && d <= MSV_F64; }
// ceil(n / d) for lengths up to 2^32 - 1: the sum is formed in 64 bits, so
// n + d - 1 cannot wrap.
inline uint32_t div_up(uint64_t n, uint64_t d) { return (uint32_t)((n + d - 1) / d); }
inline bool dtype_is_float(msv_dtype d) { return d == MSV_F32 || d == MSV_F64; }

// ---- caching allocator (one per executor) ---------------------------------
class Allocator {
 public:
  explicit Allocator(int device) : device_(device) {}
  // Returns nullptr and sets the error on failure.
  void* alloc(size_t bytes);
  void free(void* p);
  void trim();  // caller synchronises first
  size_t cached_bytes() const { return cached_; }

 private:
  static size_t round(size_t bytes);
  int device_;
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> sizes_;
  size_t cached_ = 0;
};

}  // namespace msv

struct msv_executor {
  uint64_t id;
  int device;
  int num_cus;
  hipStream_t stream;
  // Device signature for tuning entries: architecture name and CU count.
  char signature[96];
  // Second stream for work one call runs beside its main work (created on
  // first use, default priority), and the events that fork it from and join
  // it back to `stream`.
  hipStream_t aux_stream = nullptr;
  hipEvent_t fork_event = nullptr;
  hipEvent_t join_event = nullptr;
  std::atomic<int> refs;
  bool destroyed;
  msv::Allocator allocator;
  // Decoupled-lookback state shared by every lookback kernel on this
  // executor's stream: a tile counter (for kernels that draw tile ids), a
  // growable status-word array, and the generation of the next launch.
  uint64_t* d_counter;
  uint64_t* d_status;
  size_t status_words;
  uint32_t lookback_gen;
  // Tuning options (msv_executor_set_option).
  int64_t opt_seg_launch = MSV_SEG_LAUNCH_PERSISTENT;
  int64_t opt_autotune = 1;
  int64_t opt_segsort_config = -1;
  int64_t opt_radix_onesweep = -1;
  int64_t opt_segsort_wave = 1;
  // Resident blocks per CU of each batched job-kernel family (0 = not yet
  // queried). Per executor, so per device; atomic because server scheduler
  // threads and user threads may launch batches on one executor at once
  // (they store the same value).
  std::atomic<int> batch_per_cu[8] = {};
  // HIP graph capture (msv_executor_begin_capture): while capturing, call
  // scratch and pinned staging are kept for the graph's lifetime.
  bool capturing = false;
  std::vector<void*> capture_keep;
  std::vector<void*> capture_pinned;
  explicit msv_executor(int dev) : allocator(dev) {}
};

struct msv_vec {
  std::atomic<int> refs;
  msv_executor* ex;
  uint64_t owner_id;
  msv_dtype dtype;
  uint32_t capacity;
  void* ptr;
  bool owned;
  uint32_t host_len;   // logical length when d_len == nullptr
  uint32_t* d_len;     // device-resident logical length (owned), or nullptr
};

struct msv_segmentation {
  msv_vec* offsets;     // S+1 offsets, retained
  uint32_t segment_count;
  uint32_t value_count;
};

namespace msv {

void executor_retain(msv_executor* ex);
void executor_release(msv_executor* ex);

// Resolved view of a msv_slice.
struct View {
  msv_vec* vec = nullptr;
  void* ptr = nullptr;  // element pointer at slice offset
  uint32_t len = 0;
  msv_dtype dtype = MSV_U32;
};

// Fails with MSV_ERR_UNSUPPORTED when the executor is capturing a graph (for
// calls that need a host synchronisation).
msv_status no_capture(msv_executor* ex, const char* what);

// Validates ownership/range and resolves the length (synchronising only for a
// MSV_WHOLE slice of a vector whose length is device-resident).
msv_status resolve(msv_executor* ex, const msv_slice& s, View* out);

// Creates a new owned output vector.
msv_status new_vec(msv_executor* ex, msv_dtype dtype, uint32_t capacity, bool device_len,
                   msv_vec** out);

// Per-call launch context: scratch memory returned to the executor's cache at
// the end of the call (stream order makes reuse by later work safe).
class Ctx {
 public:
  explicit Ctx(msv_executor* ex) : ex_(ex) {}
  ~Ctx();
  hipStream_t stream() const { return ex_->stream; }
  // Makes the auxiliary stream wait for the work enqueued on stream() so far
  // and returns it (nullptr + error on failure). Every fork is followed by a
  // join before the call returns; scratch and look-back state the forked work
  // uses are requested before the fork.
  hipStream_t fork();
  // Makes stream() wait for the work enqueued on the auxiliary stream.
  msv_status join();
  int num_cus() const { return ex_->num_cus; }
  msv_executor* ex() const { return ex_; }
  // nullptr + error on failure.
  void* scratch(size_t bytes);
  // Host staging for a host-to-device copy: the caller's memory when not
  // capturing, else a pinned copy kept for the graph's lifetime.
  const void* staging(const void* host, size_t bytes);
  // Lookback state sized for `words` status words, and a fresh generation
  // for one launch.
  msv_status lookback(size_t words, uint64_t** counter, uint64_t** status, uint32_t* gen);
  // Arena mode (entry points that allocate nothing): scratch() carves
  // 256-byte-aligned pieces from the device region [base, base + bytes) and
  // fails with MSV_ERR_LENGTH_TOO_LARGE when it is exhausted; lookback()
  // takes its words from the same region with generation `gen` (while
  // capturing: cleared by a kernel node and generation 1).
  void use_arena(void* base, size_t bytes, uint32_t gen) {
    arena_ = (char*)base;
    arena_bytes_ = bytes;
    arena_used_ = 0;
    arena_gen_ = gen;
  }

 private:
  msv_executor* ex_;
  std::vector<void*> scratch_;
  char* arena_ = nullptr;
  size_t arena_bytes_ = 0;
  size_t arena_used_ = 0;
  uint32_t arena_gen_ = 0;
};

// Arena bytes one scratch() request of `bytes` takes (scratch() keeps every
// piece 256-byte aligned).
inline size_t arena_round(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

// ---- kernel module entry points (implemented in the .hip modules) ----------
// All take device pointers and lengths already validated by the API layer.
msv_status scan_run(Ctx& c, msv_dtype dt, msv_op op, bool exclusive, const void* in, void* out,
                    uint32_t n, const void* init_host);
msv_status reduce_run(Ctx& c, msv_dtype dt, msv_op op, const void* in, uint32_t n,
                      const void* init_host, void* out_host);
// reduce into device memory (no host sync).
msv_status reduce_run_device(Ctx& c, msv_dtype dt, msv_op op, const void* in, uint32_t n,
                             const void* init_host, void* d_out);
// copy_where with a built-in predicate on the input itself; the count goes to
// the device word out_len.
msv_status select_pred_run(Ctx& c, msv_dtype dt, const void* in, const msv_predicate* p,
                           bool invert, uint32_t n, void* out, uint32_t* out_len);
msv_status gather_run(Ctx& c, msv_dtype dt, const void* values, uint32_t vlen,
                      const uint32_t* idx, void* out, uint32_t n);
msv_status scatter_run(Ctx& c, msv_dtype dt, const void* values, const uint32_t* idx,
                       uint32_t n, void* out, uint32_t olen);
// Whether scatter_run takes the binned form for these sizes.
bool scatter_binned(msv_dtype dt, uint32_t n, uint32_t olen);
msv_status iota_run(Ctx& c, uint32_t* out, uint32_t n, uint32_t start);
// Seeded synthetic input of the library generator: full-range bits for
// integers, [0.5, 1) for floats.
msv_status uniform_run(Ctx& c, msv_dtype dt, void* out, uint32_t n, uint64_t seed);

// The PCG hash of the seeded streams (massively/src/api/util/random.rs:12-22)
// and the key a 64-bit seed selects (element i of a stream is
// pcg_hash32(i ^ key), mapped to the dtype's range).
__host__ __device__ inline uint32_t pcg_hash32(uint32_t input) {
  const uint32_t state = input * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
inline uint32_t seed_key(uint64_t seed) {
  return pcg_hash32((uint32_t)seed ^ pcg_hash32((uint32_t)(seed >> 32)));
}
// Every job through the device-wide kernels, one after another (the
// one-launch-per-job form the batched path is compared with).
msv_status run_jobs_plain(Ctx& c, const msv_job* jobs, uint32_t count);
// mode: 0 total order, 1 Less, 2 Greater; vsize 0 (keys only), 4 or 8.
msv_status radix_sort_run(Ctx& c, msv_dtype kdt, int mode, size_t vsize, const void* keys,
                          const void* vals, uint32_t n, void* out_k, void* out_v);
// The by-key scans over runs of equal keys (a run head is an element whose
// key differs from its predecessor's): RUNS_INCLUSIVE / RUNS_EXCLUSIVE write
// one value per element; RUNS_REDUCE writes each run's first key to out_keys
// and init (+) the run's values to out, and the run count to both device
// lengths.
enum RunsKind : int { RUNS_INCLUSIVE = 0, RUNS_EXCLUSIVE = 1, RUNS_REDUCE = 2 };
msv_status runs_scan_run(Ctx& c, msv_dtype kdt, const void* keys, msv_dtype vdt, msv_op op,
                         int kind, const void* in, uint32_t n, void* out, const void* init,
                         void* out_keys, uint32_t* len_keys, uint32_t* len_vals);
// unique_by_key: the values at run heads, compacted; the count goes to *out_len.
msv_status unique_runs_run(Ctx& c, msv_dtype kdt, const void* keys, size_t vsize,
                           const void* vals, uint32_t n, void* out, uint32_t* out_len);
// Once per device, with `device` current: resolves the device address of the
// static scratch array of the no-allocation entry points and clears it.
msv_status noalloc_setup(int device, hipStream_t stream);

// Autotuned choices, per device signature, family (a short word such as
// "segsort") and key: -1 when absent. With MSV_TUNING_CACHE set, the file it
// names is read at first use and rewritten after every new entry.
int tune_get(const msv_executor* ex, const char* family, uint32_t key);
void tune_put(const msv_executor* ex, const char* family, uint32_t key, int choice);
// The same lookup for the current device (outside any executor).
int tune_get_current(const char* family, uint32_t key);
// Enqueues a batch of validated-on-entry jobs (msv_run_jobs).
msv_status run_jobs(Ctx& c, const msv_job* jobs, uint32_t count);

// Batches whose launches are captured once and replayed (the serving loop).
// A batch table holds up to `capacity` small jobs and as many generation
// entries; its launches read everything from the table, so their grids and
// arguments do not depend on the batch.
//
// A generation entry: before the jobs run, the seeded stream of `seed_key`
// (the bits uniform_run writes) is written to `in` (unless null) and the
// positions 0..n-1 to `iota` (unless null); after them, the word at
// `count_src` (unless null) is copied to the batch's counts at `count_index`.
struct BatchGen {
  void* in;
  uint32_t* iota;
  const uint32_t* count_src;
  uint32_t n, dtype, key, count_index;
  uint32_t chunk0;  // generation chunks before this entry (filled by batch_table_fill)
  uint32_t pad;
};
size_t batch_table_bytes(uint32_t capacity);
// Whether a job runs inside the batched kernels (else device-wide).
bool batch_job_small(const msv_job& j);
// Validates `jobs` (all small, at most `capacity`) and writes the table for
// them and `gens` into `host` (batch_table_bytes(capacity) bytes). Sort jobs
// above the tile size get element scratch from `c`. *mask: the job families
// present (bit f).
msv_status batch_table_fill(Ctx& c, const msv_job* jobs, uint32_t count, const BatchGen* gens,
                            uint32_t ngen, uint32_t capacity, char* host, uint32_t* mask);
// Enqueues on `stream`: the upload of the host table to `dev`, the input
// generation, one persistent launch per family in `mask` (autotuned
// grids, resident-size until tuned), the count gather into d_counts and its download to h_counts
// (capacity words each). Capturable: no memset, no allocation.
// Resolves the batched kernels' occupancy (before a capture of the enqueue).
void batch_occupancy_warm(msv_executor* ex);
// Tuning-table key of a family's grid for a batch of `jobs` jobs of mean
// length `mean_len`.
uint32_t batch_tune_key(int family, uint32_t jobs, uint32_t mean_len);
// Times each grid candidate for the families in `mask` whose shape class in
// the host table has no tuned grid (at least 16 jobs, no job writing what a
// job of its family reads), running them on `stream` from `dev` (the table is
// uploaded and its inputs generated first), and records the fastest.
// Synchronises with the stream when it times anything; a no-op while
// capturing or with autotuning off.
msv_status batch_autotune(msv_executor* ex, hipStream_t stream, const char* host, char* dev,
                          uint32_t capacity, uint32_t mask);
// `mask` with the tuned grid choice of each family present (2 bits per
// family from bit 8): batches with the same code launch the same grids.
uint32_t batch_shape_code(msv_executor* ex, const char* host, uint32_t mask);
msv_status batch_table_enqueue(msv_executor* ex, hipStream_t stream, const char* host, char* dev,
                               uint32_t capacity, uint32_t mask, uint32_t* d_counts,
                               uint32_t* h_counts);
// Stable sort of every segment of offsets[0..nseg] on its own (mode 1 Less,
// 2 Greater); vsize 0, 4 or 8.
msv_status seg_sort_run(Ctx& c, msv_dtype kdt, int mode, const void* keys, const void* vals,
                        size_t vsize, uint32_t n, const uint32_t* offsets, uint32_t nseg,
                        void* out_k, void* out_v);

}  // namespace msv
