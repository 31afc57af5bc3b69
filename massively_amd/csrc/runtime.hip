// Executor, DeviceVec and DeviceSlice of the C ABI
// (massively/src/core/runtime.rs:21-545, core/extent.rs, core/error.rs).
#include <cstring>

#include "msv_internal.hpp"

namespace msv {

// ---- errors ---------------------------------------------------------------
static thread_local msv_error g_err = {0, 0, 0, {0}};

void set_error(msv_status code, uint64_t left, uint64_t right, const char* fmt, ...) {
  g_err.code = code;
  g_err.left = left;
  g_err.right = right;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err.message, sizeof(g_err.message), fmt, ap);
  va_end(ap);
}

void clear_error() {
  g_err.code = MSV_OK;
  g_err.left = g_err.right = 0;
  g_err.message[0] = 0;
}

// ---- allocator ------------------------------------------------------------
size_t Allocator::round(size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (bytes < (1u << 20)) {
    size_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  const size_t g = 2u << 20;
  return (bytes + g - 1) / g * g;
}

void* Allocator::alloc(size_t bytes) {
  const size_t r = round(bytes);
  {
    // Best fit among the cached blocks of at most twice the rounded size: with
    // requests of varied sizes an exact match is rare, and every miss would
    // leave one more block in the cache (until hipMalloc fails and the whole
    // cache is freed block by block, which stalls for seconds).
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.lower_bound(r);
    if (it != free_.end() && it->first <= 2 * r) {
      void* p = it->second;
      cached_ -= it->first;
      free_.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, r);
  if (e != hipSuccess) {
    // Give cached blocks back and retry once.
    (void)hipGetLastError();
    (void)hipDeviceSynchronize();
    trim();
    e = hipMalloc(&p, r);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(MSV_ERR_OUT_OF_MEMORY, bytes, 0, "device allocation of %zu bytes failed: %s", r,
              hipGetErrorString(e));
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(mu_);
  sizes_[p] = r;
  return p;
}

void Allocator::free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  free_.emplace(it->second, p);
  cached_ += it->second;
}

void Allocator::trim() {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    sizes_.erase(kv.second);
  }
  free_.clear();
  cached_ = 0;
}

// ---- executor lifetime ------------------------------------------------------
static std::atomic<uint64_t> g_next_executor_id{1};

void executor_retain(msv_executor* ex) { ex->refs.fetch_add(1); }

void executor_release(msv_executor* ex) {
  if (ex->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(ex->device);
  (void)hipStreamSynchronize(ex->stream);
  if (ex->d_status) (void)hipFree(ex->d_status);
  if (ex->d_counter) (void)hipFree(ex->d_counter);
  ex->allocator.trim();
  if (ex->aux_stream) {
    (void)hipStreamSynchronize(ex->aux_stream);
    (void)hipStreamDestroy(ex->aux_stream);
  }
  if (ex->fork_event) (void)hipEventDestroy(ex->fork_event);
  if (ex->join_event) (void)hipEventDestroy(ex->join_event);
  (void)hipStreamDestroy(ex->stream);
  delete ex;
}

msv_status no_capture(msv_executor* ex, const char* what) {
  if (ex && ex->capturing) {
    set_error(MSV_ERR_UNSUPPORTED, 0, 0, "%s needs a host synchronisation; not allowed while "
              "capturing a graph", what);
    return MSV_ERR_UNSUPPORTED;
  }
  return MSV_OK;
}

// ---- views ------------------------------------------------------------------
msv_status resolve(msv_executor* ex, const msv_slice& s, View* out) {
  if (!ex) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (!s.vec) {
    set_error(MSV_ERR_UNBOUND_COLUMN, 0, 0, "slice has no device vector");
    return MSV_ERR_UNBOUND_COLUMN;
  }
  msv_vec* v = s.vec;
  if (v->owner_id != ex->id) {
    set_error(MSV_ERR_FOREIGN_EXECUTOR, 0, 0, "executor %llu does not own this device data",
              (unsigned long long)ex->id);
    return MSV_ERR_FOREIGN_EXECUTOR;
  }
  uint32_t logical = v->host_len;
  if (v->d_len && s.len == MSV_WHOLE) {
    MSV_TRY(no_capture(ex, "a whole-length slice of a vector with a device-resident length"));
    MSV_HIP(hipSetDevice(ex->device));
    MSV_HIP(hipMemcpyAsync(&logical, v->d_len, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ex->stream));
    MSV_HIP(hipStreamSynchronize(ex->stream));
    if (logical > v->capacity) logical = v->capacity;
  }
  if (s.offset > logical && s.len == MSV_WHOLE) {
    set_error(MSV_ERR_INVALID_ARGUMENT, s.offset, logical, "slice start %u beyond length %u",
              s.offset, logical);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  const uint32_t len = (s.len == MSV_WHOLE) ? logical - s.offset : s.len;
  if ((uint64_t)s.offset + len > v->capacity) {
    set_error(MSV_ERR_INVALID_ARGUMENT, (uint64_t)s.offset + len, v->capacity,
              "slice range %u..%llu exceeds capacity %u", s.offset,
              (unsigned long long)((uint64_t)s.offset + len), v->capacity);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  out->vec = v;
  out->dtype = v->dtype;
  out->len = len;
  out->ptr = (char*)v->ptr + (size_t)s.offset * dtype_size(v->dtype);
  return MSV_OK;
}

msv_status new_vec(msv_executor* ex, msv_dtype dtype, uint32_t capacity, bool device_len,
                   msv_vec** out) {
  if (!out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  MSV_HIP(hipSetDevice(ex->device));
  const size_t bytes = (size_t)capacity * dtype_size(dtype);
  void* p = ex->allocator.alloc(bytes ? bytes : 1);
  if (!p) return MSV_ERR_OUT_OF_MEMORY;
  uint32_t* dl = nullptr;
  if (device_len) {
    dl = (uint32_t*)ex->allocator.alloc(sizeof(uint32_t));
    if (!dl) {
      ex->allocator.free(p);
      return MSV_ERR_OUT_OF_MEMORY;
    }
  }
  msv_vec* v = new msv_vec;
  v->refs = 1;
  v->ex = ex;
  v->owner_id = ex->id;
  v->dtype = dtype;
  v->capacity = capacity;
  v->ptr = p;
  v->owned = true;
  v->host_len = capacity;
  v->d_len = dl;
  executor_retain(ex);
  *out = v;
  return MSV_OK;
}

hipStream_t Ctx::fork() {
  msv_executor* ex = ex_;
  hipError_t e = hipSuccess;
  if (!ex->aux_stream) e = hipStreamCreateWithFlags(&ex->aux_stream, hipStreamNonBlocking);
  if (e == hipSuccess && !ex->fork_event)
    e = hipEventCreateWithFlags(&ex->fork_event, hipEventDisableTiming);
  if (e == hipSuccess && !ex->join_event)
    e = hipEventCreateWithFlags(&ex->join_event, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ex->fork_event, ex->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(ex->aux_stream, ex->fork_event, 0);
  if (e != hipSuccess) {
    set_error(MSV_ERR_LAUNCH, 0, 0, "stream fork failed: %s", hipGetErrorString(e));
    return nullptr;
  }
  return ex->aux_stream;
}

msv_status Ctx::join() {
  MSV_HIP(hipEventRecord(ex_->join_event, ex_->aux_stream));
  MSV_HIP(hipStreamWaitEvent(ex_->stream, ex_->join_event, 0));
  return MSV_OK;
}

Ctx::~Ctx() {
  for (void* p : scratch_) ex_->allocator.free(p);
}

void* Ctx::scratch(size_t bytes) {
  if (arena_) {
    const size_t need = arena_round(bytes);
    if (need > arena_bytes_ - arena_used_) {
      set_error(MSV_ERR_LENGTH_TOO_LARGE, arena_used_ + need, arena_bytes_,
                "static scratch exhausted: %zu bytes needed, %zu available", arena_used_ + need,
                arena_bytes_);
      return nullptr;
    }
    void* p = arena_ + arena_used_;
    arena_used_ += need;
    return p;
  }
  void* p = ex_->allocator.alloc(bytes);
  if (p) (ex_->capturing ? ex_->capture_keep : scratch_).push_back(p);
  return p;
}

const void* Ctx::staging(const void* host, size_t bytes) {
  if (!ex_->capturing) return host;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    set_error(MSV_ERR_OUT_OF_MEMORY, bytes, 0, "pinned staging allocation failed");
    return nullptr;
  }
  memcpy(p, host, bytes);
  ex_->capture_pinned.push_back(p);
  return p;
}

static __global__ void zero_u64_kernel(uint64_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

msv_status Ctx::lookback(size_t words, uint64_t** counter, uint64_t** status, uint32_t* gen) {
  msv_executor* ex = ex_;
  if (arena_) {
    uint64_t* p = (uint64_t*)scratch((words + 1) * 8);
    if (!p) return MSV_ERR_LENGTH_TOO_LARGE;
    *counter = p;
    *status = p + 1;
    if (ex->capturing) {
      // Replays reuse one generation, so the words are cleared before each.
      hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)((words + 1 + 255) / 256)), dim3(256), 0,
                         ex->stream, p, words + 1);
      MSV_HIP(hipGetLastError());
      *gen = 1;
    } else {
      *gen = arena_gen_;
    }
    return MSV_OK;
  }
  if (ex->capturing) {
    // A captured launch gets its own counter and status words, cleared by a
    // kernel node before it on every replay, and a fixed generation.
    uint64_t* p = (uint64_t*)scratch((words + 1) * 8);
    if (!p) return MSV_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)((words + 1 + 255) / 256)), dim3(256), 0,
                       ex->stream, p, words + 1);
    MSV_HIP(hipGetLastError());
    *counter = p;
    *status = p + 1;
    *gen = 1;
    return MSV_OK;
  }
  // Generations are 29 bits; before one would repeat, every status word is
  // cleared so no stale word can match a live generation.
  if (ex->lookback_gen >= (1u << 29) - 1) {
    if (ex->d_status) MSV_HIP(hipMemsetAsync(ex->d_status, 0, ex->status_words * 8, ex->stream));
    ex->lookback_gen = 1;
  }
  if (words > ex->status_words) {
    size_t grow = ex->status_words ? ex->status_words : (1u << 16);
    while (grow < words) grow *= 2;
    if (ex->d_status) {
      MSV_HIP(hipStreamSynchronize(ex->stream));
      MSV_HIP(hipFree(ex->d_status));
      ex->d_status = nullptr;
      ex->status_words = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, grow * 8);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error(MSV_ERR_OUT_OF_MEMORY, grow * 8, 0, "lookback status allocation failed");
      return MSV_ERR_OUT_OF_MEMORY;
    }
    MSV_HIP(hipMemsetAsync(p, 0, grow * 8, ex->stream));
    ex->d_status = (uint64_t*)p;
    ex->status_words = grow;
  }
  *counter = ex->d_counter;
  *status = ex->d_status;
  *gen = ex->lookback_gen++;
  return MSV_OK;
}

}  // namespace msv

using namespace msv;

extern "C" {

uint32_t msv_abi_version(void) { return MSV_ABI_VERSION; }

void msv_last_error(msv_error* out) {
  if (out) *out = g_err;
}

msv_status msv_executor_create(int32_t device, msv_executor** out) {
  clear_error();
  if (!out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  int count = 0;
  MSV_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    set_error(MSV_ERR_INVALID_ARGUMENT, (uint64_t)device, (uint64_t)count,
              "device %d not present (%d devices)", device, count);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MSV_HIP(hipGetDeviceProperties(&prop, device));
  msv_executor* ex = new msv_executor(device);
  ex->id = g_next_executor_id.fetch_add(1);
  ex->device = device;
  ex->num_cus = prop.multiProcessorCount;
  snprintf(ex->signature, sizeof ex->signature, "%s/%d", prop.gcnArchName, prop.multiProcessorCount);
  for (char* q = ex->signature; *q; ++q)
    if (*q == ' ') *q = '_';
  ex->refs = 1;
  ex->destroyed = false;
  ex->d_counter = nullptr;
  ex->d_status = nullptr;
  ex->status_words = 0;
  ex->lookback_gen = 1;
  hipError_t e = hipStreamCreateWithFlags(&ex->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&ex->d_counter, 8);
  if (e == hipSuccess) e = hipMemsetAsync(ex->d_counter, 0, 8, ex->stream);
  if (e != hipSuccess) {
    set_error(MSV_ERR_LAUNCH, 0, 0, "executor creation failed: %s", hipGetErrorString(e));
    delete ex;
    return MSV_ERR_LAUNCH;
  }
  // Per-device set-up of the static scratch of the no-allocation entry points.
  const msv_status ns = noalloc_setup(device, ex->stream);
  if (ns != MSV_OK) {
    (void)hipStreamDestroy(ex->stream);
    (void)hipFree(ex->d_counter);
    delete ex;
    return ns;
  }
  *out = ex;
  return MSV_OK;
}

msv_status msv_executor_destroy(msv_executor* ex) {
  clear_error();
  if (!ex || ex->destroyed) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null or destroyed executor");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  ex->destroyed = true;
  executor_release(ex);
  return MSV_OK;
}

msv_status msv_executor_sync(msv_executor* ex) {
  clear_error();
  if (!ex) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(no_capture(ex, "Executor::sync"));
  MSV_HIP(hipSetDevice(ex->device));
  MSV_HIP(hipStreamSynchronize(ex->stream));
  return MSV_OK;
}

uint64_t msv_executor_id(const msv_executor* ex) { return ex ? ex->id : 0; }

void* msv_executor_stream(msv_executor* ex) { return ex ? (void*)ex->stream : nullptr; }

msv_status msv_executor_trim(msv_executor* ex) {
  clear_error();
  if (!ex) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(no_capture(ex, "trim"));
  MSV_HIP(hipSetDevice(ex->device));
  MSV_HIP(hipStreamSynchronize(ex->stream));
  ex->allocator.trim();
  return MSV_OK;
}

// ---- graph capture ------------------------------------------------------------
struct msv_graph {
  msv_executor* ex;
  hipGraph_t graph;
  hipGraphExec_t exec;
  std::vector<void*> keep;
  std::vector<void*> pinned;
};

msv_status msv_executor_begin_capture(msv_executor* ex) {
  clear_error();
  if (!ex || ex->capturing) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor or already capturing");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_HIP(hipSetDevice(ex->device));
  MSV_HIP(hipStreamBeginCapture(ex->stream, hipStreamCaptureModeRelaxed));
  ex->capturing = true;
  return MSV_OK;
}

msv_status msv_executor_end_capture(msv_executor* ex, msv_graph** out) {
  clear_error();
  if (!ex || !ex->capturing || !out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "executor is not capturing or null output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  ex->capturing = false;
  auto* g = new msv_graph();
  g->ex = ex;
  g->keep.swap(ex->capture_keep);
  g->pinned.swap(ex->capture_pinned);
  hipError_t e = hipStreamEndCapture(ex->stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    for (void* p : g->keep) ex->allocator.free(p);
    for (void* p : g->pinned) (void)hipHostFree(p);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    set_error(MSV_ERR_LAUNCH, 0, 0, "graph capture failed: %s", hipGetErrorString(e));
    return MSV_ERR_LAUNCH;
  }
  executor_retain(ex);
  *out = g;
  return MSV_OK;
}

msv_status msv_graph_launch(msv_graph* g) {
  clear_error();
  if (!g) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null graph");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_HIP(hipSetDevice(g->ex->device));
  MSV_HIP(hipGraphLaunch(g->exec, g->ex->stream));
  return MSV_OK;
}

msv_status msv_graph_destroy(msv_graph* g) {
  clear_error();
  if (!g) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null graph");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  msv_executor* ex = g->ex;
  (void)hipSetDevice(ex->device);
  (void)hipStreamSynchronize(ex->stream);  // no replay may still use the memory
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  for (void* p : g->keep) ex->allocator.free(p);
  for (void* p : g->pinned) (void)hipHostFree(p);
  delete g;
  executor_release(ex);
  return MSV_OK;
}

msv_status msv_executor_set_option(msv_executor* ex, msv_option opt, int64_t value) {
  clear_error();
  if (!ex) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  switch (opt) {
    case MSV_OPT_SEG_LAUNCH:
      if (value != MSV_SEG_LAUNCH_PERSISTENT && value != MSV_SEG_LAUNCH_PER_CLASS) break;
      ex->opt_seg_launch = value;
      return MSV_OK;
    case MSV_OPT_AUTOTUNE:
      if (value != 0 && value != 1) break;
      ex->opt_autotune = value;
      return MSV_OK;
    case MSV_OPT_SEGSORT_CONFIG:
      if (value < -1 || value >= (int64_t)msv_segsort_config_count()) break;
      ex->opt_segsort_config = value;
      return MSV_OK;
    case MSV_OPT_RADIX_ONESWEEP:
      if (value < -1 || value > 1) break;
      ex->opt_radix_onesweep = value;
      return MSV_OK;
    case MSV_OPT_SEGSORT_WAVE:
      if (value != 0 && value != 1) break;
      ex->opt_segsort_wave = value;
      return MSV_OK;
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, (uint64_t)opt, (uint64_t)value, "bad option %d = %lld",
            (int)opt, (long long)value);
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status msv_executor_get_option(msv_executor* ex, msv_option opt, int64_t* value) {
  clear_error();
  if (!ex || !value) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor or output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  switch (opt) {
    case MSV_OPT_SEG_LAUNCH: *value = ex->opt_seg_launch; return MSV_OK;
    case MSV_OPT_AUTOTUNE: *value = ex->opt_autotune; return MSV_OK;
    case MSV_OPT_SEGSORT_CONFIG: *value = ex->opt_segsort_config; return MSV_OK;
    case MSV_OPT_RADIX_ONESWEEP: *value = ex->opt_radix_onesweep; return MSV_OK;
    case MSV_OPT_SEGSORT_WAVE: *value = ex->opt_segsort_wave; return MSV_OK;
  }
  set_error(MSV_ERR_INVALID_ARGUMENT, (uint64_t)opt, 0, "bad option %d", (int)opt);
  return MSV_ERR_INVALID_ARGUMENT;
}

msv_status msv_alloc(msv_executor* ex, msv_dtype dtype, uint32_t len, msv_vec** out) {
  clear_error();
  if (!ex || !dtype_valid(dtype)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor or bad dtype");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return new_vec(ex, dtype, len, false, out);
}

msv_status msv_to_device(msv_executor* ex, msv_dtype dtype, const void* host, uint32_t len,
                         msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "to_device"));
  if (!ex || !dtype_valid(dtype) || (len && !host)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor/host pointer or bad dtype");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, dtype, len, false, out));
  if (len) {
    hipError_t e = hipMemcpyAsync((*out)->ptr, host, (size_t)len * dtype_size(dtype),
                                  hipMemcpyHostToDevice, ex->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ex->stream);
    if (e != hipSuccess) {
      msv_vec_release(*out);
      *out = nullptr;
      set_error(MSV_ERR_LAUNCH, 0, 0, "host-to-device copy failed: %s", hipGetErrorString(e));
      return MSV_ERR_LAUNCH;
    }
  }
  return MSV_OK;
}

msv_status msv_to_host(msv_executor* ex, msv_slice src, void* host, uint32_t host_cap,
                       uint32_t* out_len) {
  clear_error();
  MSV_TRY(no_capture(ex, "to_host"));
  View v;
  MSV_TRY(resolve(ex, src, &v));
  uint32_t len = v.len;
  if (src.vec->d_len && src.len != MSV_WHOLE) {
    // Explicit ranges are clipped to the device-resident length as well.
    uint32_t logical = 0;
    MSV_HIP(hipMemcpyAsync(&logical, src.vec->d_len, 4, hipMemcpyDeviceToHost, ex->stream));
    MSV_HIP(hipStreamSynchronize(ex->stream));
    if (src.offset >= logical) len = 0;
    else if ((uint64_t)src.offset + len > logical) len = logical - src.offset;
  }
  if (len > host_cap) {
    set_error(MSV_ERR_OUTPUT_TOO_SHORT, len, host_cap,
              "output length %u is shorter than input length %u", host_cap, len);
    return MSV_ERR_OUTPUT_TOO_SHORT;
  }
  if (len) {
    if (!host) {
      set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null host pointer");
      return MSV_ERR_INVALID_ARGUMENT;
    }
    MSV_HIP(hipSetDevice(ex->device));
    MSV_HIP(hipMemcpyAsync(host, v.ptr, (size_t)len * dtype_size(v.dtype), hipMemcpyDeviceToHost,
                           ex->stream));
  }
  MSV_HIP(hipStreamSynchronize(ex->stream));
  if (out_len) *out_len = len;
  return MSV_OK;
}

msv_status msv_wrap_device(msv_executor* ex, msv_dtype dtype, void* device_ptr, uint32_t len,
                           msv_vec** out) {
  clear_error();
  if (!ex || !dtype_valid(dtype) || !out || (len && !device_ptr)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad wrap arguments");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  msv_vec* v = new msv_vec;
  v->refs = 1;
  v->ex = ex;
  v->owner_id = ex->id;
  v->dtype = dtype;
  v->capacity = len;
  v->ptr = device_ptr;
  v->owned = false;
  v->host_len = len;
  v->d_len = nullptr;
  executor_retain(ex);
  *out = v;
  return MSV_OK;
}

msv_status msv_vec_retain(msv_vec* v) {
  if (!v) return MSV_ERR_INVALID_ARGUMENT;
  v->refs.fetch_add(1);
  return MSV_OK;
}

msv_status msv_vec_release(msv_vec* v) {
  if (!v) return MSV_ERR_INVALID_ARGUMENT;
  if (v->refs.fetch_sub(1) != 1) return MSV_OK;
  msv_executor* ex = v->ex;
  if (v->owned) ex->allocator.free(v->ptr);
  if (v->d_len) ex->allocator.free(v->d_len);
  delete v;
  executor_release(ex);
  return MSV_OK;
}

msv_status msv_vec_len(msv_executor* ex, msv_vec* v, uint32_t* out_len) {
  clear_error();
  msv_slice s = {v, 0, MSV_WHOLE};
  View view;
  MSV_TRY(resolve(ex, s, &view));
  if (out_len) *out_len = view.len;
  return MSV_OK;
}

uint32_t msv_vec_capacity(const msv_vec* v) { return v ? v->capacity : 0; }
int32_t msv_vec_dtype(const msv_vec* v) { return v ? (int32_t)v->dtype : -1; }
void* msv_vec_device_ptr(const msv_vec* v) { return v ? v->ptr : nullptr; }
void* msv_vec_device_len_ptr(const msv_vec* v) { return v ? (void*)v->d_len : nullptr; }

}  // extern "C"
