// Serving loop (no counterpart in the reference): continuous batching of
// scan / reduce / copy_where / sort_by_key requests over several streams,
// admitted in arrival order by an HBM budget.
//
// One scheduler thread owns every HIP call of the loop. Each step it
//  1. retires the batches whose completion event has fired (requests become
//     done; their inputs and scratch go back to the budget, their outputs stay
//     device-resident until released);
//  2. forms new batches while a stream is free: the pending head and the
//     requests behind it, in arrival order, while each fits the budget; a
//     request too large for the batched kernels gets a batch of its own (it
//     fills the device by itself), and a batch of small requests stops at the
//     first large one;
//  3. enqueues each batch on its stream (inputs created on the device, one
//     msv_run_jobs call, the copy_where counts copied to pinned memory, an
//     event) and returns to step 1 without waiting.
// While batches run, new arrivals are admitted onto the other streams, so
// their work overlaps on the device. A large request fills the device by
// itself, so while small requests wait at most half of the streams run large
// requests; small requests behind a large head that cannot be admitted (that
// cap, or the budget) are taken from up to kServeBypass positions behind it,
// so they do not wait for the large ones (bounded bypass: the head keeps its
// place and is admitted as soon as a large slot and the budget allow).
// MSV_SERVE_SERIAL instead runs one request at a time through the device-wide
// kernels, and MSV_SERVE_PER_JOB runs one request per launch through the
// device-wide kernels on every stream (concurrency without batching): the two
// forms the batched mode is measured against.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>

#include "msv_internal.hpp"

namespace msv {

constexpr uint32_t kServeSmallStream = 1u << 17;  // the batched kernels' limits
constexpr uint32_t kServeSmallSort = 1u << 15;
constexpr uint32_t kServeMaxStreams = 16;
constexpr size_t kServeBypass = 1024;  // pending positions a small batch may look past

struct Request {
  uint64_t id = 0;
  msv_request r{};
  std::vector<char> host;  // copy of the caller's input, kept until done
  uint32_t state = 0;      // 0 queued, 1 running, 2 done, 3 failed
  msv_status status = MSV_OK;
  double t_submit = 0, t_admit = 0, t_done = 0;
  uint64_t bytes = 0;      // budget held while running
  uint64_t out_bytes = 0;  // budget held after completion (outputs)
  int slot = -1;
  void* d_in = nullptr;
  void* d_in2 = nullptr;
  void* d_out = nullptr;
  void* d_out2 = nullptr;
  uint32_t out_len = 0;
  uint32_t count_index = 0;  // copy_where: index in the batch's pinned counts
  uint32_t readers = 0;      // msv_server_result copies in flight (release waits for 0)
};

struct Slot {
  msv_executor* ex = nullptr;
  hipEvent_t ev = nullptr;
  uint32_t* h_counts = nullptr;  // pinned, max_batch entries
  bool busy = false;
  std::vector<Request*> reqs;
  // Batched mode: the batch table (pinned host copy and device copy), the
  // device counts, and one captured graph of the batch launches per set of
  // job families present.
  char* h_table = nullptr;
  char* d_table = nullptr;
  uint32_t* d_counts = nullptr;
  std::map<uint32_t, hipGraphExec_t> graphs;
};

}  // namespace msv

struct msv_server {
  msv_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<msv::Request*> pending;
  std::unordered_map<uint64_t, std::unique_ptr<msv::Request>> all;
  std::vector<msv::Slot> slots;
  uint64_t next_id = 1;
  uint64_t used = 0;
  uint64_t outstanding = 0;  // submitted and not yet done
  bool stop = false;
  std::thread worker;
};

namespace msv {

static double now_us(const msv_server* s) {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - s->t0).count();
}

static bool is_small(const msv_request& r) {
  return r.kind == MSV_JOB_SORT_BY_KEY ? r.n <= kServeSmallSort : r.n <= kServeSmallStream;
}

// Device bytes a request holds while it runs (inputs, outputs, library
// scratch) and after it is done (outputs).
static void request_bytes(const msv_request& r, uint64_t* run, uint64_t* keep) {
  const uint64_t n = r.n, es = dtype_size((msv_dtype)r.dtype), vb = r.value_bytes;
  const uint64_t slack = 1u << 20;
  switch (r.kind) {
    case MSV_JOB_REDUCE:
      *keep = 256;
      *run = n * es + *keep + slack;
      break;
    case MSV_JOB_COPY_WHERE:
      *keep = n * es + 256;
      *run = n * es + *keep + slack;
      break;
    case MSV_JOB_SORT_BY_KEY:
      *keep = n * (es + vb);
      // inputs + outputs + two element ping-pong buffers + count matrices
      *run = 4 * n * (es + vb) + 2 * (n / 2048 + 1) * 1024 + slack;
      break;
    default:
      *keep = n * es;
      *run = 2 * n * es + slack;
      break;
  }
}

static void* salloc(msv_executor* ex, uint64_t bytes) {
  return ex->allocator.alloc(bytes ? bytes : 16);
}

// Enqueues a batch of small requests on slot k by replaying the slot's
// captured graph for the families present (capturing it on first use): the
// seeded inputs, the sort positions and the copy_where counts are produced by
// the graph's own kernels from the batch table.
static msv_status launch_batch_graph(msv_server* s, int k, Ctx& c, std::vector<msv_job>& jobs,
                                     std::vector<BatchGen>& gens) {
  Slot& sl = s->slots[k];
  msv_executor* ex = sl.ex;
  const uint32_t cap = s->cfg.max_batch;
  uint32_t mask = 0;
  MSV_TRY(batch_table_fill(c, jobs.data(), (uint32_t)jobs.size(), gens.data(),
                           (uint32_t)gens.size(), cap, sl.h_table, &mask));
  // The first batch of a shape class times the grid candidates (one host
  // synchronisation per class); graphs are kept per family set and grids.
  batch_occupancy_warm(ex);
  MSV_TRY(batch_autotune(ex, c.stream(), sl.h_table, sl.d_table, cap, mask));
  const uint32_t code = batch_shape_code(ex, sl.h_table, mask);
  auto it = sl.graphs.find(code);
  if (it == sl.graphs.end()) {
    MSV_HIP(hipStreamBeginCapture(c.stream(), hipStreamCaptureModeThreadLocal));
    const msv_status st =
        batch_table_enqueue(ex, c.stream(), sl.h_table, sl.d_table, cap, mask, sl.d_counts,
                            sl.h_counts);
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(c.stream(), &graph);
    MSV_TRY(st);
    MSV_HIP(ee);
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MSV_HIP(ie);
    it = sl.graphs.emplace(code, exec).first;
  }
  MSV_HIP(hipGraphLaunch(it->second, c.stream()));
  return MSV_OK;
}

// Enqueues one batch on slot k (scheduler thread, lock not held).
static msv_status launch_batch(msv_server* s, int k) {
  Slot& sl = s->slots[k];
  msv_executor* ex = sl.ex;
  MSV_HIP(hipSetDevice(ex->device));
  Ctx c(ex);
  // Batched mode replays a graph for batches of small requests (a large
  // request always has a batch of its own).
  bool graph = s->cfg.mode == MSV_SERVE_BATCHED && sl.h_table != nullptr;
  for (Request* q : sl.reqs) graph = graph && is_small(q->r);
  std::vector<BatchGen> gens;
  std::vector<msv_job> jobs;
  jobs.reserve(sl.reqs.size());
  uint32_t ncounts = 0;
  for (Request* q : sl.reqs) {
    const msv_request& r = q->r;
    const msv_dtype dt = (msv_dtype)r.dtype;
    const size_t es = dtype_size(dt);
    q->slot = k;
    q->d_in = salloc(ex, (uint64_t)r.n * es);
    if (!q->d_in) return MSV_ERR_OUT_OF_MEMORY;
    BatchGen g{};
    g.n = r.n;
    g.dtype = r.dtype;
    if (!q->host.empty()) {
      MSV_HIP(hipMemcpyAsync(q->d_in, q->host.data(), (size_t)r.n * es, hipMemcpyHostToDevice,
                             c.stream()));
    } else if (graph) {
      g.in = q->d_in;
      g.key = seed_key(r.seed);
    } else {
      MSV_TRY(uniform_run(c, dt, q->d_in, r.n, r.seed));
    }
    msv_job j{};
    j.kind = r.kind;
    j.dtype = r.dtype;
    j.op = r.op;
    j.n = r.n;
    j.invert = r.invert;
    j.pred = r.pred;
    j.init = r.init;
    j.in = q->d_in;
    switch (r.kind) {
      case MSV_JOB_REDUCE:
        q->d_out = salloc(ex, 256);
        if (!q->d_out) return MSV_ERR_OUT_OF_MEMORY;
        break;
      case MSV_JOB_COPY_WHERE:
        q->d_out = salloc(ex, (uint64_t)r.n * es);
        q->d_out2 = salloc(ex, 256);
        if (!q->d_out || !q->d_out2) return MSV_ERR_OUT_OF_MEMORY;
        q->count_index = ncounts++;
        g.count_src = (const uint32_t*)q->d_out2;
        g.count_index = q->count_index;
        break;
      case MSV_JOB_SORT_BY_KEY:
        q->d_out2 = salloc(ex, (uint64_t)r.n * es);  // sorted keys
        if (!q->d_out2) return MSV_ERR_OUT_OF_MEMORY;
        if (r.value_bytes) {
          q->d_in2 = salloc(ex, (uint64_t)r.n * 4);
          q->d_out = salloc(ex, (uint64_t)r.n * 4);
          if (!q->d_in2 || !q->d_out) return MSV_ERR_OUT_OF_MEMORY;
          if (graph) g.iota = (uint32_t*)q->d_in2;
          else MSV_TRY(iota_run(c, (uint32_t*)q->d_in2, r.n, 0));
          j.value_bytes = 4;
        }
        break;
      default:
        q->d_out = salloc(ex, (uint64_t)r.n * es);
        if (!q->d_out) return MSV_ERR_OUT_OF_MEMORY;
        break;
    }
    j.in2 = q->d_in2;
    j.out = q->d_out;
    j.out2 = q->d_out2;
    jobs.push_back(j);
    if (g.in || g.iota || g.count_src) gens.push_back(g);
  }
  if (graph) {
    MSV_TRY(launch_batch_graph(s, k, c, jobs, gens));
    MSV_HIP(hipEventRecord(sl.ev, c.stream()));
    return MSV_OK;
  }
  if (s->cfg.mode != MSV_SERVE_BATCHED) {
    MSV_TRY(run_jobs_plain(c, jobs.data(), (uint32_t)jobs.size()));
  } else {
    MSV_TRY(run_jobs(c, jobs.data(), (uint32_t)jobs.size()));
  }
  for (Request* q : sl.reqs) {
    if (q->r.kind == MSV_JOB_COPY_WHERE)
      MSV_HIP(hipMemcpyAsync(&sl.h_counts[q->count_index], q->d_out2, 4, hipMemcpyDeviceToHost,
                             c.stream()));
  }
  MSV_HIP(hipEventRecord(sl.ev, c.stream()));
  return MSV_OK;
}

// Frees the buffers a finished request no longer needs (scheduler thread).
static void drop_inputs(msv_server* s, Request* q) {
  std::vector<char>().swap(q->host);
  if (q->slot < 0) return;
  msv_executor* ex = s->slots[q->slot].ex;
  if (q->d_in) ex->allocator.free(q->d_in);
  if (q->d_in2) ex->allocator.free(q->d_in2);
  q->d_in = q->d_in2 = nullptr;
  if (q->r.kind == MSV_JOB_COPY_WHERE && q->d_out2) {
    ex->allocator.free(q->d_out2);
    q->d_out2 = nullptr;
  }
}

static void drop_outputs(msv_server* s, Request* q) {
  if (q->slot < 0) return;
  msv_executor* ex = s->slots[q->slot].ex;
  if (q->d_out) ex->allocator.free(q->d_out);
  if (q->d_out2) ex->allocator.free(q->d_out2);
  q->d_out = q->d_out2 = nullptr;
}

static void scheduler(msv_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  hipSetDevice(s->cfg.device);
  for (;;) {
    bool progress = false;
    // 1. retire
    for (size_t k = 0; k < s->slots.size(); ++k) {
      Slot& sl = s->slots[k];
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.ev);
      if (q == hipErrorNotReady) continue;
      const double t = now_us(s);
      for (Request* r : sl.reqs) {
        r->t_done = t;
        if (q != hipSuccess) {
          r->state = 3;
          r->status = MSV_ERR_LAUNCH;
        } else {
          r->state = 2;
          r->out_len = r->r.kind == MSV_JOB_COPY_WHERE ? sl.h_counts[r->count_index]
                       : r->r.kind == MSV_JOB_REDUCE   ? 1u
                                                        : r->r.n;
        }
        drop_inputs(s, r);
        s->used -= r->bytes - r->out_bytes;
        --s->outstanding;
      }
      sl.reqs.clear();
      sl.busy = false;
      progress = true;
    }
    if (progress) s->cv_done.notify_all();
    // 2. admit
    while (!s->pending.empty()) {
      int k = -1;
      for (size_t i = 0; i < s->slots.size(); ++i)
        if (!s->slots[i].busy) {
          k = (int)i;
          break;
        }
      if (k < 0) break;
      Slot& sl = s->slots[k];
      const bool batched = s->cfg.mode == MSV_SERVE_BATCHED;
      const uint32_t cap = batched ? s->cfg.max_batch : 1u;
      if (!batched) {
        // One request per launch, in arrival order.
        Request* q = s->pending.front();
        if (s->used + q->bytes <= s->cfg.hbm_budget) {
          s->used += q->bytes;
          s->pending.pop_front();
          sl.reqs.push_back(q);
        }
      } else {
        uint32_t large_busy = 0;
        for (const Slot& o : s->slots) large_busy += (o.busy && !o.reqs.empty() && !is_small(o.reqs[0]->r)) ? 1u : 0u;
        const uint32_t large_cap = std::max<uint32_t>(1u, (uint32_t)s->slots.size() / 2);
        // The cap on large requests only holds while small ones wait within
        // the bypass window; with none waiting, large requests take any free
        // stream.
        bool small_waiting = false;
        for (size_t i = 0; i < s->pending.size() && i < kServeBypass && !small_waiting; ++i)
          small_waiting = is_small(s->pending[i]->r);
        Request* head = s->pending.front();
        if (!is_small(head->r) && (large_busy < large_cap || !small_waiting) &&
            s->used + head->bytes <= s->cfg.hbm_budget) {
          s->used += head->bytes;
          s->pending.pop_front();
          sl.reqs.push_back(head);
        } else {
          // A batch of small requests in arrival order, looking past large
          // ones (which keep their places) for at most kServeBypass positions.
          size_t i = 0;
          const bool head_small = is_small(head->r);
          while (i < s->pending.size() && i < kServeBypass && sl.reqs.size() < cap) {
            Request* q = s->pending[i];
            if (!is_small(q->r)) {
              ++i;
              continue;
            }
            if (s->used + q->bytes > s->cfg.hbm_budget) break;
            s->used += q->bytes;
            s->pending.erase(s->pending.begin() + (ptrdiff_t)i);
            sl.reqs.push_back(q);
          }
          (void)head_small;
        }
      }
      if (sl.reqs.empty()) break;  // nothing admissible now
      const double t = now_us(s);
      for (Request* q : sl.reqs) {
        q->state = 1;
        q->t_admit = t;
      }
      sl.busy = true;
      lk.unlock();
      const msv_status st = launch_batch(s, k);
      lk.lock();
      if (st != MSV_OK) {
        // Fail the batch: nothing was recorded on its event.
        const double t2 = now_us(s);
        for (Request* q : sl.reqs) {
          q->state = 3;
          q->status = st;
          q->t_done = t2;
          drop_inputs(s, q);
          drop_outputs(s, q);
          s->used -= q->bytes;
          q->bytes = q->out_bytes = 0;
          --s->outstanding;
        }
        sl.reqs.clear();
        sl.busy = false;
        s->cv_done.notify_all();
      }
      progress = true;
    }
    bool any_busy = false;
    for (const Slot& sl : s->slots) any_busy |= sl.busy;
    if (s->stop && s->pending.empty() && !any_busy) break;
    if (!progress) {
      if (any_busy) {
        // Poll the events without holding the lock (submitters stay
        // unblocked); a timed sleep would add tens of microseconds of latency.
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
      } else {
        s->cv_work.wait(lk);
      }
    }
  }
}

}  // namespace msv

using namespace msv;

extern "C" {

msv_status msv_server_create(const msv_server_config* cfg, msv_server** out) {
  clear_error();
  if (!cfg || !out || cfg->streams == 0 || cfg->streams > kServeMaxStreams ||
      cfg->max_batch == 0 || cfg->mode > MSV_SERVE_PER_JOB || cfg->hbm_budget == 0) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad server configuration");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  auto* s = new msv_server();
  s->cfg = *cfg;
  if (s->cfg.mode == MSV_SERVE_SERIAL) s->cfg.streams = 1;
  if (s->cfg.mode == MSV_SERVE_PER_JOB) s->cfg.max_batch = 1;
  s->t0 = std::chrono::steady_clock::now();
  s->slots.resize(s->cfg.streams);
  for (Slot& sl : s->slots) {
    msv_status st = msv_executor_create(cfg->device, &sl.ex);
    if (st == MSV_OK && hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess)
      st = MSV_ERR_LAUNCH;
    if (st == MSV_OK && hipHostMalloc((void**)&sl.h_counts, sizeof(uint32_t) * s->cfg.max_batch) !=
                            hipSuccess)
      st = MSV_ERR_OUT_OF_MEMORY;
    if (st == MSV_OK && s->cfg.mode == MSV_SERVE_BATCHED) {
      const size_t tb = batch_table_bytes(s->cfg.max_batch);
      if (hipHostMalloc((void**)&sl.h_table, tb) != hipSuccess ||
          hipMalloc((void**)&sl.d_table, tb) != hipSuccess ||
          hipMalloc((void**)&sl.d_counts, sizeof(uint32_t) * s->cfg.max_batch) != hipSuccess) {
        (void)hipGetLastError();
        st = MSV_ERR_OUT_OF_MEMORY;
      }
    }
    if (st != MSV_OK) {
      for (Slot& t : s->slots) {
        if (t.h_table) hipHostFree(t.h_table);
        if (t.d_table) hipFree(t.d_table);
        if (t.d_counts) hipFree(t.d_counts);
        if (t.h_counts) hipHostFree(t.h_counts);
        if (t.ev) hipEventDestroy(t.ev);
        if (t.ex) msv_executor_destroy(t.ex);
      }
      delete s;
      set_error(st, 0, 0, "server resources could not be created");
      return st;
    }
  }
  s->worker = std::thread(scheduler, s);
  *out = s;
  return MSV_OK;
}

msv_status msv_server_destroy(msv_server* s) {
  clear_error();
  if (!s) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null server");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->worker.join();
  for (auto& kv : s->all) drop_outputs(s, kv.second.get());
  for (Slot& sl : s->slots) {
    for (auto& kv : sl.graphs) (void)hipGraphExecDestroy(kv.second);
    if (sl.h_table) hipHostFree(sl.h_table);
    if (sl.d_table) hipFree(sl.d_table);
    if (sl.d_counts) hipFree(sl.d_counts);
    hipHostFree(sl.h_counts);
    hipEventDestroy(sl.ev);
    msv_executor_destroy(sl.ex);
  }
  delete s;
  return MSV_OK;
}

msv_status msv_server_submit(msv_server* s, const msv_request* r, uint64_t* ticket) {
  clear_error();
  if (!s || !r || !ticket) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null argument");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  // Validate the job shape up front (same checks as msv_run_jobs).
  if (r->kind < MSV_JOB_INCLUSIVE_SCAN || r->kind > MSV_JOB_SORT_BY_KEY || !dtype_valid((int)r->dtype) ||
      (r->kind == MSV_JOB_SORT_BY_KEY && ((r->value_bytes != 0 && r->value_bytes != 4) ||
                                          r->op > MSV_ORDER_GREATER)) ||
      (r->kind != MSV_JOB_SORT_BY_KEY && r->kind != MSV_JOB_COPY_WHERE &&
       (r->op > MSV_OP_XOR || (r->op == MSV_OP_XOR && dtype_is_float((msv_dtype)r->dtype)))) ||
      (r->kind == MSV_JOB_COPY_WHERE && r->pred.kind > MSV_PRED_BITS_EQ)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, r->kind, r->dtype, "bad request");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  auto q = std::make_unique<Request>();
  q->r = *r;
  if (r->host_input) {
    const size_t b = (size_t)r->n * dtype_size((msv_dtype)r->dtype);
    q->host.assign((const char*)r->host_input, (const char*)r->host_input + b);
  }
  q->r.host_input = nullptr;
  request_bytes(q->r, &q->bytes, &q->out_bytes);
  if (q->bytes > s->cfg.hbm_budget) {
    set_error(MSV_ERR_OUT_OF_MEMORY, q->bytes, s->cfg.hbm_budget,
              "request needs %llu bytes, budget is %llu", (unsigned long long)q->bytes,
              (unsigned long long)s->cfg.hbm_budget);
    return MSV_ERR_OUT_OF_MEMORY;
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    q->id = s->next_id++;
    q->t_submit = now_us(s);
    *ticket = q->id;
    s->pending.push_back(q.get());
    ++s->outstanding;
    s->all.emplace(q->id, std::move(q));
  }
  s->cv_work.notify_one();
  return MSV_OK;
}

static Request* find(msv_server* s, uint64_t t) {
  auto it = s->all.find(t);
  return it == s->all.end() ? nullptr : it->second.get();
}

msv_status msv_server_wait(msv_server* s, uint64_t ticket, msv_request_info* info) {
  clear_error();
  if (!s) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null server");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  Request* q = find(s, ticket);
  if (!q) {
    set_error(MSV_ERR_INVALID_ARGUMENT, ticket, 0, "unknown ticket %llu", (unsigned long long)ticket);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  s->cv_done.wait(lk, [&] { return q->state >= 2; });
  if (info) {
    info->state = q->state;
    info->status = q->status;
    info->submit_us = q->t_submit;
    info->admit_us = q->t_admit;
    info->done_us = q->t_done;
    info->out_len = q->out_len;
    info->reserved = 0;
    info->bytes = q->bytes;
  }
  return MSV_OK;
}

msv_status msv_server_result(msv_server* s, uint64_t ticket, void* out, uint64_t out_bytes,
                             void* out2, uint64_t out2_bytes) {
  clear_error();
  if (!s) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null server");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  void *src = nullptr, *src2 = nullptr;
  uint64_t need = 0, need2 = 0;
  Request* q = nullptr;
  {
    std::lock_guard<std::mutex> g(s->mu);
    q = find(s, ticket);
    if (!q || q->state != 2) {
      set_error(MSV_ERR_INVALID_ARGUMENT, ticket, 0, "ticket is not a finished request");
      return MSV_ERR_INVALID_ARGUMENT;
    }
    const uint64_t es = dtype_size((msv_dtype)q->r.dtype);
    if (q->r.kind == MSV_JOB_SORT_BY_KEY) {
      src = q->d_out2;
      need = (uint64_t)q->r.n * es;
      src2 = q->d_out;
      need2 = q->r.value_bytes ? (uint64_t)q->r.n * 4 : 0;
    } else {
      src = q->d_out;
      need = (uint64_t)q->out_len * es;
    }
    if ((need && (!out || out_bytes < need)) || (out2 && src2 && out2_bytes < need2)) {
      set_error(MSV_ERR_OUTPUT_TOO_SHORT, need, out_bytes, "output buffer too short");
      return MSV_ERR_OUTPUT_TOO_SHORT;
    }
    // The outputs stay allocated while this copy runs: msv_server_release of
    // the same ticket waits until no reader is left.
    ++q->readers;
  }
  hipError_t e = hipSetDevice(s->cfg.device);
  if (e == hipSuccess && need) e = hipMemcpy(out, src, need, hipMemcpyDeviceToHost);
  if (e == hipSuccess && out2 && src2 && need2) e = hipMemcpy(out2, src2, need2, hipMemcpyDeviceToHost);
  {
    std::lock_guard<std::mutex> g(s->mu);
    --q->readers;
  }
  s->cv_done.notify_all();
  if (e != hipSuccess) {
    set_error(MSV_ERR_LAUNCH, 0, 0, "copying the result: %s", hipGetErrorString(e));
    return MSV_ERR_LAUNCH;
  }
  return MSV_OK;
}

msv_status msv_server_release(msv_server* s, uint64_t ticket) {
  clear_error();
  if (!s) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null server");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  {
    std::unique_lock<std::mutex> lk(s->mu);
    Request* q = find(s, ticket);
    if (!q || q->state < 2) {
      set_error(MSV_ERR_INVALID_ARGUMENT, ticket, 0, "ticket is not a finished request");
      return MSV_ERR_INVALID_ARGUMENT;
    }
    s->cv_done.wait(lk, [&] {
      Request* r = find(s, ticket);
      return !r || r->readers == 0;
    });
    q = find(s, ticket);  // a concurrent release of the same ticket may have won
    if (!q) {
      set_error(MSV_ERR_INVALID_ARGUMENT, ticket, 0, "ticket is not a finished request");
      return MSV_ERR_INVALID_ARGUMENT;
    }
    drop_outputs(s, q);
    s->used -= q->out_bytes;
    s->all.erase(ticket);
  }
  s->cv_work.notify_one();  // budget freed: the head may fit now
  return MSV_OK;
}

msv_status msv_server_drain(msv_server* s) {
  clear_error();
  if (!s) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null server");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
  return MSV_OK;
}

double msv_server_now_us(msv_server* s) { return s ? now_us(s) : 0.0; }

}  // extern "C"
