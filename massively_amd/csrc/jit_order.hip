// User comparators (BinaryPredicateOp), compiled at run time for gfx950
// (hiprtc).
//
// The reference compiles a user BinaryPredicateOp (core/ordering/mod.rs:53-56)
// into its ordering kernels: as `Less` into the block sort and merge passes of
// sort / sort_by_key (core/ordering/mod.rs:325-748; api/algorithm/
// by_key.rs:193-250, ordering.rs:64-76), into merge and merge_by_key
// (core/merge.rs:24-214) and the set operations (core/set.rs:48-207), and into
// seg::Sort (seg/mod.rs:379-425); as `Equal` into the run detection of the
// by-key family (api/algorithm/by_key.rs:359-811) and seg::Unique
// (seg/mod.rs:597-640). Here a comparator is a C++ expression of `a` and `b`
// (nonzero = true) over the key type:
//   MSV_USER_LESS: a comes strictly before b. It becomes the ordering type of
//     the merge-path kernels (merge_kernels.hpp): the comparison sort (block
//     sort of 4096-element tiles, then pairwise merge passes), merge and
//     merge_by_key, and the set operations. Ties keep input order everywhere
//     (the left run / left input first), as in the reference.
//   MSV_USER_EQUAL: a and b belong to one run. One kernel flags the run heads
//     (i == 0 or !equal(k[i-1], k[i])); an inclusive scan of the flags gives
//     each element its run number, and the library's by-key kernels then run
//     on the run numbers (equal run numbers <=> same run), so every by-key
//     algorithm takes a user equality without its own compiled copy.
// seg::Sort with a user Less is, as in the reference, two stable global sorts:
// the comparison sort of (key, position), then the radix sort by segment id
// of the sorted positions; keys and values are gathered once at the end.
#include "jit_internal.hpp"
#include "merge_kernels.hpp"

namespace msv {

__global__ void cmp_fill_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

// seg[i] = the segment of element i: the last s with offsets[s] <= i.
__global__ void cmp_seg_ids_kernel(const uint32_t* __restrict__ offsets, uint32_t nseg,
                                   uint32_t n, uint32_t* __restrict__ seg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = nseg + 1;  // first offset > i, in [0, nseg + 1]
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid + 1;
    else hi = mid;
  }
  seg[i] = lo - 1;
}

// Every non-empty segment's first element is a run head.
__global__ void cmp_mark_segment_heads_kernel(const uint32_t* __restrict__ offsets,
                                              uint32_t nseg, uint32_t n,
                                              uint32_t* __restrict__ flags) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg && offsets[s] < n) flags[offsets[s]] = 1u;
}

// New offsets after keeping the run heads: the number of heads before each
// old offset (run numbers are 1-based inclusive counts of heads).
__global__ void cmp_head_offsets_kernel(const uint32_t* __restrict__ offsets, uint32_t m,
                                        const uint32_t* __restrict__ runs, uint32_t n,
                                        uint32_t* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const uint32_t o = offsets[s];
  out[s] = o < n ? runs[o] - 1u : (n ? runs[n - 1] : 0u);
}

static uint32_t vs_index(size_t vs) { return vs == 0 ? 0u : (vs == 4 ? 1u : 2u); }

msv_status cmp_gen_source(int32_t kind, msv_dtype dtype, const char* expr, std::string& src,
                          std::vector<std::string>& exprs) {
  if ((kind != MSV_USER_LESS && kind != MSV_USER_EQUAL) || !expr || dtype < MSV_U32 ||
      dtype > MSV_F64) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad comparator kind, dtype or expression");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  const char* T = jit_c_type(dtype);
  src = "#include \"device_common.hpp\"\n#include \"merge_kernels.hpp\"\n"
        "#ifndef INFINITY\n#define INFINITY __builtin_inff()\n#endif\n"
        "#ifndef NAN\n#define NAN __builtin_nanf(\"\")\n#endif\n"
        "namespace msv {\nusing UT = ";
  src += T;
  src += ";\n";
  exprs.clear();
  if (kind == MSV_USER_LESS) {
    src += "struct UserLess {\n  __device__ static inline bool less(UT a, UT b) { return (bool)(";
    src += expr;
    src += "); }\n};\n}  // namespace msv\n";
    for (int vs : {0, 4, 8})
      exprs.push_back(jit_fmt("msv::block_sort_kernel<%s, msv::UserLess, %d>", T, vs));
    exprs.push_back(jit_fmt("msv::sort_pass_partition_kernel<%s, msv::UserLess>", T));
    for (int vs : {0, 4, 8})
      exprs.push_back(jit_fmt("msv::sort_pass_kernel<%s, msv::UserLess, %d>", T, vs));
    exprs.push_back(jit_fmt("msv::merge_partition_kernel<%s, msv::UserLess, false>", T));
    exprs.push_back(jit_fmt("msv::merge_partition_kernel<%s, msv::UserLess, true>", T));
    for (int vs : {0, 4, 8})
      exprs.push_back(jit_fmt("msv::merge_tile_kernel<%s, msv::UserLess, %d>", T, vs));
    for (int k : {SET_UNION, SET_INTERSECTION, SET_DIFFERENCE})
      exprs.push_back(jit_fmt("msv::set_tile_kernel<%s, msv::UserLess, %d>", T, k));
  } else {
    src += "struct UserEqual {\n  __device__ static inline bool eq(UT a, UT b) { return (bool)(";
    src += expr;
    src += "); }\n};\n}  // namespace msv\n";
    src += "extern \"C\" __global__ __launch_bounds__(256) void msv_run_heads(\n"
           "    const msv::UT* __restrict__ keys, uint32_t n, uint32_t* __restrict__ flags) {\n"
           "  const uint32_t i = blockIdx.x * 256u + threadIdx.x;\n"
           "  if (i >= n) return;\n"
           "  flags[i] = (i == 0 || !msv::UserEqual::eq(keys[i - 1], keys[i])) ? 1u : 0u;\n"
           "}\n";
  }
  return MSV_OK;
}

msv_status cmp_op_create(msv_executor* ex, int32_t kind, msv_dtype dtype, const char* expr,
                         msv_user_op** out) {
  std::string src;
  std::vector<std::string> exprs;
  MSV_TRY(cmp_gen_source(kind, dtype, expr, src, exprs));
  msv_user_op* op = new msv_user_op;
  op->kind = kind;
  op->dtype = dtype;
  op->owner_id = ex->id;
  op->device = ex->device;
  std::vector<hipFunction_t> fns;
  msv_status s = jit_build(op, src, exprs, fns);
  if (s == MSV_OK && kind == MSV_USER_LESS) {
    size_t at = 0;
    for (int v = 0; v < 3; ++v) op->f_block_sort[v] = fns[at++];
    op->f_pass_part = fns[at++];
    for (int v = 0; v < 3; ++v) op->f_pass[v] = fns[at++];
    op->f_merge_part = fns[at++];
    op->f_set_part = fns[at++];
    for (int v = 0; v < 3; ++v) op->f_merge[v] = fns[at++];
    for (int k = 0; k < 3; ++k) op->f_set[k] = fns[at++];
  } else if (s == MSV_OK) {
    hipError_t e = hipModuleGetFunction(&op->f_heads, op->module, "msv_run_heads");
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "run-head kernel: %s", hipGetErrorString(e));
      s = MSV_ERR_LAUNCH;
    }
  }
  if (s != MSV_OK) {
    if (op->module) (void)hipModuleUnload(op->module);
    delete op;
    return s;
  }
  *out = op;
  return MSV_OK;
}

// ---- drivers ------------------------------------------------------------------
// Comparison sort of n keys (and values of vs bytes) into out_k / out_v.
static msv_status cmp_sort(Ctx& c, const msv_user_op* op, const void* keys, const void* vals,
                           size_t vs, uint32_t n, void* out_k, void* out_v) {
  if (n == 0) return MSV_OK;
  const size_t es = dtype_size(op->dtype);
  const uint32_t vi = vs_index(vs);
  const uint32_t sort_tiles = div_up(n, kSortTile);
  int passes = 0;
  for (uint64_t w = kSortTile; w < n; w *= 2) ++passes;
  // Ping-pong between the output and one scratch pair, starting where the
  // last pass ends in the output.
  void *tk = nullptr, *tv = nullptr;
  if (passes > 0) {
    tk = c.scratch((size_t)n * es);
    tv = vs ? c.scratch((size_t)n * vs) : nullptr;
    if (!tk || (vs && !tv)) return MSV_ERR_OUT_OF_MEMORY;
  }
  void* dk = (passes & 1) ? tk : out_k;
  void* dv = (passes & 1) ? tv : out_v;
  {
    void* a[] = {(void*)&keys, (void*)&vals, (void*)&n, (void*)&dk, (void*)&dv};
    MSV_TRY(jit_launch(op->f_block_sort[vi], sort_tiles, kSortBlock, c.stream(), a));
  }
  if (passes == 0) return MSV_OK;
  uint32_t ntiles = div_up(n, kMergeTile);
  uint32_t* sa = (uint32_t*)c.scratch(sizeof(uint32_t) * ntiles);
  if (!sa) return MSV_ERR_OUT_OF_MEMORY;
  const void* sk = dk;
  const void* sv = dv;
  uint32_t width = kSortTile;
  for (int p = 0; p < passes; ++p) {
    void* ok = (sk == out_k) ? tk : out_k;
    void* ov = (sk == out_k) ? tv : out_v;
    {
      void* a[] = {(void*)&sk, (void*)&n, (void*)&width, (void*)&ntiles, (void*)&sa};
      MSV_TRY(jit_launch(op->f_pass_part, div_up(ntiles, 256), 256, c.stream(), a));
    }
    {
      void* a[] = {(void*)&sk, (void*)&sv, (void*)&n, (void*)&width, (void*)&sa, (void*)&ok,
                   (void*)&ov};
      MSV_TRY(jit_launch(op->f_pass[vi], ntiles, kMergeBlock, c.stream(), a));
    }
    sk = ok;
    sv = ov;
    width *= 2;
  }
  return MSV_OK;
}

static msv_status cmp_partition(Ctx& c, hipFunction_t f, const void* A, uint32_t na,
                                const void* B, uint32_t nb, uint32_t ntiles, uint32_t tile,
                                uint32_t** sa, uint32_t** sb) {
  *sa = (uint32_t*)c.scratch(sizeof(uint32_t) * (ntiles + 1));
  *sb = (uint32_t*)c.scratch(sizeof(uint32_t) * (ntiles + 1));
  if (!*sa || !*sb) return MSV_ERR_OUT_OF_MEMORY;
  void* a[] = {(void*)&A, (void*)&na, (void*)&B, (void*)&nb, (void*)&ntiles, (void*)&tile,
               (void*)sa, (void*)sb};
  return jit_launch(f, div_up((uint64_t)ntiles + 1, 256), 256, c.stream(), a);
}

static msv_status cmp_merge(Ctx& c, const msv_user_op* op, const void* A, uint32_t na,
                            const void* B, uint32_t nb, const void* Av, const void* Bv, size_t vs,
                            void* out_k, void* out_v) {
  const uint64_t total = (uint64_t)na + nb;
  if (total == 0) return MSV_OK;
  uint32_t ntiles = (uint32_t)((total + kMergeTile - 1) / kMergeTile);
  uint32_t *sa, *sb;
  MSV_TRY(cmp_partition(c, op->f_merge_part, A, na, B, nb, ntiles, kMergeTile, &sa, &sb));
  void* a[] = {(void*)&A,  (void*)&na, (void*)&B,  (void*)&nb,    (void*)&Av,
               (void*)&Bv, (void*)&sa, (void*)&sb, (void*)&out_k, (void*)&out_v};
  return jit_launch(op->f_merge[vs_index(vs)], ntiles, kMergeBlock, c.stream(), a);
}

static msv_status cmp_set(Ctx& c, const msv_user_op* op, int kind, const void* A, uint32_t na,
                          const void* B, uint32_t nb, void* out, uint32_t* out_len) {
  const uint64_t total = (uint64_t)na + nb;
  if (total == 0) {
    hipLaunchKernelGGL(cmp_fill_u32_kernel, dim3(1), dim3(1), 0, c.stream(), out_len, 0u);
    MSV_HIP(hipGetLastError());
    return MSV_OK;
  }
  uint32_t ntiles = (uint32_t)((total + kSetSpan - 1) / kSetSpan);
  uint32_t *sa, *sb;
  MSV_TRY(cmp_partition(c, op->f_set_part, A, na, B, nb, ntiles, kSetSpan, &sa, &sb));
  uint64_t *counter, *status;
  uint32_t gen;
  MSV_TRY(c.lookback(ntiles, &counter, &status, &gen));
  void* a[] = {(void*)&A,   (void*)&na,      (void*)&B,       (void*)&nb,
               (void*)&sa,  (void*)&sb,      (void*)&out,     (void*)&out_len,
               (void*)&counter, (void*)&status, (void*)&gen, (void*)&ntiles};
  return jit_launch(op->f_set[kind], ntiles, kSetBlock, c.stream(), a);
}

// Run numbers of a key column under a user equality: runs[i] = 1 + the
// number of run heads in (0, i]. With offsets, every non-empty segment's
// first element is a head too. `flags` (n words) is left holding the heads.
static msv_status eq_runs(Ctx& c, const msv_user_op* op, const void* keys, uint32_t n,
                          const uint32_t* offsets, uint32_t nseg, uint32_t* flags,
                          uint32_t* runs) {
  void* a[] = {(void*)&keys, (void*)&n, (void*)&flags};
  MSV_TRY(jit_launch(op->f_heads, div_up(n, 256), 256, c.stream(), a));
  if (offsets && nseg) {
    hipLaunchKernelGGL(cmp_mark_segment_heads_kernel, dim3(div_up(nseg, 256)), dim3(256), 0,
                       c.stream(), offsets, nseg, n, flags);
    MSV_HIP(hipGetLastError());
  }
  return scan_run(c, MSV_U32, MSV_OP_ADD, false, flags, runs, n, nullptr);
}

}  // namespace msv

using namespace msv;

static msv_status same_len(const View& a, const View& b) {
  if (a.len != b.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, a.len, b.len, "input lengths differ: %u != %u", a.len,
              b.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  return MSV_OK;
}

static msv_status total_fits(uint32_t a, uint32_t b) {
  if ((uint64_t)a + b > 0xffffffffull) {
    set_error(MSV_ERR_LENGTH_TOO_LARGE, (uint64_t)a + b, 0, "length does not fit in u32: %llu",
              (unsigned long long)((uint64_t)a + b));
    return MSV_ERR_LENGTH_TOO_LARGE;
  }
  return MSV_OK;
}

static void release_all(msv_vec** a, msv_vec** b = nullptr) {
  if (a && *a) {
    msv_vec_release(*a);
    *a = nullptr;
  }
  if (b && *b) {
    msv_vec_release(*b);
    *b = nullptr;
  }
}

// Stable sort of keys (and values) by a user Less; outputs owned by the caller.
static msv_status sort_user_api(msv_executor* ex, msv_slice keys, const msv_slice* values,
                                const msv_user_op* less, msv_vec** out_values,
                                msv_vec** out_keys) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  if (values) {
    MSV_TRY(resolve(ex, *values, &v));
    MSV_TRY(same_len(k, v));
  }
  MSV_TRY(jit_check_op(ex, less, MSV_USER_LESS, k.dtype, false));
  if ((values && !out_values) || (!values && !out_keys)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  void* ok = nullptr;
  msv_status s = MSV_OK;
  if (out_keys) {
    s = new_vec(ex, k.dtype, k.len, false, out_keys);
    if (s == MSV_OK) ok = (*out_keys)->ptr;
  } else {
    ok = c.scratch((size_t)k.len * dtype_size(k.dtype) + 1);
    if (!ok) s = MSV_ERR_OUT_OF_MEMORY;
  }
  if (s == MSV_OK && values) s = new_vec(ex, v.dtype, v.len, false, out_values);
  if (s == MSV_OK)
    s = cmp_sort(c, less, k.ptr, values ? v.ptr : nullptr, values ? dtype_size(v.dtype) : 0,
                 k.len, ok, values ? (*out_values)->ptr : nullptr);
  if (s != MSV_OK) release_all(out_keys, values ? out_values : nullptr);
  return s;
}

// The run numbers of `keys` under a user Equal, in c's scratch.
static msv_status runs_of(Ctx& c, msv_executor* ex, const View& k, const msv_user_op* equal,
                          uint32_t** runs, uint32_t** flags) {
  MSV_TRY(jit_check_op(ex, equal, MSV_USER_EQUAL, k.dtype, false));
  *flags = (uint32_t*)c.scratch((size_t)k.len * 4 + 4);
  *runs = (uint32_t*)c.scratch((size_t)k.len * 4 + 4);
  if (!*flags || !*runs) return MSV_ERR_OUT_OF_MEMORY;
  if (k.len == 0) return MSV_OK;
  return eq_runs(c, equal, k.ptr, k.len, nullptr, 0, *flags, *runs);
}

static bool op_ok(msv_dtype dt, int op) {
  if (op < MSV_OP_ADD || op > MSV_OP_XOR) return false;
  return !(op == MSV_OP_XOR && dtype_is_float(dt));
}

static msv_status scan_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                   const msv_user_op* equal, const void* init, msv_op op,
                                   bool exclusive, msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  MSV_TRY(resolve(ex, values, &v));
  MSV_TRY(same_len(k, v));
  if (!op_ok(v.dtype, op) || !out || (exclusive && !init)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad operator, init or output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  uint32_t *runs, *flags;
  MSV_TRY(runs_of(c, ex, k, equal, &runs, &flags));
  MSV_TRY(new_vec(ex, v.dtype, v.len, false, out));
  msv_status s = MSV_OK;
  if (k.len)
    s = runs_scan_run(c, MSV_U32, runs, v.dtype, op, exclusive ? RUNS_EXCLUSIVE : RUNS_INCLUSIVE,
                      v.ptr, v.len, (*out)->ptr, init, nullptr, nullptr, nullptr);
  if (s != MSV_OK) release_all(out);
  return s;
}

extern "C" {

msv_status msv_sort_user(msv_executor* ex, msv_slice input, const msv_user_op* less,
                         msv_vec** out) {
  return sort_user_api(ex, input, nullptr, less, nullptr, out);
}

msv_status msv_sort_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                const msv_user_op* less, msv_vec** out_values,
                                msv_vec** out_keys) {
  return sort_user_api(ex, keys, &values, less, out_values, out_keys);
}

msv_status msv_merge_user(msv_executor* ex, msv_slice left, msv_slice right,
                          const msv_user_op* less, msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View a, b;
  MSV_TRY(resolve(ex, left, &a));
  MSV_TRY(resolve(ex, right, &b));
  if (a.dtype != b.dtype || !out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "left and right dtypes differ or null output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(jit_check_op(ex, less, MSV_USER_LESS, a.dtype, false));
  MSV_TRY(total_fits(a.len, b.len));
  MSV_TRY(new_vec(ex, a.dtype, a.len + b.len, false, out));
  Ctx c(ex);
  msv_status s = cmp_merge(c, less, a.ptr, a.len, b.ptr, b.len, nullptr, nullptr, 0, (*out)->ptr,
                           nullptr);
  if (s != MSV_OK) release_all(out);
  return s;
}

msv_status msv_merge_by_key_user(msv_executor* ex, msv_slice left_keys, msv_slice left_values,
                                 msv_slice right_keys, msv_slice right_values,
                                 const msv_user_op* less, msv_vec** out_keys,
                                 msv_vec** out_values) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View lk, lv, rk, rv;
  MSV_TRY(resolve(ex, left_keys, &lk));
  MSV_TRY(resolve(ex, left_values, &lv));
  MSV_TRY(resolve(ex, right_keys, &rk));
  MSV_TRY(resolve(ex, right_values, &rv));
  MSV_TRY(same_len(lk, lv));
  MSV_TRY(same_len(rk, rv));
  if (lk.dtype != rk.dtype || lv.dtype != rv.dtype || !out_keys || !out_values) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "dtype mismatch or null output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(jit_check_op(ex, less, MSV_USER_LESS, lk.dtype, false));
  MSV_TRY(total_fits(lk.len, rk.len));
  MSV_TRY(new_vec(ex, lk.dtype, lk.len + rk.len, false, out_keys));
  msv_status s = new_vec(ex, lv.dtype, lv.len + rv.len, false, out_values);
  if (s == MSV_OK) {
    Ctx c(ex);
    s = cmp_merge(c, less, lk.ptr, lk.len, rk.ptr, rk.len, lv.ptr, rv.ptr, dtype_size(lv.dtype),
                  (*out_keys)->ptr, (*out_values)->ptr);
  }
  if (s != MSV_OK) release_all(out_keys, out_values);
  return s;
}

msv_status msv_set_op_user(msv_executor* ex, msv_set_kind kind, msv_slice left, msv_slice right,
                           const msv_user_op* less, msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  if (kind < MSV_SET_UNION || kind > MSV_SET_DIFFERENCE || !out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "unknown set operation %d or null output",
              (int)kind);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  View a, b;
  MSV_TRY(resolve(ex, left, &a));
  MSV_TRY(resolve(ex, right, &b));
  if (a.dtype != b.dtype) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "left and right dtypes differ");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(jit_check_op(ex, less, MSV_USER_LESS, a.dtype, false));
  MSV_TRY(total_fits(a.len, b.len));
  const uint32_t cap = kind == MSV_SET_UNION ? a.len + b.len : a.len;
  MSV_TRY(new_vec(ex, a.dtype, cap, true, out));
  Ctx c(ex);
  msv_status s = cmp_set(c, less, (int)kind, a.ptr, a.len, b.ptr, b.len, (*out)->ptr,
                         (*out)->d_len);
  if (s != MSV_OK) release_all(out);
  return s;
}

msv_status msv_seg_sort_user(msv_executor* ex, msv_slice keys, msv_slice values,
                             const msv_segmentation* seg, const msv_user_op* less,
                             msv_vec** out_keys, msv_vec** out_values) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  const bool has_v = values.vec != nullptr;
  if (has_v) {
    MSV_TRY(resolve(ex, values, &v));
    MSV_TRY(same_len(k, v));
  }
  if (!seg || !seg->offsets || !out_keys || (has_v && !out_values)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null segmentation or output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (seg->offsets->owner_id != ex->id) {
    set_error(MSV_ERR_FOREIGN_EXECUTOR, 0, 0, "executor does not own this segmentation");
    return MSV_ERR_FOREIGN_EXECUTOR;
  }
  if (seg->value_count != k.len) {
    set_error(MSV_ERR_LENGTH_MISMATCH, seg->value_count, k.len, "input lengths differ: %u != %u",
              seg->value_count, k.len);
    return MSV_ERR_LENGTH_MISMATCH;
  }
  MSV_TRY(jit_check_op(ex, less, MSV_USER_LESS, k.dtype, false));
  const uint32_t n = k.len, nseg = seg->segment_count;
  MSV_TRY(new_vec(ex, k.dtype, n, false, out_keys));
  msv_status s = has_v ? new_vec(ex, v.dtype, n, false, out_values) : MSV_OK;
  Ctx c(ex);
  // 1) positions sorted by key; 2) their segment ids; 3) positions stably
  // sorted by segment id; 4) keys and values gathered through them.
  uint32_t *pos = nullptr, *pos1 = nullptr, *ids = nullptr, *ids1 = nullptr, *pos2 = nullptr;
  void* k1 = nullptr;
  if (s == MSV_OK && n > 0) {
    pos = (uint32_t*)c.scratch((size_t)n * 4);
    pos1 = (uint32_t*)c.scratch((size_t)n * 4);
    ids = (uint32_t*)c.scratch((size_t)n * 4);
    ids1 = (uint32_t*)c.scratch((size_t)n * 4);
    pos2 = (uint32_t*)c.scratch((size_t)n * 4);
    k1 = c.scratch((size_t)n * dtype_size(k.dtype));
    if (!pos || !pos1 || !ids || !ids1 || !pos2 || !k1) s = MSV_ERR_OUT_OF_MEMORY;
    if (s == MSV_OK) s = iota_run(c, pos, n, 0);
    if (s == MSV_OK) s = cmp_sort(c, less, k.ptr, pos, 4, n, k1, pos1);
    if (s == MSV_OK) {
      hipLaunchKernelGGL(cmp_seg_ids_kernel, dim3(div_up(n, 256)), dim3(256), 0, c.stream(),
                         (const uint32_t*)seg->offsets->ptr, nseg, n, ids);
      s = hipGetLastError() == hipSuccess ? MSV_OK : MSV_ERR_LAUNCH;
    }
    if (s == MSV_OK) s = gather_run(c, MSV_U32, ids, n, pos1, ids1, n);
    // ids is free again: it receives the (unused) sorted segment ids.
    if (s == MSV_OK) s = radix_sort_run(c, MSV_U32, ORD_TOTAL, 4, ids1, pos1, n, ids, pos2);
    if (s == MSV_OK) s = gather_run(c, k.dtype, k.ptr, n, pos2, (*out_keys)->ptr, n);
    if (s == MSV_OK && has_v) s = gather_run(c, v.dtype, v.ptr, n, pos2, (*out_values)->ptr, n);
  }
  if (s != MSV_OK) release_all(out_keys, has_v ? out_values : nullptr);
  return s;
}

msv_status msv_inclusive_scan_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                          const msv_user_op* equal, msv_op op, msv_vec** out) {
  return scan_by_key_user(ex, keys, values, equal, nullptr, op, false, out);
}

msv_status msv_exclusive_scan_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                          const msv_user_op* equal, const void* init,
                                          msv_op op, msv_vec** out) {
  return scan_by_key_user(ex, keys, values, equal, init, op, true, out);
}

msv_status msv_reduce_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                  const msv_user_op* equal, const void* init, msv_op op,
                                  msv_vec** out_keys, msv_vec** out_values) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  MSV_TRY(resolve(ex, values, &v));
  MSV_TRY(same_len(k, v));
  if (!op_ok(v.dtype, op) || !init || !out_keys || !out_values) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad operator, init or output handles");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  uint32_t *runs, *flags;
  MSV_TRY(runs_of(c, ex, k, equal, &runs, &flags));
  MSV_TRY(new_vec(ex, k.dtype, k.len, true, out_keys));
  msv_status s = new_vec(ex, v.dtype, v.len, true, out_values);
  const uint32_t n = k.len;
  if (s == MSV_OK && n == 0) {
    hipLaunchKernelGGL(cmp_fill_u32_kernel, dim3(1), dim3(1), 0, c.stream(), (*out_keys)->d_len, 0u);
    hipLaunchKernelGGL(cmp_fill_u32_kernel, dim3(1), dim3(1), 0, c.stream(), (*out_values)->d_len,
                       0u);
    s = hipGetLastError() == hipSuccess ? MSV_OK : MSV_ERR_LAUNCH;
  } else if (s == MSV_OK) {
    // The by-key reduction over run numbers; its run keys (run numbers) go
    // to scratch, and each run's first key is taken from the keys.
    uint32_t* run_keys = (uint32_t*)c.scratch((size_t)n * 4);
    uint32_t* run_len = (uint32_t*)c.scratch(4);
    if (!run_keys || !run_len) s = MSV_ERR_OUT_OF_MEMORY;
    if (s == MSV_OK)
      s = runs_scan_run(c, MSV_U32, runs, v.dtype, op, RUNS_REDUCE, v.ptr, n, (*out_values)->ptr,
                        init, run_keys, run_len, (*out_values)->d_len);
    if (s == MSV_OK)
      s = unique_runs_run(c, MSV_U32, runs, dtype_size(k.dtype), k.ptr, n, (*out_keys)->ptr,
                          (*out_keys)->d_len);
  }
  if (s != MSV_OK) release_all(out_keys, out_values);
  return s;
}

msv_status msv_unique_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                  const msv_user_op* equal, msv_vec** out_values) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View k, v;
  MSV_TRY(resolve(ex, keys, &k));
  MSV_TRY(resolve(ex, values, &v));
  MSV_TRY(same_len(k, v));
  if (!out_values) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output handle");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  uint32_t *runs, *flags;
  MSV_TRY(runs_of(c, ex, k, equal, &runs, &flags));
  MSV_TRY(new_vec(ex, v.dtype, v.len, true, out_values));
  msv_status s;
  if (k.len == 0) {
    hipLaunchKernelGGL(cmp_fill_u32_kernel, dim3(1), dim3(1), 0, c.stream(), (*out_values)->d_len,
                       0u);
    s = hipGetLastError() == hipSuccess ? MSV_OK : MSV_ERR_LAUNCH;
  } else {
    s = unique_runs_run(c, MSV_U32, runs, dtype_size(v.dtype), v.ptr, k.len, (*out_values)->ptr,
                        (*out_values)->d_len);
  }
  if (s != MSV_OK) release_all(out_values);
  return s;
}

msv_status msv_seg_unique_user(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                               const msv_user_op* equal, msv_vec** out_values,
                               msv_segmentation** out_seg) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
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
  MSV_TRY(jit_check_op(ex, equal, MSV_USER_EQUAL, v.dtype, false));
  const uint32_t n = v.len, nseg = seg->segment_count;
  msv_vec *vals = nullptr, *offs = nullptr;
  MSV_TRY(new_vec(ex, v.dtype, n, false, &vals));
  msv_status s = new_vec(ex, MSV_U32, nseg + 1, false, &offs);
  Ctx c(ex);
  uint32_t* flags = s == MSV_OK ? (uint32_t*)c.scratch((size_t)n * 4 + 4) : nullptr;
  uint32_t* runs = s == MSV_OK ? (uint32_t*)c.scratch((size_t)n * 4 + 4) : nullptr;
  uint32_t* d_count = s == MSV_OK ? (uint32_t*)c.scratch(4) : nullptr;
  if (s == MSV_OK && (!flags || !runs || !d_count)) s = MSV_ERR_OUT_OF_MEMORY;
  const uint32_t* offsets = (const uint32_t*)seg->offsets->ptr;
  if (s == MSV_OK && n > 0) {
    s = eq_runs(c, equal, v.ptr, n, offsets, nseg, flags, runs);
    if (s == MSV_OK)
      s = unique_runs_run(c, MSV_U32, runs, dtype_size(v.dtype), v.ptr, n, vals->ptr, d_count);
  } else if (s == MSV_OK) {
    hipLaunchKernelGGL(cmp_fill_u32_kernel, dim3(1), dim3(1), 0, c.stream(), d_count, 0u);
    s = hipGetLastError() == hipSuccess ? MSV_OK : MSV_ERR_LAUNCH;
  }
  if (s == MSV_OK) {
    hipLaunchKernelGGL(cmp_head_offsets_kernel, dim3(div_up((uint64_t)nseg + 1, 256)), dim3(256), 0,
                       c.stream(), offsets, nseg + 1, runs, n, (uint32_t*)offs->ptr);
    s = hipGetLastError() == hipSuccess ? MSV_OK : MSV_ERR_LAUNCH;
  }
  uint32_t kept = 0;
  if (s == MSV_OK) {
    hipError_t e = hipMemcpyAsync(&kept, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, ex->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ex->stream);
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "seg unique failed: %s", hipGetErrorString(e));
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

}  // extern "C"
