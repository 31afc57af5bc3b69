// User-defined operators and lazy inputs, compiled at run time for gfx950
// (hiprtc).
//
// The reference's algorithms take user operators as CubeCL traits compiled
// into its kernels (ReductionOp::apply, core/reduce.rs:63-66; PredicateOp::apply,
// core/predicate.rs:40-43), and lazy inputs (api/lazy.rs: counting, constant,
// stride, map, permute, reverse, repeat_each, tile, adjacent_difference) as
// read expressions over up to 13 device slots that are compiled into the
// consuming kernel (core/eval.rs:21-107, core/read.rs). Here:
//   reduction:  an expression of `a` and `b` over the element type (associative;
//               commutativity is not needed: every tree and scan below
//               combines in index order), plus an identity expression; or a
//               built-in operator code (add / min / max / xor);
//   predicate:  an expression of `x` (nonzero = keep), or the built-in
//               predicates (msv_predicate, given per call);
//   read source (optional): C++ defining
//                 UT msv_read(const LazyArgs& la, size_t i)
//               the element i of a lazy input, reading device slots with
//               MSV_SLOT(k, type, index) (0 past the slot's end) and run-time
//               parameters la.q[0..15]. A source turns the operator's kernels
//               into kernels that read through it: the lazy input is computed
//               inside the consuming kernel and never stored.
// msv_lazy_op_create compiles, once, the library's own kernel templates
// instantiated with these: the same device headers the library is built from
// are embedded in it (jit_headers.inc), a user reduction becomes a
// specialisation Combine<T, kOpUser>, a user predicate a functor UserPred, a
// read source the functor LazySrc that the kernels take in place of their input
// pointer, and hiprtc's name expressions give the instantiations' symbols in
// the loaded module. Reduce and the scans' tile partials use an
// order-preserving tile reduction (rows in index order); the scans use the
// fixed-shape (prefixed) passes of the float scans, which keep the order of
// combination; copy_where uses the one-pass compaction with the predicate on
// the input itself; map writes the lazy input out.
#include "jit_internal.hpp"
#include "scan_kernels.hpp"
#include "select_kernels.hpp"
#include "jit_headers.inc"

namespace msv {
constexpr int kOpUser = 100;  // Combine<T, kOpUser> is the user's operator
constexpr int kMapBlock = 256;
// Rows of the long tiles of the one-pass scan over a lazy input: half of the
// built-in scan's (a lazy tile's per-element fallback for the last tile holds
// more registers; 16 rows of u32 take 134 VGPRs, 32 rows 265).
inline int lazy_scan_rows_large(size_t es) {
  return (es == 4 ? scan_rows_large<uint32_t>() : scan_rows_large<uint64_t>()) / 2;
}

// Host image of the device LazyArgs (same layout: kernels take it by value).
struct LazyArgsHost {
  const void* p[MSV_LAZY_SLOTS];
  uint32_t n[MSV_LAZY_SLOTS];
  uint64_t q[MSV_LAZY_PARAMS];
};
static_assert(sizeof(LazyArgsHost) == 288, "LazyArgs layout");
}  // namespace msv

using namespace msv;

const char* msv::jit_c_type(msv_dtype dt) {
  switch (dt) {
    case MSV_U32: return "uint32_t";
    case MSV_I32: return "int32_t";
    case MSV_F32: return "float";
    case MSV_U64: return "uint64_t";
    case MSV_I64: return "int64_t";
    default: return "double";
  }
}

std::string msv::jit_fmt(const char* f, ...) {
  va_list ap, ap2;
  va_start(ap, f);
  va_copy(ap2, ap);
  const int n = vsnprintf(nullptr, 0, f, ap);
  va_end(ap);
  std::string out(n > 0 ? (size_t)n : 0, '\0');
  if (n > 0) vsnprintf(&out[0], (size_t)n + 1, f, ap2);
  va_end(ap2);
  return out;
}

msv_status msv::jit_compile(const std::string& src, const std::vector<std::string>& exprs,
                            std::vector<char>& code, std::vector<std::string>& lowered) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "msv_user_op.hip", kJitHeaderCount,
                          kJitHeaderBodies, kJitHeaderNames) != HIPRTC_SUCCESS) {
    set_error(MSV_ERR_LAUNCH, 0, 0, "hiprtcCreateProgram failed");
    return MSV_ERR_LAUNCH;
  }
  for (const std::string& e : exprs) hiprtcAddNameExpression(prog, e.c_str());
  const char* opts[] = {"--offload-arch=gfx950", "-O3",     "-std=c++17", "-Xclang",
                        "-target-feature",       "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    // The first diagnostic is what the caller needs (the log names the line).
    const size_t at = log.find("error:");
    const std::string first = log.substr(at == std::string::npos ? 0 : at, 200);
    hiprtcDestroyProgram(&prog);
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "user operator does not compile: %s", first.c_str());
    return MSV_ERR_INVALID_ARGUMENT;
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  code.resize(code_size);
  hiprtcGetCode(prog, code.data());
  lowered.clear();
  for (const std::string& e : exprs) {
    const char* name = nullptr;
    hiprtcGetLoweredName(prog, e.c_str(), &name);
    lowered.push_back(name ? name : "");
  }
  hiprtcDestroyProgram(&prog);
  for (const std::string& l : lowered) {
    if (l.empty()) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "a kernel instantiation has no symbol");
      return MSV_ERR_LAUNCH;
    }
  }
  return MSV_OK;
}

msv_status msv::jit_build(msv_user_op* op, const std::string& src,
                          const std::vector<std::string>& exprs, std::vector<hipFunction_t>& fns) {
  std::vector<char> code;
  std::vector<std::string> lowered;
  MSV_TRY(jit_compile(src, exprs, code, lowered));
  MSV_HIP(hipSetDevice(op->device));
  MSV_HIP(hipModuleLoadData(&op->module, code.data()));
  fns.clear();
  for (const std::string& l : lowered) {
    hipFunction_t f = nullptr;
    MSV_HIP(hipModuleGetFunction(&f, op->module, l.c_str()));
    fns.push_back(f);
  }
  return MSV_OK;
}

msv_status msv::jit_launch(hipFunction_t f, uint32_t grid, uint32_t block, hipStream_t s,
                           void** args) {
  MSV_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, s, args, nullptr));
  return MSV_OK;
}

// A value of the op's element type, passed by address as a kernel argument.
union Scalar {
  uint32_t u4;
  uint64_t u8;
  unsigned char bytes[8];
};

static Scalar scalar_of(const void* p, size_t size) {
  Scalar v{};
  memcpy(v.bytes, p, size);
  return v;
}

msv_status msv::jit_check_op(msv_executor* ex, const msv_user_op* op, int32_t kind,
                             msv_dtype dt, bool lazy) {
  if (!op || op->kind != kind) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null user operator or wrong kind");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (op->owner_id != ex->id) {
    set_error(MSV_ERR_FOREIGN_EXECUTOR, 0, 0, "executor does not own this user operator");
    return MSV_ERR_FOREIGN_EXECUTOR;
  }
  if (op->dtype != dt) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "user operator was compiled for another dtype");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  if (lazy && !op->lazy) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator was compiled without a read source");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  // A predicate's compaction kernels read either a pointer or the read source.
  if (kind == MSV_USER_PREDICATE && !lazy && op->lazy) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "predicate was compiled for a lazy input");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return MSV_OK;
}

// Resolves the slots of a lazy input (each an owned slice of this executor)
// into device pointers and lengths.
static msv_status resolve_lazy(msv_executor* ex, const msv_lazy_args* a, LazyArgsHost* h) {
  if (!a || a->nslots > MSV_LAZY_SLOTS) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null lazy input or more than %d slots",
              MSV_LAZY_SLOTS);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  memset(h, 0, sizeof(*h));
  for (uint32_t k = 0; k < a->nslots; ++k) {
    View v;
    MSV_TRY(resolve(ex, a->slots[k], &v));
    h->p[k] = v.ptr;
    h->n[k] = v.len;
  }
  memcpy(h->q, a->params, sizeof(h->q));
  return MSV_OK;
}

// One input of the fixed-shape passes: a device pointer or the lazy source.
struct Input {
  const void* ptr = nullptr;
  const LazyArgsHost* lz = nullptr;
  void* arg() const { return lz ? (void*)lz : (void*)&ptr; }
};

// Fixed-shape scan of n elements (the float scans' passes): tile partials, an
// exclusive scan of them (recursively, over a pointer), then the prefixed tile
// scan.
static msv_status user_scan(Ctx& c, const msv_user_op* op, Input in, void* out, uint32_t n,
                            bool exclusive, Scalar init) {
  const size_t es = dtype_size(op->dtype);
  const uint32_t tile = (uint32_t)kScanBlock * kScanRowsSmall * (uint32_t)(16 / es);
  const uint32_t ntiles = div_up(n, tile);
  bool aligned = (((in.lz ? 0 : (uintptr_t)in.ptr) | (uintptr_t)out) & 15u) == 0;
  const void* prefixes = nullptr;
  Scalar ident = scalar_of(&op->identity, es);
  if (ntiles > 1) {
    void* partials = c.scratch((size_t)ntiles * es);
    void* pre = c.scratch((size_t)ntiles * es);
    if (!partials || !pre) return MSV_ERR_OUT_OF_MEMORY;
    bool no = false;
    void* a[] = {in.arg(), (void*)&n, (void*)&aligned, (void*)&partials, (void*)&no, (void*)&ident};
    MSV_TRY(jit_launch(in.lz ? op->f_partials_lz : op->f_partials, ntiles, kScanBlock, c.stream(), a));
    Input p;
    p.ptr = partials;
    MSV_TRY(user_scan(c, op, p, pre, ntiles, true, ident));
    prefixes = pre;
  }
  uint64_t* status = nullptr;
  uint32_t gen = 0;
  void* a[] = {in.arg(),        (void*)&out,    (void*)&n,   (void*)&init,
               (void*)&aligned, (void*)&status, (void*)&gen, (void*)&prefixes};
  hipFunction_t f = exclusive ? (in.lz ? op->f_scan_ex_lz : op->f_scan_ex)
                              : (in.lz ? op->f_scan_in_lz : op->f_scan_in);
  return jit_launch(f, ntiles, kScanBlock, c.stream(), a);
}

// Scan of a slice (`input`) or of a lazy input (`lz`).
static msv_status user_scan_api(msv_executor* ex, const msv_slice* input, const msv_lazy_args* lz,
                                const void* init, const msv_user_op* op, bool exclusive,
                                msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View in;
  LazyArgsHost h;
  Input src;
  if (lz) {
    if (!op) return jit_check_op(ex, op, MSV_USER_REDUCTION, MSV_U32, true);
    MSV_TRY(resolve_lazy(ex, lz, &h));
    in.len = lz->len;
    in.dtype = op->dtype;
    src.lz = &h;
  } else {
    MSV_TRY(resolve(ex, *input, &in));
    src.ptr = in.ptr;
  }
  MSV_TRY(jit_check_op(ex, op, MSV_USER_REDUCTION, in.dtype, lz != nullptr));
  if (!out || (exclusive && !init)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output or init");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, in.dtype, in.len, false, out));
  if (in.len == 0) return MSV_OK;
  Ctx c(ex);
  const size_t es = dtype_size(in.dtype);
  Scalar iv = exclusive ? scalar_of(init, es) : scalar_of(&op->identity, es);
  msv_status s;
  if (lz && op->f_lb[exclusive][0]) {
    // One pass with decoupled look-back (integer results do not depend on
    // the order of combination), as the built-in scans of device vectors.
    const bool large = in.len >= kScanLargeMin;
    const int rows = large ? lazy_scan_rows_large(es) : kScanRowsSmall;
    const uint32_t tile = (uint32_t)kScanBlock * (uint32_t)rows * (uint32_t)(16 / es);
    uint32_t n = in.len, ntiles = div_up(n, tile);
    uint64_t *counter, *status;
    uint32_t gen;
    s = c.lookback((size_t)ntiles * (es == 4 ? 1 : 2), &counter, &status, &gen);
    if (s == MSV_OK) {
      void* o = (*out)->ptr;
      bool aligned = ((uintptr_t)o & 15u) == 0;
      const void* prefixes = nullptr;
      void* a[] = {(void*)&h,        (void*)&o,      (void*)&n,   (void*)&iv,
                   (void*)&aligned, (void*)&status, (void*)&gen, (void*)&prefixes};
      s = jit_launch(op->f_lb[exclusive][large], ntiles, kScanBlock, c.stream(), a);
    }
  } else {
    s = user_scan(c, op, src, (*out)->ptr, in.len, exclusive, iv);
  }
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

static msv_status user_reduce_api(msv_executor* ex, const msv_slice* input,
                                  const msv_lazy_args* lz, const void* init,
                                  const msv_user_op* op, void* out_host) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View in;
  LazyArgsHost h;
  if (lz) {
    if (!op) return jit_check_op(ex, op, MSV_USER_REDUCTION, MSV_U32, true);
    MSV_TRY(resolve_lazy(ex, lz, &h));
    in.len = lz->len;
    in.dtype = op->dtype;
  } else {
    MSV_TRY(resolve(ex, *input, &in));
  }
  MSV_TRY(jit_check_op(ex, op, MSV_USER_REDUCTION, in.dtype, lz != nullptr));
  if (!init || !out_host) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null init or output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  const size_t es = dtype_size(in.dtype);
  const uint32_t tile = (uint32_t)kReduceBlock * kReduceRows * (uint32_t)(16 / es);
  void* d_result = c.scratch(es);
  if (!d_result) return MSV_ERR_OUT_OF_MEMORY;
  // Levels of the fixed tree; the last one applies op(init, total). The first
  // level reads the input (or the lazy source), the others the partials.
  Input src;
  if (lz) src.lz = &h;
  else src.ptr = in.ptr;
  uint32_t m = in.len;
  bool aligned = lz || ((uintptr_t)in.ptr & 15u) == 0;
  Scalar iv = scalar_of(init, es);
  for (;;) {
    const uint32_t tiles = m == 0 ? 1 : div_up(m, tile);
    bool last = tiles == 1;
    void* dst = last ? d_result : c.scratch((size_t)tiles * es);
    if (!dst) return MSV_ERR_OUT_OF_MEMORY;
    void* a[] = {src.arg(), (void*)&m, (void*)&aligned, (void*)&dst, (void*)&last, (void*)&iv};
    MSV_TRY(jit_launch(src.lz ? op->f_reduce_lz : op->f_reduce, tiles, kReduceBlock, c.stream(), a));
    if (last) break;
    src = Input();
    src.ptr = dst;
    m = tiles;
    aligned = true;
  }
  MSV_HIP(hipMemcpyAsync(out_host, d_result, es, hipMemcpyDeviceToHost, c.stream()));
  MSV_HIP(hipStreamSynchronize(c.stream()));
  return MSV_OK;
}

// The built-in predicate's kernel argument (Pred<S>) for the op's dtype.
struct PredArg {
  alignas(8) unsigned char bytes[32];
};

static PredArg pred_arg(msv_dtype dt, const msv_predicate* p) {
  PredArg a{};
  switch (dt) {
    case MSV_U32: { auto v = make_pred<uint32_t>(p); memcpy(a.bytes, &v, sizeof(v)); break; }
    case MSV_I32: { auto v = make_pred<int32_t>(p); memcpy(a.bytes, &v, sizeof(v)); break; }
    case MSV_F32: { auto v = make_pred<float>(p); memcpy(a.bytes, &v, sizeof(v)); break; }
    case MSV_U64: { auto v = make_pred<uint64_t>(p); memcpy(a.bytes, &v, sizeof(v)); break; }
    case MSV_I64: { auto v = make_pred<int64_t>(p); memcpy(a.bytes, &v, sizeof(v)); break; }
    default: { auto v = make_pred<double>(p); memcpy(a.bytes, &v, sizeof(v)); break; }
  }
  return a;
}
static_assert(sizeof(Pred<double>) <= sizeof(PredArg), "predicate argument size");

static msv_status user_select_api(msv_executor* ex, const msv_slice* input,
                                  const msv_lazy_args* lz, const msv_user_op* pred,
                                  const msv_predicate* builtin, msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View in;
  LazyArgsHost h;
  if (lz) {
    if (!pred) return jit_check_op(ex, pred, MSV_USER_PREDICATE, MSV_U32, true);
    MSV_TRY(resolve_lazy(ex, lz, &h));
    in.len = lz->len;
    in.dtype = pred->dtype;
  } else {
    MSV_TRY(resolve(ex, *input, &in));
  }
  MSV_TRY(jit_check_op(ex, pred, MSV_USER_PREDICATE, in.dtype, lz != nullptr));
  if (!out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, in.dtype, in.len, true, out));
  Ctx c(ex);
  msv_status s = MSV_OK;
  uint32_t* out_len = (*out)->d_len;
  uint32_t n = in.len;
  if (n == 0) {
    hipError_t e = hipMemsetAsync(out_len, 0, sizeof(uint32_t), ex->stream);
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "%s", hipGetErrorString(e));
      s = MSV_ERR_LAUNCH;
    }
  } else {
    const bool large = n >= kSelLargeMin;
    const uint32_t tile = (uint32_t)kSelBlock *
        (large ? (dtype_size(in.dtype) == 4 ? sel_rows_large<uint32_t>() : sel_rows_large<uint64_t>())
               : kSelRowsSmall) * 4;
    uint32_t ntiles = div_up(n, tile);
    uint64_t *counter, *status;
    uint32_t gen;
    s = c.lookback(ntiles, &counter, &status, &gen);
    if (s == MSV_OK) {
      const void* ip = in.ptr;
      const void* stencil = nullptr;  // form 2 reads the input only
      void* op_ = (*out)->ptr;
      // The user functor is empty (one byte); the built-in one is Pred<T>.
      PredArg pa = pred->builtin ? pred_arg(in.dtype, builtin) : PredArg{};
      bool invert = false, aligned = lz || ((uintptr_t)in.ptr & 15u) == 0;
      const uint32_t* part_total = nullptr;
      void* a[] = {lz ? (void*)&h : (void*)&ip,
                   (void*)&stencil, (void*)pa.bytes, (void*)&invert,
                   (void*)&n,       (void*)&aligned, (void*)&op_,     (void*)&out_len,
                   (void*)&status,  (void*)&gen,    (void*)&ntiles,   (void*)&part_total};
      s = jit_launch(large ? pred->f_sel_large : pred->f_sel_small, ntiles, kSelBlock, c.stream(), a);
    }
  }
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

// Checks the arguments of an operator and generates its source and the name
// expressions of its kernels, in the order the create call stores them.
static msv_status gen_source(int32_t kind, msv_dtype dtype, const char* read_src,
                             int32_t builtin_op, const char* expr, const char* identity,
                             bool* builtin_out, std::string& src,
                             std::vector<std::string>& exprs) {
  const bool builtin = (kind == MSV_USER_REDUCTION && builtin_op >= 0) ||
                       (kind == MSV_USER_PREDICATE && !expr);
  const bool bad_kind = kind != MSV_USER_REDUCTION && kind != MSV_USER_PREDICATE &&
                        kind != MSV_LAZY_MAP;
  const bool is_float = dtype == MSV_F32 || dtype == MSV_F64;
  if (bad_kind || dtype < MSV_U32 || dtype > MSV_F64 ||
      (kind == MSV_USER_REDUCTION && !builtin && (!expr || !identity)) ||
      (kind == MSV_USER_REDUCTION && builtin && (builtin_op > OP_XOR || (builtin_op == OP_XOR && is_float))) ||
      (kind == MSV_LAZY_MAP && !read_src) || (builtin && !read_src)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0,
              "bad operator kind, dtype, built-in code, expression or read source");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  *builtin_out = builtin;
  const bool lazy = read_src != nullptr;
  const char* T = jit_c_type(dtype);
  const bool wide = dtype_size(dtype) == 8;
  // hiprtc has the device math functions but not <math.h>'s constants.
  src = "#include \"device_common.hpp\"\n#include \"scan_kernels.hpp\"\n"
        "#include \"select_kernels.hpp\"\n"
        "#ifndef INFINITY\n#define INFINITY __builtin_inff()\n#endif\n"
        "#ifndef NAN\n#define NAN __builtin_nanf(\"\")\n#endif\n"
        "namespace msv {\nusing UT = ";
  src += T;
  src += ";\n}  // namespace msv\n";
  if (lazy) {
    src += jit_fmt("namespace msv {\nstruct LazyArgs {\n  const void* p[%d];\n  uint32_t n[%d];\n"
               "  uint64_t q[%d];\n};\nstatic_assert(sizeof(LazyArgs) == 288, \"layout\");\n",
               MSV_LAZY_SLOTS, MSV_LAZY_SLOTS, MSV_LAZY_PARAMS);
    src += "#define MSV_SLOT(k, D, i) ((size_t)(i) < (size_t)la.n[k] ? "
           "((const D*)la.p[k])[(size_t)(i)] : (D)0)\n";
    src += read_src;
    // K consecutive elements: the source's own msv_read_v (vector slot reads)
    // or element by element.
    // Without them: element by element, and no tile is read unchecked.
    if (std::string(read_src).find("msv_read_v") == std::string::npos)
      src += "\ntemplate <int K>\n__device__ inline void msv_read_v(const LazyArgs& la, size_t i, UT (&o)[K]) {\n"
             "#pragma unroll\n  for (int k = 0; k < K; ++k) o[k] = msv_read(la, i + k);\n}\n";
    if (std::string(read_src).find("msv_tile_ok") == std::string::npos)
      src += "\ntemplate <int K>\n__device__ inline void msv_read_vf(const LazyArgs& la, size_t i, UT (&o)[K]) {\n"
             "  msv_read_v<K>(la, i, o);\n}\n"
             "__device__ inline bool msv_tile_ok(const LazyArgs&, size_t, size_t) { return false; }\n";
    src += "\nstruct LazySrc : LazySource {\n  LazyArgs la;\n"
           "  __device__ inline UT operator()(size_t i) const { return msv_read(la, i); }\n"
           "  template <int K>\n"
           "  __device__ inline void load(size_t i, UT (&v)[K]) const { msv_read_v<K>(la, i, v); }\n"
           "  template <int K>\n"
           "  __device__ inline void load_fast(size_t i, UT (&v)[K]) const { msv_read_vf<K>(la, i, v); }\n"
           "  __device__ inline bool tile_ok(size_t lo, size_t hi) const { return msv_tile_ok(la, lo, hi); }\n"
           "};\n"
           "struct LazySrcBits : LazySource {\n  LazyArgs la;\n"
           "  using U = typename Bits<UT>::U;\n"
           "  __device__ inline U operator()(size_t i) const { return bit_cast<U>(msv_read(la, i)); }\n"
           "  template <int K>\n  __device__ inline void load(size_t i, U (&v)[K]) const {\n"
           "    UT t[K];\n    msv_read_v<K>(la, i, t);\n"
           "#pragma unroll\n    for (int k = 0; k < K; ++k) v[k] = bit_cast<U>(t[k]);\n  }\n"
           "  template <int K>\n  __device__ inline void load_fast(size_t i, U (&v)[K]) const {\n"
           "    UT t[K];\n    msv_read_vf<K>(la, i, t);\n"
           "#pragma unroll\n    for (int k = 0; k < K; ++k) v[k] = bit_cast<U>(t[k]);\n  }\n"
           "  __device__ inline bool tile_ok(size_t lo, size_t hi) const { return msv_tile_ok(la, lo, hi); }\n"
           "};\n"
           "static_assert(sizeof(LazySrc) == 288 && sizeof(LazySrcBits) == 288, \"layout\");\n"
           "}  // namespace msv\n";
  }
  exprs.clear();
  if (kind == MSV_USER_REDUCTION) {
    const int code = builtin ? builtin_op : kOpUser;
    if (!builtin) {
      src += jit_fmt("namespace msv {\ntemplate <> struct Combine<UT, %d> {\n", kOpUser);
      src += "  __host__ __device__ static inline UT apply(UT a, UT b) { return (UT)(";
      src += expr;
      src += "); }\n  __host__ __device__ static inline UT identity() { return (UT)(";
      src += identity;
      src += "); }\n};\n}  // namespace msv\n";
    }
    src += jit_fmt("extern \"C\" __global__ void msv_user_identity(msv::UT* o) { *o = msv::Combine<msv::UT, %d>::identity(); }\n", code);
    exprs.push_back(jit_fmt("msv::reduce_tile_ordered_kernel<%s, %d, %d, %d, true>", T, code, kReduceBlock, kReduceRows));
    exprs.push_back(jit_fmt("msv::reduce_tile_ordered_kernel<%s, %d, %d, %d, false>", T, code, kScanBlock, kScanRowsSmall));
    exprs.push_back(jit_fmt("msv::scan_tile_kernel<%s, %d, true, %d, %d, %d>", T, code, (int)SCAN_PREFIXED, kScanBlock, kScanRowsSmall));
    exprs.push_back(jit_fmt("msv::scan_tile_kernel<%s, %d, false, %d, %d, %d>", T, code, (int)SCAN_PREFIXED, kScanBlock, kScanRowsSmall));
    if (lazy && builtin && !dtype_is_float(dtype)) {
      const int large = lazy_scan_rows_large(dtype_size(dtype));
      for (int excl = 0; excl < 2; ++excl)
        for (int rows : {kScanRowsSmall, large})
          exprs.push_back(jit_fmt("msv::scan_tile_kernel<%s, %d, %s, %d, %d, %d, msv::LazySrc>", T, code,
                              excl ? "true" : "false", (int)SCAN_LOOKBACK, kScanBlock, rows));
    }
    if (lazy) {
      exprs.push_back(jit_fmt("msv::reduce_tile_ordered_kernel<%s, %d, %d, %d, true, msv::LazySrc>", T, code, kReduceBlock, kReduceRows));
      exprs.push_back(jit_fmt("msv::reduce_tile_ordered_kernel<%s, %d, %d, %d, false, msv::LazySrc>", T, code, kScanBlock, kScanRowsSmall));
      exprs.push_back(jit_fmt("msv::scan_tile_kernel<%s, %d, true, %d, %d, %d, msv::LazySrc>", T, code, (int)SCAN_PREFIXED, kScanBlock, kScanRowsSmall));
      exprs.push_back(jit_fmt("msv::scan_tile_kernel<%s, %d, false, %d, %d, %d, msv::LazySrc>", T, code, (int)SCAN_PREFIXED, kScanBlock, kScanRowsSmall));
    }
  } else if (kind == MSV_USER_PREDICATE) {
    std::string P = "msv::Pred<msv::UT>";
    if (!builtin) {
      src += "namespace msv {\nstruct UserPred {\n  __device__ inline bool operator()(UT x) const { return (bool)(";
      src += expr;
      src += "); }\n};\n}  // namespace msv\n";
      P = "msv::UserPred";
    }
    const char* U = wide ? "uint64_t" : "uint32_t";
    const char* S = lazy ? ", msv::LazySrcBits" : "";
    exprs.push_back(jit_fmt("msv::select_kernel<%s, %s, 2, %d, false, %s%s>", U, T, kSelRowsSmall, P.c_str(), S));
    exprs.push_back(jit_fmt("msv::select_kernel<%s, %s, 2, %d, false, %s%s>", U, T,
                        wide ? sel_rows_large<uint64_t>() : sel_rows_large<uint32_t>(), P.c_str(), S));
  } else {
    src += jit_fmt("extern \"C\" __global__ __launch_bounds__(%d) void msv_lazy_map(msv::LazySrc s,\n"
               "    msv::UT* __restrict__ out, uint32_t n) {\n"
               "  constexpr int V = 16 / sizeof(msv::UT);\n"
               "  const size_t i0 = ((size_t)blockIdx.x * %d + threadIdx.x) * V;\n"
               "  if (i0 >= n) return;\n"
               "  msv::Vec16<msv::UT> w;\n"
               "  if (i0 + V <= n) {\n"
               "    s.load<V>(i0, w.v);\n"
               "    *reinterpret_cast<msv::Vec16<msv::UT>*>(out + i0) = w;\n"
               "  } else {\n"
               "    for (int k = 0; k < V; ++k) if (i0 + k < n) out[i0 + k] = s(i0 + k);\n"
               "  }\n}\n",
               kMapBlock, kMapBlock);
  }
  return MSV_OK;
}

extern "C" {

msv_status msv_lazy_op_create(msv_executor* ex, int32_t kind, msv_dtype dtype,
                              const char* read_src, int32_t builtin_op, const char* expr,
                              const char* identity, msv_user_op** out) {
  clear_error();
  if (!ex || !out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor or output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  std::string src;
  std::vector<std::string> exprs;
  bool builtin = false;
  MSV_TRY(gen_source(kind, dtype, read_src, builtin_op, expr, identity, &builtin, src, exprs));
  msv_user_op* op = new msv_user_op;
  op->kind = kind;
  op->dtype = dtype;
  op->owner_id = ex->id;
  op->device = ex->device;
  op->lazy = read_src != nullptr;
  op->builtin = builtin;
  std::vector<hipFunction_t> fns;
  msv_status s = jit_build(op, src, exprs, fns);
  if (s == MSV_OK && kind == MSV_USER_REDUCTION) {
    op->f_reduce = fns[0];
    op->f_partials = fns[1];
    op->f_scan_ex = fns[2];
    op->f_scan_in = fns[3];
    size_t at = 4;
    if (op->lazy && op->builtin && !dtype_is_float(dtype)) {
      op->f_lb[0][0] = fns[at++];
      op->f_lb[0][1] = fns[at++];
      op->f_lb[1][0] = fns[at++];
      op->f_lb[1][1] = fns[at++];
    }
    if (op->lazy) {
      op->f_reduce_lz = fns[at++];
      op->f_partials_lz = fns[at++];
      op->f_scan_ex_lz = fns[at++];
      op->f_scan_in_lz = fns[at++];
    }
    // The identity's bits, evaluated once on the device.
    hipFunction_t fid = nullptr;
    void* d = nullptr;
    hipError_t e = hipModuleGetFunction(&fid, op->module, "msv_user_identity");
    if (e == hipSuccess) e = hipMalloc(&d, 8);
    if (e == hipSuccess) {
      void* a[] = {(void*)&d};
      e = hipModuleLaunchKernel(fid, 1, 1, 1, 1, 1, 1, 0, ex->stream, a, nullptr);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&op->identity, d, dtype_size(dtype), hipMemcpyDeviceToHost, ex->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ex->stream);
    if (d) (void)hipFree(d);
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "user operator identity: %s", hipGetErrorString(e));
      s = MSV_ERR_LAUNCH;
    }
  } else if (s == MSV_OK && kind == MSV_USER_PREDICATE) {
    op->f_sel_small = fns[0];
    op->f_sel_large = fns[1];
  } else if (s == MSV_OK) {
    hipError_t e = hipModuleGetFunction(&op->f_map, op->module, "msv_lazy_map");
    if (e != hipSuccess) {
      set_error(MSV_ERR_LAUNCH, 0, 0, "lazy map kernel: %s", hipGetErrorString(e));
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

msv_status msv_user_op_create(msv_executor* ex, int32_t kind, msv_dtype dtype, const char* expr,
                              const char* identity, msv_user_op** out) {
  clear_error();
  if (kind == MSV_USER_LESS || kind == MSV_USER_EQUAL) {
    if (!ex || !out) {
      set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null executor or output");
      return MSV_ERR_INVALID_ARGUMENT;
    }
    return cmp_op_create(ex, kind, dtype, expr, out);
  }
  if (!expr || (kind != MSV_USER_REDUCTION && kind != MSV_USER_PREDICATE)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad user operator kind, dtype or expression");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return msv_lazy_op_create(ex, kind, dtype, nullptr, -1, expr, identity, out);
}

msv_status msv_lazy_op_check(int32_t kind, msv_dtype dtype, const char* read_src,
                             int32_t builtin_op, const char* expr, const char* identity) {
  clear_error();
  std::string src;
  std::vector<std::string> exprs;
  bool builtin = false;
  if (kind == MSV_USER_LESS || kind == MSV_USER_EQUAL) {
    if (read_src) {
      set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "a comparator takes no read source");
      return MSV_ERR_INVALID_ARGUMENT;
    }
    MSV_TRY(cmp_gen_source(kind, dtype, expr, src, exprs));
  } else {
    MSV_TRY(gen_source(kind, dtype, read_src, builtin_op, expr, identity, &builtin, src, exprs));
  }
  std::vector<char> code;
  std::vector<std::string> lowered;
  return jit_compile(src, exprs, code, lowered);
}

msv_status msv_user_op_destroy(msv_user_op* op) {
  if (!op) return MSV_ERR_INVALID_ARGUMENT;
  if (op->module) {
    (void)hipSetDevice(op->device);
    (void)hipModuleUnload(op->module);
  }
  delete op;
  return MSV_OK;
}

msv_status msv_reduce_user(msv_executor* ex, msv_slice input, const void* init,
                           const msv_user_op* op, void* out_host) {
  return user_reduce_api(ex, &input, nullptr, init, op, out_host);
}

msv_status msv_inclusive_scan_user(msv_executor* ex, msv_slice input, const msv_user_op* op,
                                   msv_vec** out) {
  return user_scan_api(ex, &input, nullptr, nullptr, op, false, out);
}

msv_status msv_exclusive_scan_user(msv_executor* ex, msv_slice input, const void* init,
                                   const msv_user_op* op, msv_vec** out) {
  return user_scan_api(ex, &input, nullptr, init, op, true, out);
}

msv_status msv_copy_where_user(msv_executor* ex, msv_slice input, const msv_user_op* pred,
                               msv_vec** out) {
  if (pred && pred->builtin) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "operator has no user predicate");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return user_select_api(ex, &input, nullptr, pred, nullptr, out);
}

msv_status msv_reduce_lazy(msv_executor* ex, const msv_lazy_args* input, const void* init,
                           const msv_user_op* op, void* out_host) {
  return user_reduce_api(ex, nullptr, input, init, op, out_host);
}

msv_status msv_inclusive_scan_lazy(msv_executor* ex, const msv_lazy_args* input,
                                   const msv_user_op* op, msv_vec** out) {
  return user_scan_api(ex, nullptr, input, nullptr, op, false, out);
}

msv_status msv_exclusive_scan_lazy(msv_executor* ex, const msv_lazy_args* input, const void* init,
                                   const msv_user_op* op, msv_vec** out) {
  return user_scan_api(ex, nullptr, input, init, op, true, out);
}

msv_status msv_copy_where_lazy(msv_executor* ex, const msv_lazy_args* input,
                               const msv_user_op* pred, const msv_predicate* builtin,
                               msv_vec** out) {
  return user_select_api(ex, nullptr, input, pred, builtin, out);
}

msv_status msv_map_lazy(msv_executor* ex, const msv_lazy_args* input, const msv_user_op* op,
                        msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a lazy input"));
  if (!op) return jit_check_op(ex, op, MSV_LAZY_MAP, MSV_U32, true);
  LazyArgsHost h;
  MSV_TRY(resolve_lazy(ex, input, &h));
  MSV_TRY(jit_check_op(ex, op, MSV_LAZY_MAP, op->dtype, true));
  if (!out) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  uint32_t n = input->len;
  MSV_TRY(new_vec(ex, op->dtype, n, false, out));
  if (n == 0) return MSV_OK;
  const uint32_t per_block = (uint32_t)kMapBlock * (uint32_t)(16 / dtype_size(op->dtype));
  const uint32_t blocks = (uint32_t)(((uint64_t)n + per_block - 1) / per_block);
  void* o = (*out)->ptr;
  void* a[] = {(void*)&h, (void*)&o, (void*)&n};
  msv_status s = jit_launch(op->f_map, blocks, kMapBlock, ex->stream, a);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

}  // extern "C"
