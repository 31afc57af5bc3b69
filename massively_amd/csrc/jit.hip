// User-defined operators, compiled at run time for gfx950 (hiprtc).
//
// The reference's algorithms take user operators as CubeCL traits compiled
// into its kernels (ReductionOp::apply, core/reduce.rs:63-66; PredicateOp::apply,
// core/predicate.rs:40-43; runtime compilation of the read expressions,
// core/eval.rs:21-107). Here a user operator is a C++ expression over the
// element type T:
//   reduction:  an expression of `a` and `b` (associative; commutativity is
//               not needed: every tree and scan below combines in index
//               order), plus an identity expression;
//   predicate:  an expression of `x` (nonzero = keep).
// msv_user_op_create compiles, once, the library's own kernel templates
// instantiated with that operator: the same device headers the library is
// built from are embedded in it (jit_headers.inc), the operator becomes a
// specialisation Combine<T, kOpUser> or a functor UserPred, and hiprtc's name
// expressions give the instantiations' symbols in the loaded module. Reduce
// and the scans' tile partials use an order-preserving tile reduction (rows in
// index order); the scans use the fixed-shape
// (prefixed) passes of the float scans, which keep the order of combination;
// copy_where uses the one-pass compaction with the user predicate on the
// input itself.
#include <hip/hiprtc.h>

#include <string>

#include "msv_internal.hpp"
#include "scan_kernels.hpp"
#include "select_kernels.hpp"
#include "jit_headers.inc"

namespace msv {
constexpr int kOpUser = 100;  // Combine<T, kOpUser> is the user's operator
}  // namespace msv

struct msv_user_op {
  int32_t kind;
  msv_dtype dtype;
  uint64_t owner_id;
  int device;
  hipModule_t module = nullptr;
  // reduction: reduce (512 x 16), tile partials (256 x 4), exclusive / inclusive prefixed scans
  hipFunction_t f_reduce = nullptr, f_partials = nullptr, f_scan_ex = nullptr, f_scan_in = nullptr;
  uint64_t identity = 0;  // bits of the identity, for host-side padding values
  // predicate: compaction kernels for short and long inputs
  hipFunction_t f_sel_small = nullptr, f_sel_large = nullptr;
};

using namespace msv;

static const char* c_type(msv_dtype dt) {
  switch (dt) {
    case MSV_U32: return "uint32_t";
    case MSV_I32: return "int32_t";
    case MSV_F32: return "float";
    case MSV_U64: return "uint64_t";
    case MSV_I64: return "int64_t";
    default: return "double";
  }
}

static std::string fmt(const char* f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

// Compiles `src`, loads the code object and looks up each name expression.
static msv_status jit_build(msv_user_op* op, const std::string& src,
                            const std::vector<std::string>& exprs,
                            std::vector<hipFunction_t>& fns) {
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
  std::vector<char> code(code_size);
  hiprtcGetCode(prog, code.data());
  std::vector<std::string> lowered;
  for (const std::string& e : exprs) {
    const char* name = nullptr;
    hiprtcGetLoweredName(prog, e.c_str(), &name);
    lowered.push_back(name ? name : "");
  }
  hiprtcDestroyProgram(&prog);
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

static msv_status launch(hipFunction_t f, uint32_t grid, uint32_t block, hipStream_t s,
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

static msv_status check_op(msv_executor* ex, const msv_user_op* op, int32_t kind,
                           msv_dtype dt) {
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
  return MSV_OK;
}

// Fixed-shape scan of n elements (the float scans' passes): tile partials, an
// exclusive scan of them (recursively), then the prefixed tile scan.
static msv_status user_scan(Ctx& c, const msv_user_op* op, const void* in, void* out, uint32_t n,
                            bool exclusive, Scalar init) {
  const size_t es = dtype_size(op->dtype);
  const uint32_t tile = (uint32_t)kScanBlock * kScanRowsSmall * (uint32_t)(16 / es);
  const uint32_t ntiles = (n + tile - 1) / tile;
  bool aligned = (((uintptr_t)in | (uintptr_t)out) & 15u) == 0;
  const void* prefixes = nullptr;
  Scalar ident = scalar_of(&op->identity, es);
  if (ntiles > 1) {
    void* partials = c.scratch((size_t)ntiles * es);
    void* pre = c.scratch((size_t)ntiles * es);
    if (!partials || !pre) return MSV_ERR_OUT_OF_MEMORY;
    bool no = false;
    void* a[] = {(void*)&in, (void*)&n, (void*)&aligned, (void*)&partials, (void*)&no, (void*)&ident};
    MSV_TRY(launch(op->f_partials, ntiles, kScanBlock, c.stream(), a));
    MSV_TRY(user_scan(c, op, partials, pre, ntiles, true, ident));
    prefixes = pre;
  }
  uint64_t* status = nullptr;
  uint32_t gen = 0;
  void* a[] = {(void*)&in,    (void*)&out,    (void*)&n,   (void*)&init,
               (void*)&aligned, (void*)&status, (void*)&gen, (void*)&prefixes};
  return launch(exclusive ? op->f_scan_ex : op->f_scan_in, ntiles, kScanBlock, c.stream(), a);
}

static msv_status user_scan_api(msv_executor* ex, msv_slice input, const void* init,
                                const msv_user_op* op, bool exclusive, msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View in;
  MSV_TRY(resolve(ex, input, &in));
  MSV_TRY(check_op(ex, op, MSV_USER_REDUCTION, in.dtype));
  if (!out || (exclusive && !init)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null output or init");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  MSV_TRY(new_vec(ex, in.dtype, in.len, false, out));
  if (in.len == 0) return MSV_OK;
  Ctx c(ex);
  const size_t es = dtype_size(in.dtype);
  const Scalar iv = exclusive ? scalar_of(init, es) : scalar_of(&op->identity, es);
  msv_status s = user_scan(c, op, in.ptr, (*out)->ptr, in.len, exclusive, iv);
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

extern "C" {

msv_status msv_user_op_create(msv_executor* ex, int32_t kind, msv_dtype dtype, const char* expr,
                              const char* identity, msv_user_op** out) {
  clear_error();
  if (!ex || !expr || !out || (kind != MSV_USER_REDUCTION && kind != MSV_USER_PREDICATE) ||
      (kind == MSV_USER_REDUCTION && !identity) || dtype < MSV_U32 || dtype > MSV_F64) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "bad user operator kind, dtype or expression");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  msv_user_op* op = new msv_user_op;
  op->kind = kind;
  op->dtype = dtype;
  op->owner_id = ex->id;
  op->device = ex->device;
  const char* T = c_type(dtype);
  const bool wide = dtype_size(dtype) == 8;
  // hiprtc has the device math functions but not <math.h>'s constants.
  std::string src = "#include \"device_common.hpp\"\n#include \"scan_kernels.hpp\"\n"
                    "#include \"select_kernels.hpp\"\n"
                    "#ifndef INFINITY\n#define INFINITY __builtin_inff()\n#endif\n"
                    "#ifndef NAN\n#define NAN __builtin_nanf(\"\")\n#endif\n"
                    "namespace msv {\nusing UT = ";
  src += T;
  src += ";\n";
  std::vector<std::string> exprs;
  if (kind == MSV_USER_REDUCTION) {
    src += fmt("template <> struct Combine<UT, %d> {\n", kOpUser);
    src += "  __host__ __device__ static inline UT apply(UT a, UT b) { return (UT)(";
    src += expr;
    src += "); }\n  __host__ __device__ static inline UT identity() { return (UT)(";
    src += identity;
    src += "); }\n};\n}  // namespace msv\n";
    src += fmt("extern \"C\" __global__ void msv_user_identity(msv::UT* o) { *o = msv::Combine<msv::UT, %d>::identity(); }\n", kOpUser);
    exprs.push_back(fmt("msv::reduce_tile_ordered_kernel<%s, %d, %d, %d>", T, kOpUser, kReduceBlock, kReduceRows));
    exprs.push_back(fmt("msv::reduce_tile_ordered_kernel<%s, %d, %d, %d>", T, kOpUser, kScanBlock, kScanRowsSmall));
    exprs.push_back(fmt("msv::scan_tile_kernel<%s, %d, true, %d, %d, %d>", T, kOpUser, (int)SCAN_PREFIXED, kScanBlock, kScanRowsSmall));
    exprs.push_back(fmt("msv::scan_tile_kernel<%s, %d, false, %d, %d, %d>", T, kOpUser, (int)SCAN_PREFIXED, kScanBlock, kScanRowsSmall));
  } else {
    src += "struct UserPred {\n  __device__ inline bool operator()(UT x) const { return (bool)(";
    src += expr;
    src += "); }\n};\n}  // namespace msv\n";
    const char* U = wide ? "uint64_t" : "uint32_t";
    exprs.push_back(fmt("msv::select_kernel<%s, %s, 2, %d, false, msv::UserPred>", U, T, kSelRowsSmall));
    exprs.push_back(fmt("msv::select_kernel<%s, %s, 2, %d, false, msv::UserPred>", U, T, kSelRowsLarge));
  }
  std::vector<hipFunction_t> fns;
  msv_status s = jit_build(op, src, exprs, fns);
  if (s == MSV_OK && kind == MSV_USER_REDUCTION) {
    op->f_reduce = fns[0];
    op->f_partials = fns[1];
    op->f_scan_ex = fns[2];
    op->f_scan_in = fns[3];
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
  } else if (s == MSV_OK) {
    op->f_sel_small = fns[0];
    op->f_sel_large = fns[1];
  }
  if (s != MSV_OK) {
    if (op->module) (void)hipModuleUnload(op->module);
    delete op;
    return s;
  }
  *out = op;
  return MSV_OK;
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
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View in;
  MSV_TRY(resolve(ex, input, &in));
  MSV_TRY(check_op(ex, op, MSV_USER_REDUCTION, in.dtype));
  if (!init || !out_host) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null init or output");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  Ctx c(ex);
  const size_t es = dtype_size(in.dtype);
  const uint32_t tile = (uint32_t)kReduceBlock * kReduceRows * (uint32_t)(16 / es);
  void* d_result = c.scratch(es);
  if (!d_result) return MSV_ERR_OUT_OF_MEMORY;
  // Levels of the fixed tree; the last one applies op(init, total).
  const void* src = in.ptr;
  uint32_t m = in.len;
  bool aligned = ((uintptr_t)in.ptr & 15u) == 0;
  Scalar iv = scalar_of(init, es);
  for (;;) {
    const uint32_t tiles = m == 0 ? 1 : (m + tile - 1) / tile;
    bool last = tiles == 1;
    void* dst = last ? d_result : c.scratch((size_t)tiles * es);
    if (!dst) return MSV_ERR_OUT_OF_MEMORY;
    void* a[] = {(void*)&src, (void*)&m, (void*)&aligned, (void*)&dst, (void*)&last, (void*)&iv};
    MSV_TRY(launch(op->f_reduce, tiles, kReduceBlock, c.stream(), a));
    if (last) break;
    src = dst;
    m = tiles;
    aligned = true;
  }
  MSV_HIP(hipMemcpyAsync(out_host, d_result, es, hipMemcpyDeviceToHost, c.stream()));
  MSV_HIP(hipStreamSynchronize(c.stream()));
  return MSV_OK;
}

msv_status msv_inclusive_scan_user(msv_executor* ex, msv_slice input, const msv_user_op* op,
                                   msv_vec** out) {
  return user_scan_api(ex, input, nullptr, op, false, out);
}

msv_status msv_exclusive_scan_user(msv_executor* ex, msv_slice input, const void* init,
                                   const msv_user_op* op, msv_vec** out) {
  return user_scan_api(ex, input, init, op, true, out);
}

msv_status msv_copy_where_user(msv_executor* ex, msv_slice input, const msv_user_op* pred,
                               msv_vec** out) {
  clear_error();
  MSV_TRY(no_capture(ex, "a user-defined operator"));
  View in;
  MSV_TRY(resolve(ex, input, &in));
  MSV_TRY(check_op(ex, pred, MSV_USER_PREDICATE, in.dtype));
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
    const uint32_t tile = (uint32_t)kSelBlock * (large ? kSelRowsLarge : kSelRowsSmall) * 4;
    uint32_t ntiles = (n + tile - 1) / tile;
    uint64_t *counter, *status;
    uint32_t gen;
    s = c.lookback(ntiles, &counter, &status, &gen);
    if (s == MSV_OK) {
      const void* ip = in.ptr;
      void* op_ = (*out)->ptr;
      unsigned char pred_obj = 0;  // the functor is empty
      bool invert = false, aligned = ((uintptr_t)in.ptr & 15u) == 0;
      const uint32_t* part_total = nullptr;
      void* a[] = {(void*)&ip,      (void*)&ip,     (void*)&pred_obj, (void*)&invert,
                   (void*)&n,       (void*)&aligned, (void*)&op_,     (void*)&out_len,
                   (void*)&status,  (void*)&gen,    (void*)&ntiles,   (void*)&part_total};
      s = launch(large ? pred->f_sel_large : pred->f_sel_small, ntiles, kSelBlock, c.stream(), a);
    }
  }
  if (s != MSV_OK) {
    msv_vec_release(*out);
    *out = nullptr;
  }
  return s;
}

}  // extern "C"
