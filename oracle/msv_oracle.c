/*
 * msv_oracle.c -- CPU restatement of massively's reference CPU path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle the GPU kernels are
 * checked against and the CPU baseline bench.py times; nothing in the product
 * library links or calls it.
 *
 * Each function restates a sequential loop of the reference's oracle crate
 * (paths relative to the massively repository):
 *   verification/oracle/src/vector.rs  reduce :33-42, scans :44-74,
 *     copy/remove_where :255-271, gather/scatter :285-311, sort/sort_by_key
 *     :483-524 (stable via sort_by_less :666-680), merge :493-513,
 *     merge_by_key :526-558, set ops :595-664
 *   verification/oracle/src/seg.rs     sort :18-39, scans :48-88, reduce :155-169
 * plus GPU-side contracts the oracle crate leaves implicit: integer addition
 * wraps modulo 2^bits; radix order is integer order / IEEE total order
 * (massively/src/core/radix.rs:13-118); exclusive scans start from init
 * (core/scan.rs:1209-1233); flags are true iff nonzero (flag.rs:17-19); the
 * seeded PCG input streams of massively/src/api/util/random.rs:12-22,331-396.
 *
 * Built-in operators (the closed set the engine serves): ADD, MIN
 * (b<a ? b : a), MAX (a<b ? b : a), XOR (integers).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { DT_U32 = 0, DT_I32 = 1, DT_F32 = 2, DT_U64 = 3, DT_I64 = 4, DT_F64 = 5 };
enum { OP_ADD = 0, OP_MIN = 1, OP_MAX = 2, OP_XOR = 3 };
/* Sort/merge comparison modes: total order (radix_sort_by_key), Less, Greater. */
enum { ORD_TOTAL = 0, ORD_LESS = 1, ORD_GREATER = 2 };

static size_t dt_size(int dt) { return (dt == DT_U32 || dt == DT_I32 || dt == DT_F32) ? 4 : 8; }

/* ---- operators (integer add wraps) ----------------------------------- */
#define INT_OPS(T, U, SUF)                                   \
  static inline T apply_##SUF(int op, T a, T b) {            \
    switch (op) {                                            \
      case OP_ADD: return (T)((U)a + (U)b);                  \
      case OP_MIN: return (b < a) ? b : a;                   \
      case OP_MAX: return (a < b) ? b : a;                   \
      default: return (T)((U)a ^ (U)b);                      \
    }                                                        \
  }
#define FLT_OPS(T, SUF)                                      \
  static inline T apply_##SUF(int op, T a, T b) {            \
    switch (op) {                                            \
      case OP_ADD: return a + b;                             \
      case OP_MIN: return (b < a) ? b : a;                   \
      default: return (a < b) ? b : a;                       \
    }                                                        \
  }
INT_OPS(uint32_t, uint32_t, u32)
INT_OPS(int32_t, uint32_t, i32)
INT_OPS(uint64_t, uint64_t, u64)
INT_OPS(int64_t, uint64_t, i64)
FLT_OPS(float, f32)
FLT_OPS(double, f64)

/* ---- scans and reduce (vector.rs:33-74, seg.rs:48-88,155-169) --------- */
#define SCAN_FNS(T, SUF)                                                          \
  static void incl_##SUF(int op, const T* x, T* y, uint64_t n) {                  \
    T acc = 0;                                                                    \
    for (uint64_t i = 0; i < n; ++i) {                                            \
      acc = (i == 0) ? x[0] : apply_##SUF(op, acc, x[i]);                         \
      y[i] = acc;                                                                 \
    }                                                                             \
  }                                                                               \
  static void excl_##SUF(int op, const T* x, T* y, uint64_t n, T init) {          \
    T acc = init;                                                                 \
    for (uint64_t i = 0; i < n; ++i) {                                            \
      y[i] = acc;                                                                 \
      acc = apply_##SUF(op, acc, x[i]);                                           \
    }                                                                             \
  }                                                                               \
  static T fold_##SUF(int op, const T* x, uint64_t n, T init) {                   \
    T acc = init;                                                                 \
    for (uint64_t i = 0; i < n; ++i) acc = apply_##SUF(op, acc, x[i]);            \
    return acc;                                                                   \
  }
SCAN_FNS(uint32_t, u32)
SCAN_FNS(int32_t, i32)
SCAN_FNS(uint64_t, u64)
SCAN_FNS(int64_t, i64)
SCAN_FNS(float, f32)
SCAN_FNS(double, f64)

#define DISPATCH(dt, STMT_U32, STMT_I32, STMT_F32, STMT_U64, STMT_I64, STMT_F64) \
  switch (dt) {                                                                   \
    case DT_U32: STMT_U32; break;                                                 \
    case DT_I32: STMT_I32; break;                                                 \
    case DT_F32: STMT_F32; break;                                                 \
    case DT_U64: STMT_U64; break;                                                 \
    case DT_I64: STMT_I64; break;                                                 \
    case DT_F64: STMT_F64; break;                                                 \
    default: return -1;                                                           \
  }

int msvo_inclusive_scan(int dt, int op, const void* in, void* out, uint64_t n) {
  DISPATCH(dt, incl_u32(op, in, out, n), incl_i32(op, in, out, n), incl_f32(op, in, out, n),
           incl_u64(op, in, out, n), incl_i64(op, in, out, n), incl_f64(op, in, out, n))
  return 0;
}

int msvo_exclusive_scan(int dt, int op, const void* in, void* out, uint64_t n, const void* init) {
  DISPATCH(dt, excl_u32(op, in, out, n, *(const uint32_t*)init),
           excl_i32(op, in, out, n, *(const int32_t*)init),
           excl_f32(op, in, out, n, *(const float*)init),
           excl_u64(op, in, out, n, *(const uint64_t*)init),
           excl_i64(op, in, out, n, *(const int64_t*)init),
           excl_f64(op, in, out, n, *(const double*)init))
  return 0;
}

/* vector.rs:33-42: fold(init, op). */
int msvo_reduce(int dt, int op, const void* in, uint64_t n, const void* init, void* out) {
  DISPATCH(dt, *(uint32_t*)out = fold_u32(op, in, n, *(const uint32_t*)init),
           *(int32_t*)out = fold_i32(op, in, n, *(const int32_t*)init),
           *(float*)out = fold_f32(op, in, n, *(const float*)init),
           *(uint64_t*)out = fold_u64(op, in, n, *(const uint64_t*)init),
           *(int64_t*)out = fold_i64(op, in, n, *(const int64_t*)init),
           *(double*)out = fold_f64(op, in, n, *(const double*)init))
  return 0;
}

/* FP64 oracle for f32 sums: exact-enough f64 sum and sum of magnitudes (the
 * error bound is stated in terms of sum |x_i|). Pairwise in f64. */
static double pairwise_f64(const float* x, uint64_t n, int absval) {
  if (n <= 64) {
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) s += absval ? fabs((double)x[i]) : (double)x[i];
    return s;
  }
  uint64_t h = n / 2;
  return pairwise_f64(x, h, absval) + pairwise_f64(x + h, n - h, absval);
}
void msvo_sum_f32_f64(const float* x, uint64_t n, double* sum, double* abs_sum) {
  *sum = pairwise_f64(x, n, 0);
  *abs_sum = pairwise_f64(x, n, 1);
}

/* ---- selection (vector.rs:255-271) ------------------------------------ */
uint64_t msvo_copy_where(int dt, const void* in, const uint32_t* stencil, uint64_t n, int invert,
                         void* out) {
  const size_t es = dt_size(dt);
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const int keep = (stencil[i] != 0) != (invert != 0);
    if (keep) memcpy((char*)out + k++ * es, (const char*)in + i * es, es);
  }
  return k;
}

/* Built-in predicates on a stencil element of dtype sdt (thresholds `a`, `b`
 * carry the bit pattern of a value of that dtype). */
enum { PR_NONZERO, PR_ZERO, PR_LT, PR_LE, PR_GT, PR_GE, PR_BITS_ANY, PR_BITS_EQ };
#define PRED_CMP(T, x, a)                      \
  switch (kind) {                              \
    case PR_NONZERO: return x != (T)0;         \
    case PR_ZERO: return x == (T)0;            \
    case PR_LT: return x < a;                  \
    case PR_LE: return x <= a;                 \
    case PR_GT: return x > a;                  \
    case PR_GE: return x >= a;                 \
  }
static int pred_eval(int sdt, const void* p, int kind, uint64_t a, uint64_t b) {
  uint64_t bits = (dt_size(sdt) == 4) ? *(const uint32_t*)p : *(const uint64_t*)p;
  if (kind == PR_BITS_ANY) return (bits & a) != 0;
  if (kind == PR_BITS_EQ) return (bits & a) == b;
  switch (sdt) {
    case DT_U32: { uint32_t x = (uint32_t)bits, t = (uint32_t)a; PRED_CMP(uint32_t, x, t) break; }
    case DT_I32: { int32_t x = (int32_t)(uint32_t)bits, t = (int32_t)(uint32_t)a; PRED_CMP(int32_t, x, t) break; }
    case DT_U64: { uint64_t x = bits, t = a; PRED_CMP(uint64_t, x, t) break; }
    case DT_I64: { int64_t x = (int64_t)bits, t = (int64_t)a; PRED_CMP(int64_t, x, t) break; }
    case DT_F32: { float x, t; uint32_t xb = (uint32_t)bits, tb = (uint32_t)a;
      memcpy(&x, &xb, 4); memcpy(&t, &tb, 4); PRED_CMP(float, x, t) break; }
    case DT_F64: { double x, t; memcpy(&x, &bits, 8); memcpy(&t, &a, 8); PRED_CMP(double, x, t) break; }
  }
  return 0;
}

uint64_t msvo_copy_where_pred(int dt, const void* in, int sdt, const void* stencil, uint64_t n,
                              int kind, uint64_t a, uint64_t b, int invert, void* out) {
  const size_t es = dt_size(dt), ss = dt_size(sdt);
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const int keep = pred_eval(sdt, (const char*)stencil + i * ss, kind, a, b) != (invert != 0);
    if (keep) memcpy((char*)out + k++ * es, (const char*)in + i * es, es);
  }
  return k;
}

/* ---- movement (vector.rs:285-311) ---------------------------------------- */
void msvo_gather(int dt, const void* values, const uint32_t* idx, uint64_t n, void* out) {
  const size_t es = dt_size(dt);
  for (uint64_t i = 0; i < n; ++i) memcpy((char*)out + i * es, (const char*)values + idx[i] * es, es);
}

void msvo_scatter(int dt, const void* values, const uint32_t* idx, uint64_t n, void* out) {
  const size_t es = dt_size(dt);
  for (uint64_t i = 0; i < n; ++i) memcpy((char*)out + (uint64_t)idx[i] * es, (const char*)values + i * es, es);
}

/* ---- ordering keys ----------------------------------------------------------
 * Maps a key to an unsigned 64-bit word whose numeric order is the requested
 * order: ORD_TOTAL = integer order / IEEE total order (core/radix.rs:13-118);
 * ORD_LESS = the comparator `<` (floats: -0 and +0 compare equal);
 * ORD_GREATER = reverse of ORD_LESS. NaN keys under ORD_LESS/GREATER are
 * outside the reference's strict-weak-order contract. */
uint64_t msvo_order_word(int dt, const void* p, int mode) {
  uint64_t w = 0;
  switch (dt) {
    case DT_U32: w = *(const uint32_t*)p; break;
    case DT_I32: w = (uint64_t)(*(const uint32_t*)p ^ 0x80000000u); break;
    case DT_U64: w = *(const uint64_t*)p; break;
    case DT_I64: w = *(const uint64_t*)p ^ 0x8000000000000000ull; break;
    case DT_F32: {
      uint32_t b = *(const uint32_t*)p;
      if (mode != ORD_TOTAL && (b & 0x7fffffffu) == 0) b = 0;
      w = (b & 0x80000000u) ? (uint64_t)(uint32_t)~b : (uint64_t)(b ^ 0x80000000u);
      break;
    }
    case DT_F64: {
      uint64_t b = *(const uint64_t*)p;
      if (mode != ORD_TOTAL && (b & 0x7fffffffffffffffull) == 0) b = 0;
      w = (b & 0x8000000000000000ull) ? ~b : (b ^ 0x8000000000000000ull);
      break;
    }
  }
  return mode == ORD_GREATER ? ~w : w;
}

typedef struct { uint64_t w; uint32_t i; } Item;

/* Stable bottom-up merge sort on (word, index): equal words keep index order,
 * exactly as sort_by_less's stable sort (vector.rs:666-680). */
static void stable_sort_items(Item* a, uint64_t n) {
  Item* tmp = (Item*)malloc(sizeof(Item) * (n ? n : 1));
  const uint64_t RUN = 32;
  for (uint64_t s = 0; s < n; s += RUN) { /* insertion sort of each run */
    uint64_t e = s + RUN < n ? s + RUN : n;
    for (uint64_t i = s + 1; i < e; ++i) {
      Item v = a[i];
      uint64_t j = i;
      while (j > s && a[j - 1].w > v.w) { a[j] = a[j - 1]; --j; }
      a[j] = v;
    }
  }
  Item *src = a, *dst = tmp;
  for (uint64_t width = RUN; width < n; width *= 2) {
    for (uint64_t lo = 0; lo < n; lo += 2 * width) {
      uint64_t mid = lo + width < n ? lo + width : n;
      uint64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
      uint64_t i = lo, j = mid, k = lo;
      while (i < mid && j < hi) dst[k++] = (src[j].w < src[i].w) ? src[j++] : src[i++];
      while (i < mid) dst[k++] = src[i++];
      while (j < hi) dst[k++] = src[j++];
    }
    Item* t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(Item) * n);
  free(tmp);
}

/* Stable sorting permutation: perm[k] = index of the k-th smallest key
 * (vector.rs:483-524 sort / sort_by_key; by_key.rs radix total order). */
void msvo_sort_perm(int dt, int mode, const void* keys, uint64_t n, uint32_t* perm) {
  const size_t es = dt_size(dt);
  Item* it = (Item*)malloc(sizeof(Item) * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) {
    it[i].w = msvo_order_word(dt, (const char*)keys + i * es, mode);
    it[i].i = (uint32_t)i;
  }
  stable_sort_items(it, n);
  for (uint64_t i = 0; i < n; ++i) perm[i] = it[i].i;
  free(it);
}

/* sort_by_key restated on pairs (vector.rs:515-524): sorts keys and values. */
void msvo_sort_by_key(int kdt, int vdt, int mode, const void* keys, const void* values,
                      uint64_t n, void* out_keys, void* out_values) {
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  msvo_sort_perm(kdt, mode, keys, n, perm);
  if (out_keys) msvo_gather(kdt, keys, perm, n, out_keys);
  if (out_values) msvo_gather(vdt, values, perm, n, out_values);
  free(perm);
}

/* merge (vector.rs:493-513): takes right only if right < left. src[k] is the
 * source of output k: index into left, or nl + index into right. */
void msvo_merge_src(int dt, int mode, const void* left, uint64_t nl, const void* right,
                    uint64_t nr, uint32_t* src) {
  const size_t es = dt_size(dt);
  uint64_t l = 0, r = 0, k = 0;
  while (l < nl && r < nr) {
    uint64_t wl = msvo_order_word(dt, (const char*)left + l * es, mode);
    uint64_t wr = msvo_order_word(dt, (const char*)right + r * es, mode);
    if (wr < wl) src[k++] = (uint32_t)(nl + r++);
    else src[k++] = (uint32_t)l++;
  }
  while (l < nl) src[k++] = (uint32_t)l++;
  while (r < nr) src[k++] = (uint32_t)(nl + r++);
}

/* set_union / set_intersection / set_difference (vector.rs:595-664). kind:
 * 0 union, 1 intersection, 2 difference. Returns the output length. */
uint64_t msvo_set_op(int dt, int mode, int kind, const void* left, uint64_t nl, const void* right,
                     uint64_t nr, void* out) {
  const size_t es = dt_size(dt);
  uint64_t l = 0, r = 0, k = 0;
#define EMIT(ptr) memcpy((char*)out + (k++) * es, (ptr), es)
  while (l < nl && r < nr) {
    const char* pl = (const char*)left + l * es;
    const char* pr = (const char*)right + r * es;
    uint64_t wl = msvo_order_word(dt, pl, mode), wr = msvo_order_word(dt, pr, mode);
    if (wl < wr) {
      if (kind != 1) EMIT(pl);
      ++l;
    } else if (wr < wl) {
      if (kind == 0) EMIT(pr);
      ++r;
    } else {
      if (kind != 2) EMIT(pl);
      ++l;
      ++r;
    }
  }
  if (kind != 1) while (l < nl) { EMIT((const char*)left + l * es); ++l; }
  if (kind == 0) while (r < nr) { EMIT((const char*)right + r * es); ++r; }
#undef EMIT
  return k;
}

/* The same walk (vector.rs:595-664) giving the source of each output: an
 * index into left, or nl + an index into right. Returns the output length. */
uint64_t msvo_set_op_src(int dt, int mode, int kind, const void* left, uint64_t nl,
                         const void* right, uint64_t nr, uint32_t* src) {
  const size_t es = dt_size(dt);
  uint64_t l = 0, r = 0, k = 0;
  while (l < nl && r < nr) {
    uint64_t wl = msvo_order_word(dt, (const char*)left + l * es, mode);
    uint64_t wr = msvo_order_word(dt, (const char*)right + r * es, mode);
    if (wl < wr) {
      if (kind != 1) src[k++] = (uint32_t)l;
      ++l;
    } else if (wr < wl) {
      if (kind == 0) src[k++] = (uint32_t)(nl + r);
      ++r;
    } else {
      if (kind != 2) src[k++] = (uint32_t)l;
      ++l;
      ++r;
    }
  }
  if (kind != 1) while (l < nl) src[k++] = (uint32_t)l++;
  if (kind == 0) while (r < nr) src[k++] = (uint32_t)(nl + r++);
  return k;
}

/* ---- seg (seg.rs:18-39, 48-88, 155-169) ------------------------------------- */
int msvo_seg_inclusive_scan(int dt, int op, const void* in, void* out, const uint32_t* offsets,
                            uint64_t nseg) {
  const size_t es = dt_size(dt);
  for (uint64_t s = 0; s < nseg; ++s) {
    uint64_t b = offsets[s], e = offsets[s + 1];
    if (msvo_inclusive_scan(dt, op, (const char*)in + b * es, (char*)out + b * es, e - b)) return -1;
  }
  return 0;
}

int msvo_seg_exclusive_scan(int dt, int op, const void* in, void* out, const uint32_t* offsets,
                            uint64_t nseg, const void* init) {
  const size_t es = dt_size(dt);
  for (uint64_t s = 0; s < nseg; ++s) {
    uint64_t b = offsets[s], e = offsets[s + 1];
    if (msvo_exclusive_scan(dt, op, (const char*)in + b * es, (char*)out + b * es, e - b, init))
      return -1;
  }
  return 0;
}

int msvo_seg_reduce(int dt, int op, const void* in, const uint32_t* offsets, uint64_t nseg,
                    const void* init, void* out) {
  const size_t es = dt_size(dt);
  for (uint64_t s = 0; s < nseg; ++s) {
    uint64_t b = offsets[s], e = offsets[s + 1];
    if (msvo_reduce(dt, op, (const char*)in + b * es, e - b, init, (char*)out + s * es)) return -1;
  }
  return 0;
}

/* ---- by-key family (verification/oracle/src/vector.rs:92-175,578-594) ----
 * Key equality is the built-in Equal: numeric == in the key dtype (so
 * -0.0 == +0.0 and a NaN key equals nothing). */
static int key_eq(int kdt, const void* keys, uint64_t i, uint64_t j) {
  const char* k = (const char*)keys;
  switch (kdt) {
    case DT_F32: {
      float a, b;
      memcpy(&a, k + i * 4, 4);
      memcpy(&b, k + j * 4, 4);
      return a == b;
    }
    case DT_F64: {
      double a, b;
      memcpy(&a, k + i * 8, 8);
      memcpy(&b, k + j * 8, 8);
      return a == b;
    }
    default: {
      const size_t es = dt_size(kdt);
      return memcmp(k + i * es, k + j * es, es) == 0;
    }
  }
}

/* End of the run that starts at b: the scans compare adjacent keys
 * (vector.rs:107, :131). */
static uint64_t adjacent_run_end(int kdt, const void* keys, uint64_t b, uint64_t n) {
  uint64_t e = b + 1;
  while (e < n && key_eq(kdt, keys, e - 1, e)) ++e;
  return e;
}

int msvo_inclusive_scan_by_key(int kdt, int vdt, int op, const void* keys, const void* values,
                               void* out, uint64_t n) {
  const size_t es = dt_size(vdt);
  for (uint64_t b = 0; b < n;) {
    const uint64_t e = adjacent_run_end(kdt, keys, b, n);
    if (msvo_inclusive_scan(vdt, op, (const char*)values + b * es, (char*)out + b * es, e - b))
      return -1;
    b = e;
  }
  return 0;
}

int msvo_exclusive_scan_by_key(int kdt, int vdt, int op, const void* keys, const void* values,
                               void* out, uint64_t n, const void* init) {
  const size_t es = dt_size(vdt);
  for (uint64_t b = 0; b < n;) {
    const uint64_t e = adjacent_run_end(kdt, keys, b, n);
    if (msvo_exclusive_scan(vdt, op, (const char*)values + b * es, (char*)out + b * es, e - b, init))
      return -1;
    b = e;
  }
  return 0;
}

/* reduce_by_key (vector.rs:144-175): a run continues while the key equals
 * the run's first key; the run's first element always belongs to it. Returns
 * the number of runs. */
int64_t msvo_reduce_by_key(int kdt, int vdt, int op, const void* keys, const void* values,
                           uint64_t n, const void* init, void* out_keys, void* out_values) {
  const size_t ks = dt_size(kdt), vs = dt_size(vdt);
  int64_t r = 0;
  for (uint64_t b = 0; b < n;) {
    uint64_t e = b + 1;
    while (e < n && key_eq(kdt, keys, b, e)) ++e;
    memcpy((char*)out_keys + r * ks, (const char*)keys + b * ks, ks);
    if (msvo_reduce(vdt, op, (const char*)values + b * vs, e - b, init, (char*)out_values + r * vs))
      return -1;
    ++r;
    b = e;
  }
  return r;
}

/* unique_by_key (vector.rs:578-594): the value of each run's first element. */
int64_t msvo_unique_by_key(int kdt, int vdt, const void* keys, const void* values, uint64_t n,
                           void* out_values) {
  const size_t vs = dt_size(vdt);
  int64_t r = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (i == 0 || !key_eq(kdt, keys, i - 1, i))
      memcpy((char*)out_values + (r++) * vs, (const char*)values + i * vs, vs);
  return r;
}

/* ---- partition (vector.rs:220-235) -----------------------------------------
 * Matching elements in input order, then failing ones in input order. Returns
 * the number of matching elements (the boundary). */
uint64_t msvo_partition(int dt, const void* in, uint64_t n, int kind, uint64_t a, uint64_t b,
                        void* out) {
  const size_t es = dt_size(dt);
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (pred_eval(dt, (const char*)in + i * es, kind, a, b))
      memcpy((char*)out + k++ * es, (const char*)in + i * es, es);
  uint64_t f = k;
  for (uint64_t i = 0; i < n; ++i)
    if (!pred_eval(dt, (const char*)in + i * es, kind, a, b))
      memcpy((char*)out + f++ * es, (const char*)in + i * es, es);
  return k;
}

/* ---- scatter_reduce (api/algorithm/scatter.rs:337-420; core/scatter_reduce.rs:
 * 130-145 combines as op(previous, proposal)) -------------------------------
 * The proposal for destination d is init (+) the values with index d folded
 * in input order; out[d] = op(out[d], proposal) for every d present. Indices
 * >= m are dropped (the GPU build never writes outside the output). */
#define SCATTER_REDUCE(T, SUF)                                                     \
  {                                                                                \
    const T* v = (const T*)values;                                                 \
    T* o = (T*)out;                                                                \
    T* acc = (T*)malloc((m ? m : 1) * sizeof(T));                                  \
    unsigned char* seen = (unsigned char*)calloc(m ? m : 1, 1);                    \
    if (!acc || !seen) { free(acc); free(seen); return -1; }                       \
    T iv;                                                                          \
    memcpy(&iv, init, sizeof(T));                                                  \
    for (uint64_t i = 0; i < n; ++i) {                                             \
      const uint32_t d = idx[i];                                                   \
      if (d >= m) continue;                                                        \
      if (!seen[d]) { acc[d] = iv; seen[d] = 1; }                                  \
      acc[d] = apply_##SUF(op, acc[d], v[i]);                                      \
    }                                                                              \
    for (uint64_t d = 0; d < m; ++d)                                               \
      if (seen[d]) o[d] = apply_##SUF(op, o[d], acc[d]);                           \
    free(acc);                                                                     \
    free(seen);                                                                    \
  }
int msvo_scatter_reduce(int dt, int op, const void* values, const uint32_t* idx, uint64_t n,
                        const void* init, void* out, uint64_t m) {
  if ((dt == DT_F32 || dt == DT_F64) && op == OP_XOR) return -1;
  DISPATCH(dt, SCATTER_REDUCE(uint32_t, u32), SCATTER_REDUCE(int32_t, i32),
           SCATTER_REDUCE(float, f32), SCATTER_REDUCE(uint64_t, u64),
           SCATTER_REDUCE(int64_t, i64), SCATTER_REDUCE(double, f64));
  return 0;
}

/* ---- compacting segment operations (seg.rs:109-153) ------------------------
 * form 0 Filter (keep the elements passing the predicate), 1 Unique (keep an
 * element unless it equals, numeric ==, the last element kept from its
 * segment), 2 Take (the first `take` elements of each segment). Writes the
 * kept values and the new offsets (nseg + 1: kept elements before each
 * segment); returns the number kept. */
static int num_equal(int dt, const void* a, const void* b) {
  switch (dt) {
    case DT_F32: { float x, y; memcpy(&x, a, 4); memcpy(&y, b, 4); return x == y; }
    case DT_F64: { double x, y; memcpy(&x, a, 8); memcpy(&y, b, 8); return x == y; }
    default: return memcmp(a, b, dt_size(dt)) == 0;
  }
}
uint64_t msvo_seg_compact(int dt, int form, const void* in, const uint32_t* offsets,
                          uint64_t nseg, int kind, uint64_t a, uint64_t b, uint64_t take,
                          void* out, uint32_t* out_offsets) {
  const size_t es = dt_size(dt);
  uint64_t k = 0;
  for (uint64_t s = 0; s < nseg; ++s) {
    out_offsets[s] = (uint32_t)k;
    const uint64_t first = k;
    for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i) {
      const char* x = (const char*)in + i * es;
      int keep;
      if (form == 0) keep = pred_eval(dt, x, kind, a, b);
      else if (form == 1) keep = k == first || !num_equal(dt, (char*)out + (k - 1) * es, x);
      else keep = i - offsets[s] < take;
      if (keep) memcpy((char*)out + k++ * es, x, es);
    }
  }
  out_offsets[nseg] = (uint32_t)k;
  return k;
}

/* Per-segment stable sorting permutation (global indices). */
void msvo_seg_sort_perm(int dt, int mode, const void* keys, const uint32_t* offsets,
                        uint64_t nseg, uint32_t* perm) {
  const size_t es = dt_size(dt);
  for (uint64_t s = 0; s < nseg; ++s) {
    uint64_t b = offsets[s], e = offsets[s + 1];
    msvo_sort_perm(dt, mode, (const char*)keys + b * es, e - b, perm + b);
    for (uint64_t i = b; i < e; ++i) perm[i] += (uint32_t)b;
  }
}

/* ---- seeded streams (massively/src/api/util/random.rs:12-22,331-396) ------- */
uint32_t msvo_pcg_hash32(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
  return (word >> 22) ^ word;
}

uint32_t msvo_seed_key(uint64_t seed) {
  return msvo_pcg_hash32((uint32_t)seed ^ msvo_pcg_hash32((uint32_t)(seed >> 32)));
}

void msvo_uniform_u32(uint32_t min, uint32_t max, uint64_t seed, uint64_t start, uint64_t n,
                      uint32_t* out) {
  const uint32_t key = msvo_seed_key(seed);
  const uint32_t span = max - min;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t v = msvo_pcg_hash32((uint32_t)(start + i) ^ key ^ 0u);
    out[i] = (span == 0xffffffffu) ? v : min + v % (span + 1u);
  }
}

void msvo_uniform_u64(uint64_t min, uint64_t max, uint64_t seed, uint64_t start, uint64_t n,
                      uint64_t* out) {
  const uint32_t key = msvo_seed_key(seed);
  const uint64_t span = max - min;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t idx = (uint32_t)(start + i);
    uint64_t lo = msvo_pcg_hash32(idx ^ key ^ 0u);
    uint64_t hi = msvo_pcg_hash32(idx ^ key ^ 0x9e3779b9u);
    uint64_t v = lo | (hi << 32);
    out[i] = (span == ~0ull) ? v : min + v % (span + 1ull);
  }
}
