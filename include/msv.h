/*
 * msv.h -- C ABI of the MI355X (gfx950) vector and segment primitive engine.
 *
 * Drop-in boundary for the `massively` crate's device-resident primitives.
 * Every function returns an msv_status; details of the last failure on the
 * calling thread are available from msv_last_error(). No C++ exception and
 * no GPU fault crosses this boundary for a rejected argument.
 *
 * Reference interfaces replaced (paths relative to the massively repository):
 *   Executor / DeviceVec / DeviceSlice   massively/src/core/runtime.rs:21-545
 *   Error                                massively/src/core/error.rs:5-13
 *   vector::*                            massively/src/api/algorithm (cited per function)
 *   seg::ForEachSegment / Segmentation   massively/src/seg/mod.rs, seg/segmentation.rs
 *
 * Lengths are MIndex (= uint32_t, massively/src/lib.rs:79-82). Flags (MFlag)
 * are uint32_t and true iff nonzero (massively/src/flag.rs:9-19).
 *
 * Threading: one executor owns one HIP stream; calls on one executor must be
 * externally serialised. Work is enqueued asynchronously; functions that
 * return a host value (msv_reduce, msv_to_host, msv_vec_len of a vector with
 * a device-resident length, segmentation constructors) synchronise.
 */
#ifndef MSV_H
#define MSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSV_ABI_VERSION 1u

/* Status codes; 1..7 map one-to-one onto massively::Error
 * (massively/src/core/error.rs:5-13). */
typedef enum msv_status {
  MSV_OK = 0,
  MSV_ERR_LENGTH_MISMATCH = 1,      /* LengthMismatch { left, right } */
  MSV_ERR_OUTPUT_TOO_SHORT = 2,     /* OutputTooShort { input, output } */
  MSV_ERR_LENGTH_TOO_LARGE = 3,     /* LengthTooLarge { len } */
  MSV_ERR_INVALID_SEGMENTATION = 4, /* InvalidSegmentation */
  MSV_ERR_UNBOUND_COLUMN = 5,       /* UnboundColumn */
  MSV_ERR_FOREIGN_EXECUTOR = 6,     /* ForeignExecutor */
  MSV_ERR_LAUNCH = 7,               /* Launch { message } (HIP runtime failure) */
  MSV_ERR_INVALID_ARGUMENT = 8,     /* null handle, bad dtype/op combination, bad range */
  MSV_ERR_UNSUPPORTED = 9,          /* valid request this build does not serve */
  MSV_ERR_OUT_OF_MEMORY = 10
} msv_status;

typedef struct msv_error {
  int32_t code;       /* msv_status */
  uint64_t left;      /* LengthMismatch.left / OutputTooShort.input / LengthTooLarge.len */
  uint64_t right;     /* LengthMismatch.right / OutputTooShort.output */
  char message[256];
} msv_error;

typedef enum msv_dtype {
  MSV_U32 = 0,
  MSV_I32 = 1,
  MSV_F32 = 2,
  MSV_U64 = 3,
  MSV_I64 = 4,
  MSV_F64 = 5
} msv_dtype;

/* Built-in ReductionOp set (massively/src/core/reduce.rs:35-66). XOR is
 * integer-only. Integer ADD wraps modulo 2^bits. MIN(a,b) = b<a ? b : a. */
typedef enum msv_op { MSV_OP_ADD = 0, MSV_OP_MIN = 1, MSV_OP_MAX = 2, MSV_OP_XOR = 3 } msv_op;

/* Built-in BinaryPredicateOp key comparisons (massively/src/core/ordering/mod.rs:29-65). */
typedef enum msv_cmp { MSV_CMP_LESS = 0, MSV_CMP_GREATER = 1 } msv_cmp;

/* Built-in PredicateOp set (massively/src/core/predicate.rs:40-52), evaluated
 * on an element x of the stencil column. Thresholds `a`/`b` hold the bit
 * pattern of a value of the stencil's dtype (zero-extended to 64 bits).
 *   NONZERO      x != 0          ZERO       x == 0
 *   LT/LE/GT/GE  x < a, x <= a, x > a, x >= a (numeric compare in the dtype)
 *   BITS_ANY     (bits(x) & a) != 0
 *   BITS_EQ      (bits(x) & a) == b                                       */
typedef enum msv_pred_kind {
  MSV_PRED_NONZERO = 0,
  MSV_PRED_ZERO = 1,
  MSV_PRED_LT = 2,
  MSV_PRED_LE = 3,
  MSV_PRED_GT = 4,
  MSV_PRED_GE = 5,
  MSV_PRED_BITS_ANY = 6,
  MSV_PRED_BITS_EQ = 7
} msv_pred_kind;

typedef struct msv_predicate {
  uint32_t kind; /* msv_pred_kind */
  uint32_t reserved;
  uint64_t a;
  uint64_t b;
} msv_predicate;

typedef struct msv_executor msv_executor; /* Executor<R> */
typedef struct msv_vec msv_vec;           /* DeviceVec<R,T> / MVec (refcounted) */
typedef struct msv_segmentation msv_segmentation; /* seg::Segmentation<R> */

/* DeviceSlice / DeviceSliceMut: a by-value view {vec, offset, len}. `len`
 * MSV_WHOLE means "to the end of the vector's logical length". */
#define MSV_WHOLE 0xffffffffu
typedef struct msv_slice {
  msv_vec* vec;
  uint32_t offset;
  uint32_t len;
} msv_slice;

/* ---- errors and versions --------------------------------------------- */
uint32_t msv_abi_version(void);
/* Copies the calling thread's last error into *out (code MSV_OK if none). */
void msv_last_error(msv_error* out);

/* ---- executor (massively/src/core/runtime.rs:21-197) ------------------- */
/* Executor::new: binds `device`, creates one stream and a caching allocator. */
msv_status msv_executor_create(int32_t device, msv_executor** out);
msv_status msv_executor_destroy(msv_executor* ex);
/* Executor::sync (runtime.rs:192-196). */
msv_status msv_executor_sync(msv_executor* ex);
/* Monotonic executor id used for ForeignExecutor checks (runtime.rs:12,49-54). */
uint64_t msv_executor_id(const msv_executor* ex);
/* The executor's HIP stream (hipStream_t), for callers that interleave their own work. */
void* msv_executor_stream(msv_executor* ex);
/* Returns cached device memory to the HIP runtime (synchronises). */
msv_status msv_executor_trim(msv_executor* ex);

/* Tuning options of the batching layer. No counterpart in the reference
 * (its launch shapes are fixed per kernel); results never depend on them.
 *  MSV_OPT_SEG_LAUNCH: how seg::Sort runs its work items: one persistent
 *    launch over a work queue (MSV_SEG_LAUNCH_PERSISTENT, default) or one
 *    launch per size class (MSV_SEG_LAUNCH_PER_CLASS).
 *  MSV_OPT_AUTOTUNE: 1 (default) times the tile-shape candidates on the first
 *    seg::Sort of each (key width, element width, mean-segment-length class),
 *    and the grid candidates of each batched job-kernel family on the first
 *    batch of each (family, job-count class, mean-length class), and keeps
 *    the fastest in the tuning table; 0 uses the built-in defaults.
 *  MSV_OPT_SEGSORT_CONFIG: -1 (default) = as above, else force candidate k
 *    (0 <= k < msv_segsort_config_count()); MSV_ERR_UNSUPPORTED if candidate
 *    k does not fit the element width.
 *  MSV_OPT_RADIX_ONESWEEP: the radix sorts' schedule: 1 = one histogram pass
 *    and one look-back pass per digit (onesweep); 0 = a counting upsweep, a
 *    scan of the counts and a downsweep per digit (reduce-then-scan); -1
 *    (default) = onesweep.
 *  MSV_OPT_SEGSORT_WAVE: 1 (default) seg::Sort sorts segments of up to 1024
 *    elements inside one wave each (bitonic network in registers), longer
 *    ones up to the tile size one block each with the tile shapes above;
 *    0 packs the short segments into tile-sized windows instead. Either way
 *    segments longer than max(tile, 4096) are sorted by many blocks each. */
typedef enum msv_option {
  MSV_OPT_SEG_LAUNCH = 1,
  MSV_OPT_AUTOTUNE = 2,
  MSV_OPT_SEGSORT_CONFIG = 3,
  MSV_OPT_RADIX_ONESWEEP = 4,
  MSV_OPT_SEGSORT_WAVE = 5
} msv_option;
#define MSV_SEG_LAUNCH_PERSISTENT 0
#define MSV_SEG_LAUNCH_PER_CLASS 1
msv_status msv_executor_set_option(msv_executor* ex, msv_option opt, int64_t value);
msv_status msv_executor_get_option(msv_executor* ex, msv_option opt, int64_t* value);
/* HIP graph capture of library calls (no counterpart in the reference): the
 * calls made on `ex` between begin and end are recorded, not run, and
 * msv_graph_launch replays them on the executor's stream. Each replay reads
 * the buffers the calls were given at capture time (write new inputs into
 * them between replays) and writes the same output vectors. Scratch memory
 * and lookback state of the captured calls belong to the graph (cleared by
 * kernel nodes on every replay, never by memset nodes). Calls that need a
 * host synchronisation (to_host, to_device, sync, trim, reduce with a host
 * result, Segmentation constructors, whole-length slices of vectors with a
 * device-resident length) fail with MSV_ERR_UNSUPPORTED while capturing. */
typedef struct msv_graph msv_graph;
msv_status msv_executor_begin_capture(msv_executor* ex);
msv_status msv_executor_end_capture(msv_executor* ex, msv_graph** out);
msv_status msv_graph_launch(msv_graph* g);
/* Waits for the executor's stream, then frees the graph and its memory. */
msv_status msv_graph_destroy(msv_graph* g);

/* Number of seg::Sort tile-shape candidates, and candidate k's shape. */
uint32_t msv_segsort_config_count(void);
msv_status msv_segsort_config_info(uint32_t k, uint32_t* block, uint32_t* items_per_thread,
                                   uint32_t* digit_bits);
/* The candidate the autotuner chose on the current device for this shape
 * (key/value bytes, launch form, total elements, segments; short segments
 * sorted inside a wave, or packed into windows for the _windows form), or -1
 * if it has not run for it. */
int32_t msv_segsort_tuned_config(uint32_t key_bytes, uint32_t value_bytes, int64_t launch,
                                 uint32_t n, uint32_t nseg);
int32_t msv_segsort_tuned_config_windows(uint32_t key_bytes, uint32_t value_bytes, int64_t launch,
                                         uint32_t n, uint32_t nseg);
/* The grid candidate the autotuner chose on the current device for a batched
 * job-kernel family (0 = scan / reduce / copy_where, 1..6 = sort_by_key by
 * (key, value) words: (1,0) (1,1) (1,2) (2,0) (2,1) (2,2)) in a batch of
 * `jobs` jobs of mean length `mean_len`, or -1 if it has not run for that
 * class. Candidate c launches the family's resident blocks per CU divided by
 * 2^c (at least one per CU). msv_run_jobs and the batched server tune a class
 * on its first batch of at least 16 jobs of that family. */
int32_t msv_batch_tuned_grid(uint32_t family, uint32_t jobs, uint32_t mean_len);
/* The autotuner's table (no counterpart in the reference, which specialises
 * kernels in its JIT compiler): one entry per (device signature "<gfx arch
 * name>/<compute units>", family, key), saved to and merged from a text file
 * so that a later process starts tuned; entries made on another kind of GPU
 * are kept but never used. With the environment variable MSV_TUNING_CACHE
 * naming a file, the library reads it at first use and rewrites it after
 * every new entry. */
msv_status msv_tuning_save(const char* path);
msv_status msv_tuning_load(const char* path);
msv_status msv_tuning_clear(void);
uint32_t msv_tuning_count(void);

/* ---- device vectors (runtime.rs:79-177, allocation.rs:852-861) -------- */
/* Executor::alloc: uninitialised vector of `len` elements. */
msv_status msv_alloc(msv_executor* ex, msv_dtype dtype, uint32_t len, msv_vec** out);
/* Executor::to_device: copies `len` host elements into a new vector. */
msv_status msv_to_device(msv_executor* ex, msv_dtype dtype, const void* host, uint32_t len,
                         msv_vec** out);
/* Executor::to_host: copies the slice (clipped to a device-resident logical
 * length, core/extent.rs:17-36) to `host` (capacity `host_cap` elements) and
 * writes the number of elements copied to *out_len. Synchronises. */
msv_status msv_to_host(msv_executor* ex, msv_slice src, void* host, uint32_t host_cap,
                       uint32_t* out_len);
/* Wraps caller-owned device memory (not freed by the library). */
msv_status msv_wrap_device(msv_executor* ex, msv_dtype dtype, void* device_ptr, uint32_t len,
                           msv_vec** out);
/* Shallow Clone (runtime.rs:275-285): shares the allocation. */
msv_status msv_vec_retain(msv_vec* v);
msv_status msv_vec_release(msv_vec* v);
/* Logical length; synchronises only if the length is device-resident. */
msv_status msv_vec_len(msv_executor* ex, msv_vec* v, uint32_t* out_len);
uint32_t msv_vec_capacity(const msv_vec* v);
int32_t msv_vec_dtype(const msv_vec* v);
void* msv_vec_device_ptr(const msv_vec* v);
/* Device address of the u32 logical length, or NULL when the length is host-known. */
void* msv_vec_device_len_ptr(const msv_vec* v);

/* ---- vector:: scans and reductions ----------------------------------- */
/* vector::inclusive_scan (api/algorithm/scan.rs:96-113; core/scan.rs:1125-1156). */
msv_status msv_inclusive_scan(msv_executor* ex, msv_slice input, msv_op op, msv_vec** out);
/* No counterpart in the reference: msv_inclusive_scan into the caller's
 * `output` (same dtype, at least len(input) elements, not overlapping the
 * input) that allocates nothing (no hipMalloc/hipFree, no allocating library
 * call), for callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation. Its look-back status words
 * and float tile partials live in a static 4 MiB array per GPU, so inputs
 * whose scan needs more are rejected with MSV_ERR_LENGTH_TOO_LARGE (u32/i32:
 * any length; u64/i64 and f32: up to about 2^31 elements; f64: up to about
 * 2^29). Results equal msv_inclusive_scan's bit for bit. Every call on one
 * GPU shares that array: calls that overlap on one GPU (on different
 * executors or streams) race and must be ordered by the caller. */
msv_status msv_inclusive_scan_noalloc(msv_executor* ex, msv_slice input, msv_op op,
                                      msv_slice output);
/* vector::exclusive_scan (api/algorithm/scan.rs:216-257; core/scan.rs:1209-1233).
 * `init` points to one host element of the input dtype. */
msv_status msv_exclusive_scan(msv_executor* ex, msv_slice input, const void* init, msv_op op,
                              msv_vec** out);
/* vector::reduce (api/algorithm/reduce.rs:32-50): result = op(init, fold(x)),
 * written to the host element `out_host`. Synchronises. */
msv_status msv_reduce(msv_executor* ex, msv_slice input, const void* init, msv_op op,
                      void* out_host);

/* ---- vector:: selection and movement -------------------------------- */
/* vector::copy_where / remove_where (api/algorithm/selection.rs:146-260): stable
 * compaction by a u32 stencil. The output has capacity len(input) and a
 * device-resident logical length (no host sync). */
msv_status msv_copy_where(msv_executor* ex, msv_slice input, msv_slice stencil, msv_vec** out);
msv_status msv_remove_where(msv_executor* ex, msv_slice input, msv_slice stencil, msv_vec** out);
/* copy_where with a built-in predicate evaluated on `stencil` (any dtype;
 * may be the input itself). `invert` != 0 gives remove_where semantics. */
msv_status msv_copy_where_pred(msv_executor* ex, msv_slice input, msv_slice stencil,
                               const msv_predicate* pred, int32_t invert, msv_vec** out);
/* vector::partition (api/algorithm/selection.rs:245-305; core/selection.rs:414-457):
 * stable partition, passing elements first, then failing ones, each in input
 * order. `*boundary` receives the number of passing elements (read back, as
 * the reference does). */
msv_status msv_partition(msv_executor* ex, msv_slice input, const msv_predicate* pred,
                         msv_vec** out, uint32_t* boundary);
/* vector::gather (api/algorithm/indexed.rs:157-208): out[i] = values[indices[i]]. */
msv_status msv_gather(msv_executor* ex, msv_slice values, msv_slice indices, msv_vec** out);
/* vector::scatter (api/algorithm/scatter.rs:225-265): output[indices[i]] = values[i].
 * Indices are unchecked by the reference; this build rejects none but never
 * writes outside `output` (out-of-range indices are dropped). */
msv_status msv_scatter(msv_executor* ex, msv_slice values, msv_slice indices, msv_slice output);
/* vector::scatter_reduce (api/algorithm/scatter.rs:337-420; core/scatter_reduce.rs):
 * output[d] = op(output[d], init (+) every values[i] with indices[i] == d), for
 * the destinations present in `indices`; others are unchanged. `init` must be
 * op's identity (the reference's contract). Out-of-range indices are dropped. */
msv_status msv_scatter_reduce(msv_executor* ex, msv_slice values, msv_slice indices,
                              const void* init, msv_op op, msv_slice output);

/* ---- vector:: ordering ---------------------------------------------- */
/* vector::radix_sort_by_key (api/algorithm/by_key.rs:252-325; core/radix.rs):
 * stable ascending sort in integer / IEEE total order. Returns the permuted
 * values; `out_keys` (nullable) also receives the sorted keys. */
msv_status msv_radix_sort_by_key(msv_executor* ex, msv_slice keys, msv_slice values,
                                 msv_vec** out_values, msv_vec** out_keys);
/* vector::radix_sort_by_key over zipped columns (core/radix.rs:499-531;
 * api/iter.rs:380-388; tests/vector/api/radix.rs:118-242): 1..3 key columns in
 * lexicographic order (leftmost primary; each column in integer / IEEE total
 * order), stable; 1..12 value columns, all reordered by the one permutation.
 * `out_values` receives nvalues new vectors. */
msv_status msv_radix_sort_by_key_columns(msv_executor* ex, const msv_slice* keys, uint32_t nkeys,
                                         const msv_slice* values, uint32_t nvalues,
                                         msv_vec** out_values);
/* Zipped rows of 1..12 columns of one length (4- and 8-byte dtypes mixed),
 * moved by one launch that reads the shared input once: the reference's
 * indexed copy over the leaves of a zipped read (core/indexed.rs:16-196) for
 * gather and scatter, and its selection copy over zipped leaves
 * (core/selection.rs:151-411) for copy_where / remove_where. Semantics per
 * column are those of msv_gather / msv_scatter / msv_copy_where(_pred); every
 * output of msv_copy_where_columns carries the same device-resident length.
 * `out` receives ncols new vectors; `outputs` are ncols slices of one length
 * (scatter outputs of 64 MiB or more per column are written column by column
 * in the binned form). */
msv_status msv_gather_columns(msv_executor* ex, const msv_slice* values, uint32_t ncols,
                              msv_slice indices, msv_vec** out);
msv_status msv_scatter_columns(msv_executor* ex, const msv_slice* values, uint32_t ncols,
                               msv_slice indices, const msv_slice* outputs);
/* pred == NULL: the stencil is u32 flags (kept where != 0); else the built-in
 * predicate applied to the stencil column. invert != 0: remove_where. */
msv_status msv_copy_where_columns(msv_executor* ex, const msv_slice* input, uint32_t ncols,
                                  msv_slice stencil, const msv_predicate* pred, int32_t invert,
                                  msv_vec** out);
/* vector::sort_by_key (api/algorithm/by_key.rs:193-250): stable sort by the
 * built-in comparison; float keys compare with IEEE `<` (-0 == +0). */
msv_status msv_sort_by_key(msv_executor* ex, msv_slice keys, msv_slice values, msv_cmp cmp,
                           msv_vec** out_values, msv_vec** out_keys);
/* vector::sort (api/algorithm/ordering.rs:64-76). */
msv_status msv_sort(msv_executor* ex, msv_slice input, msv_cmp cmp, msv_vec** out);

/* ---- vector:: merge and ordered set operations ------------------------ */
/* vector::merge (api/algorithm/merge.rs:109-190): stable, ties take left. */
msv_status msv_merge(msv_executor* ex, msv_slice left, msv_slice right, msv_cmp cmp,
                     msv_vec** out);
/* vector::merge_by_key (api/algorithm/merge.rs:195-282). */
msv_status msv_merge_by_key(msv_executor* ex, msv_slice left_keys, msv_slice left_values,
                            msv_slice right_keys, msv_slice right_values, msv_cmp cmp,
                            msv_vec** out_keys, msv_vec** out_values);
/* vector::set_union / set_intersection / set_difference
 * (api/algorithm/set.rs:44-195; multiset rules core/set.rs:18-46). Output
 * capacity: union L+R, others L; the logical length is device-resident. */
typedef enum msv_set_kind { MSV_SET_UNION = 0, MSV_SET_INTERSECTION = 1, MSV_SET_DIFFERENCE = 2 } msv_set_kind;
msv_status msv_set_op(msv_executor* ex, msv_set_kind kind, msv_slice left, msv_slice right,
                      msv_cmp cmp, msv_vec** out);

/* ---- seg:: ---------------------------------------------------------- */
/* Segmentation::from_offsets (seg/segmentation.rs:187-218): offsets must be
 * nonempty, start at 0 and be nondecreasing; the last offset is the value
 * count. Synchronises (validation result). */
msv_status msv_segmentation_from_offsets(msv_executor* ex, msv_slice offsets,
                                         msv_segmentation** out);
/* Segmentation::from_lengths (seg/segmentation.rs:224-264). */
msv_status msv_segmentation_from_lengths(msv_executor* ex, msv_slice lengths,
                                         msv_segmentation** out);
msv_status msv_segmentation_destroy(msv_segmentation* s);
/* Segmentation::offsets (shared, retained for the caller) / segment_count / value count. */
msv_status msv_segmentation_offsets(msv_segmentation* s, msv_vec** out);
uint32_t msv_segmentation_segment_count(const msv_segmentation* s);
uint32_t msv_segmentation_value_count(const msv_segmentation* s);

/* ForEachSegment(InclusiveScan(op)) (seg/mod.rs:458-572). */
msv_status msv_seg_inclusive_scan(msv_executor* ex, msv_slice values,
                                  const msv_segmentation* seg, msv_op op, msv_vec** out);
/* ForEachSegment(ExclusiveScan(op, init)): each segment starts from init. */
msv_status msv_seg_exclusive_scan(msv_executor* ex, msv_slice values,
                                  const msv_segmentation* seg, const void* init, msv_op op,
                                  msv_vec** out);
/* ForEachSegment(Reduce(op, init)) (seg/mod.rs:676-806,1314-1336): one value
 * per segment, init (+) fold(segment); empty segments yield init. */
msv_status msv_seg_reduce(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                          const void* init, msv_op op, msv_vec** out);
/* ForEachSegment(Sort(cmp)) (seg/mod.rs:379-425): stable sort inside every
 * segment by `keys`; `values` (nullable .vec) are permuted alongside. */
msv_status msv_seg_sort(msv_executor* ex, msv_slice keys, msv_slice values,
                        const msv_segmentation* seg, msv_cmp cmp, msv_vec** out_keys,
                        msv_vec** out_values);
/* The compacting segment operations (seg/mod.rs:597-640 Unique, :810-846
 * Filter, :847-880 Take, :1245-1310 owned results; seg/control.rs:555-618).
 * Each returns the kept values (an exact-length vector; the kept count is
 * read back, as the reference does) and a new segmentation with the same
 * segment count whose offsets index them. Filter keeps the elements passing
 * `pred`; Unique keeps each segment's first element and every element whose
 * value differs (numeric ==) from the previous one; Take keeps the first
 * `count` elements of every segment. */
msv_status msv_seg_filter(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                          const msv_predicate* pred, msv_vec** out_values,
                          msv_segmentation** out_seg);
msv_status msv_seg_unique(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                          msv_vec** out_values, msv_segmentation** out_seg);
msv_status msv_seg_take(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                        uint32_t count, msv_vec** out_values, msv_segmentation** out_seg);

/* ---- user-defined operators (compiled at run time) --------------------
 * The reference's algorithms take user operators as CubeCL traits compiled
 * into its kernels (ReductionOp, core/reduce.rs:63-66; PredicateOp,
 * core/predicate.rs:40-43; runtime-compiled read expressions,
 * core/eval.rs:21-107). Here a user operator is a C++ expression over the
 * element type, compiled once by msv_user_op_create (hiprtc, a fraction of a
 * second) into the library's own kernel templates:
 *   MSV_USER_REDUCTION: `expr` of `a` and `b` (associative; the scans and the
 *     reduce combine in index order, so it need not be commutative) and
 *     `identity`, an expression of the identity element;
 *   MSV_USER_PREDICATE: `expr` of `x`, nonzero = keep (`identity` unused);
 *   MSV_USER_LESS: a BinaryPredicateOp used as Less (core/ordering/mod.rs:53-56),
 *     `expr` of `a` and `b`, nonzero = a comes strictly before b (a strict
 *     weak ordering; `identity` unused);
 *   MSV_USER_EQUAL: a BinaryPredicateOp used as Equal, `expr` of `a` and `b`,
 *     nonzero = a and b are equal keys (`identity` unused).
 * The code runs on the GPU with the caller's privileges, as the reference's
 * user operators do. A compile error returns MSV_ERR_INVALID_ARGUMENT with
 * the compiler's first diagnostic in msv_last_error. An operator belongs to
 * the executor it was created on and to one dtype. */
typedef struct msv_user_op msv_user_op;
enum { MSV_USER_REDUCTION = 1, MSV_USER_PREDICATE = 2, MSV_USER_LESS = 4, MSV_USER_EQUAL = 5 };
msv_status msv_user_op_create(msv_executor* ex, int32_t kind, msv_dtype dtype, const char* expr,
                              const char* identity, msv_user_op** out);
msv_status msv_user_op_destroy(msv_user_op* op);
/* vector::reduce / inclusive_scan / exclusive_scan / copy_where with a user
 * operator (same outputs and synchronisation as the built-in forms; the scans
 * use the fixed-shape passes, so results are reproducible bit for bit). */
msv_status msv_reduce_user(msv_executor* ex, msv_slice input, const void* init,
                           const msv_user_op* op, void* out_host);
msv_status msv_inclusive_scan_user(msv_executor* ex, msv_slice input, const msv_user_op* op,
                                   msv_vec** out);
msv_status msv_exclusive_scan_user(msv_executor* ex, msv_slice input, const void* init,
                                   const msv_user_op* op, msv_vec** out);
msv_status msv_copy_where_user(msv_executor* ex, msv_slice input, const msv_user_op* pred,
                               msv_vec** out);
/* The ordering algorithms with a user Less (MSV_USER_LESS of the key dtype),
 * with the outputs and errors of their built-in forms. Every form is stable:
 * equal keys keep their input order, and merge and the set operations take
 * the left input first.
 *   sort (api/algorithm/ordering.rs:64-76), sort_by_key (by_key.rs:193-250;
 *     returns the permuted values, `out_keys` nullable);
 *   merge / merge_by_key (merge.rs:109-282; core/merge.rs:24-214);
 *   set operations (set.rs:44-195; core/set.rs:48-207): two keys are equal
 *     when neither is less than the other;
 *   ForEachSegment(Sort(less)) (seg/mod.rs:379-425): `values` nullable .vec. */
msv_status msv_sort_user(msv_executor* ex, msv_slice input, const msv_user_op* less,
                         msv_vec** out);
msv_status msv_sort_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                const msv_user_op* less, msv_vec** out_values,
                                msv_vec** out_keys);
msv_status msv_merge_user(msv_executor* ex, msv_slice left, msv_slice right,
                          const msv_user_op* less, msv_vec** out);
msv_status msv_merge_by_key_user(msv_executor* ex, msv_slice left_keys, msv_slice left_values,
                                 msv_slice right_keys, msv_slice right_values,
                                 const msv_user_op* less, msv_vec** out_keys,
                                 msv_vec** out_values);
msv_status msv_set_op_user(msv_executor* ex, msv_set_kind kind, msv_slice left, msv_slice right,
                           const msv_user_op* less, msv_vec** out);
msv_status msv_seg_sort_user(msv_executor* ex, msv_slice keys, msv_slice values,
                             const msv_segmentation* seg, const msv_user_op* less,
                             msv_vec** out_keys, msv_vec** out_values);
/* The by-key family and seg::Unique with a user Equal (MSV_USER_EQUAL of the
 * key dtype; by_key.rs:359-811, seg/mod.rs:597-640): a run is a maximal range
 * of adjacent keys with equal(k[i-1], k[i]); outputs and errors as in the
 * built-in forms. */
msv_status msv_inclusive_scan_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                          const msv_user_op* equal, msv_op op, msv_vec** out);
msv_status msv_exclusive_scan_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                          const msv_user_op* equal, const void* init,
                                          msv_op op, msv_vec** out);
msv_status msv_reduce_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                  const msv_user_op* equal, const void* init, msv_op op,
                                  msv_vec** out_keys, msv_vec** out_values);
msv_status msv_unique_by_key_user(msv_executor* ex, msv_slice keys, msv_slice values,
                                  const msv_user_op* equal, msv_vec** out_values);
msv_status msv_seg_unique_user(msv_executor* ex, msv_slice values, const msv_segmentation* seg,
                               const msv_user_op* equal, msv_vec** out_values,
                               msv_segmentation** out_seg);

/* ---- lazy inputs, fused into the consuming kernel (compiled at run time) ---
 * The reference's lazy inputs (api/lazy.rs:664-920: constant, counting, stride,
 * map, permute, reverse, repeat_each, tile, adjacent_difference, identity) are
 * read expressions over up to 13 device slots, compiled into the kernel that
 * consumes them (core/eval.rs:21-107, core/read.rs). Here a lazy input is
 * `read_src`, C++ source that defines, for the element type UT,
 *     __device__ UT msv_read(const LazyArgs& la, size_t i)
 * giving element i. It reads slot k with MSV_SLOT(k, type, index) (0 for an
 * index past the slot's length, so no read leaves a slot) and run-time
 * parameters la.q[0..15]; helper functions may precede it. msv_lazy_op_create
 * compiles it once together with the consuming operator: kind
 * MSV_USER_REDUCTION (a user expression and identity as for msv_user_op_create,
 * or expr NULL and builtin_op 0 add, 1 min, 2 max, 3 xor), MSV_USER_PREDICATE
 * (a user expression of `x`, or expr NULL for the built-in predicates, whose
 * msv_predicate is given per call) or MSV_LAZY_MAP (the lazy input written
 * out: vector::map, api/algorithm/transform.rs:91-108). The source does not
 * depend on the slots' addresses or the parameters' values, so one compiled
 * operator serves every call with the same expression shape. A reduction
 * compiled with a read source also accepts plain slices (msv_*_user). */
#define MSV_LAZY_SLOTS 13
#define MSV_LAZY_PARAMS 16
enum { MSV_LAZY_MAP = 3 };
typedef struct msv_lazy_args {
  msv_slice slots[MSV_LAZY_SLOTS]; /* owned slices of the executor; MSV_SLOT(k,...) reads slots[k] */
  uint32_t nslots;
  uint32_t len;                    /* logical length of the lazy input */
  uint64_t params[MSV_LAZY_PARAMS];
} msv_lazy_args;
msv_status msv_lazy_op_create(msv_executor* ex, int32_t kind, msv_dtype dtype,
                              const char* read_src, int32_t builtin_op, const char* expr,
                              const char* identity, msv_user_op** out);
/* Generates and compiles an operator as msv_lazy_op_create does, without a
 * device or an executor: MSV_OK, or the error msv_lazy_op_create would give. */
msv_status msv_lazy_op_check(int32_t kind, msv_dtype dtype, const char* read_src,
                             int32_t builtin_op, const char* expr, const char* identity);
/* vector::reduce / inclusive_scan / exclusive_scan (api/algorithm/reduce.rs:32,
 * scan.rs:96,216) of a lazy input; same outputs and synchronisation as the
 * msv_*_user forms. */
msv_status msv_reduce_lazy(msv_executor* ex, const msv_lazy_args* input, const void* init,
                           const msv_user_op* op, void* out_host);
msv_status msv_inclusive_scan_lazy(msv_executor* ex, const msv_lazy_args* input,
                                   const msv_user_op* op, msv_vec** out);
msv_status msv_exclusive_scan_lazy(msv_executor* ex, const msv_lazy_args* input, const void* init,
                                   const msv_user_op* op, msv_vec** out);
/* vector::copy_where with the predicate applied to the lazy input itself
 * (selection.rs:146-165, device-resident length); `builtin` is used when the
 * operator was created with expr NULL (NULL there means nonzero). */
msv_status msv_copy_where_lazy(msv_executor* ex, const msv_lazy_args* input,
                               const msv_user_op* pred, const msv_predicate* builtin,
                               msv_vec** out);
/* vector::map (api/algorithm/transform.rs:91-108): the lazy input, written out. */
msv_status msv_map_lazy(msv_executor* ex, const msv_lazy_args* input, const msv_user_op* op,
                        msv_vec** out);

/* ---- seeded synthetic inputs (massively/src/api/util/random.rs:12-22,331-396) */
/* uniform_u32(min, max, seed): pcg_hash32(i ^ seed_key(seed) ^ stream) mapped
 * to [min, max]; bit-identical to the reference generator. */
msv_status msv_random_uniform_u32(msv_executor* ex, uint32_t min, uint32_t max, uint64_t seed,
                                  msv_slice out);
msv_status msv_random_uniform_u64(msv_executor* ex, uint64_t min, uint64_t max, uint64_t seed,
                                  msv_slice out);
/* out[i] = start + i (u32). */
msv_status msv_iota_u32(msv_executor* ex, uint32_t start, msv_slice out);

/* ---- by-key family (massively/src/api/algorithm/by_key.rs:359-811) ----
 * Keys and values have equal lengths (else MSV_ERR_LENGTH_MISMATCH). The key
 * equality is the built-in Equal: numeric == in the key dtype (-0.0 == +0.0;
 * a NaN key equals nothing). A run is a maximal range of equal adjacent keys.
 *  inclusive_scan_by_key (by_key.rs:399-427): the scan restarts at each run;
 *  exclusive_scan_by_key (by_key.rs:497-546): each run starts from init;
 *  reduce_by_key (by_key.rs:618-682): per run, its first key and init (+) its
 *    values; both outputs carry a device-resident length (the run count);
 *  unique_by_key (by_key.rs:759-789): the first value of each run, with a
 *    device-resident length. */
msv_status msv_inclusive_scan_by_key(msv_executor* ex, msv_slice keys, msv_slice values, msv_op op,
                                     msv_vec** out);
msv_status msv_exclusive_scan_by_key(msv_executor* ex, msv_slice keys, msv_slice values,
                                     const void* init, msv_op op, msv_vec** out);
msv_status msv_reduce_by_key(msv_executor* ex, msv_slice keys, msv_slice values, const void* init,
                             msv_op op, msv_vec** out_keys, msv_vec** out_values);
msv_status msv_unique_by_key(msv_executor* ex, msv_slice keys, msv_slice values,
                             msv_vec** out_values);

/* ---- batched jobs (no counterpart in the reference) -------------------
 * For callers with many independent jobs (the serving loop below, or any
 * batch of small requests): every job's inputs and outputs are caller-owned
 * device memory. Jobs small enough for one workgroup (scan / reduce /
 * copy_where up to 2^17 elements, sort_by_key up to 2^15) run as one
 * persistent launch per job family; larger jobs run with the device-wide
 * kernels. All work is enqueued on the executor's stream (no host sync). A
 * job's result does not depend on which other jobs share the call. */
typedef enum msv_job_kind {
  MSV_JOB_INCLUSIVE_SCAN = 1, /* out[n] = inclusive_scan(in, op) */
  MSV_JOB_EXCLUSIVE_SCAN = 2, /* out[n] = exclusive_scan(in, init, op) */
  MSV_JOB_REDUCE = 3,         /* out[1] = reduce(in, init, op) */
  MSV_JOB_COPY_WHERE = 4,     /* out[<= n] = in where pred(in) != invert; *(uint32_t*)out2 = count */
  MSV_JOB_SORT_BY_KEY = 5     /* out2[n] = sorted keys, out[n] = values moved with them */
} msv_job_kind;
/* Key order of a sort job: radix total order (radix_sort_by_key), or the
 * Less / Greater comparators (sort_by_key). */
#define MSV_ORDER_TOTAL 0u
#define MSV_ORDER_LESS 1u
#define MSV_ORDER_GREATER 2u
typedef struct msv_job {
  uint32_t kind;        /* msv_job_kind */
  uint32_t dtype;       /* element dtype (key dtype for sort jobs) */
  uint32_t op;          /* msv_op for scans and reduce; MSV_ORDER_* for sorts */
  uint32_t n;           /* input length */
  uint32_t value_bytes; /* sort: 0 (keys only), 4 or 8 */
  uint32_t invert;      /* copy_where: non-zero keeps the elements failing pred */
  msv_predicate pred;   /* copy_where predicate on the input elements */
  uint64_t init;        /* exclusive scan / reduce init (bit pattern of dtype) */
  const void* in;       /* input (keys for sorts) */
  const void* in2;      /* sort: values, or null */
  void* out;            /* scan / copy_where: n elements; reduce: 1; sort: values */
  void* out2;           /* sort: sorted keys; copy_where: uint32_t count */
} msv_job;
msv_status msv_run_jobs(msv_executor* ex, const msv_job* jobs, uint32_t count);

/* ---- serving loop (no counterpart in the reference) -------------------
 * A request loop on its own host thread: requests are admitted in arrival
 * order while the device memory they need fits the HBM budget, their inputs
 * are created on the device (from a seed, or copied from host memory), and
 * admitted requests are batched continuously: every batch goes to the next
 * free stream as soon as one is free, so requests that arrive while others
 * run overlap with them on the device. Outputs stay device-resident until
 * fetched and released. Small requests are not held behind large ones: at
 * most half of the streams run large requests (those too big for the batched
 * kernels) at once, and small requests behind a large head that cannot start
 * yet are admitted past it (up to 1024 positions). Two forms the batched mode
 * is measured against: MSV_SERVE_SERIAL runs one request at a time through
 * the device-wide kernels (one launch sequence per job); MSV_SERVE_PER_JOB
 * runs one request per launch sequence through the device-wide kernels on
 * every stream, in arrival order. Results are the same in all modes for integer
 * and sort jobs, and the same for one request alone or in any mix within a
 * mode (floating-point scan/reduce of jobs small enough for the batched
 * kernels follow that kernel's fixed order). */
typedef struct msv_server msv_server;
#define MSV_SERVE_BATCHED 0u
#define MSV_SERVE_SERIAL 1u
#define MSV_SERVE_PER_JOB 2u
typedef struct msv_server_config {
  int32_t device;
  uint32_t streams;    /* streams (executors) batches rotate over, 1..16 */
  uint32_t max_batch;  /* requests per batch, >= 1 */
  uint32_t mode;       /* MSV_SERVE_BATCHED, MSV_SERVE_SERIAL or MSV_SERVE_PER_JOB */
  uint64_t hbm_budget; /* bytes of device memory admitted requests may hold */
} msv_server_config;
typedef struct msv_request {
  uint32_t kind;        /* msv_job_kind */
  uint32_t dtype;       /* element / key dtype */
  uint32_t op;          /* as msv_job */
  uint32_t n;
  uint32_t value_bytes; /* sort: 0, or 4 for a u32 payload = input position */
  uint32_t invert;      /* copy_where */
  msv_predicate pred;   /* copy_where */
  uint64_t init;        /* exclusive scan / reduce */
  uint64_t seed;        /* input = the library generator's uniform stream for dtype */
  const void* host_input; /* if non-null: n elements copied in instead */
} msv_request;
typedef struct msv_request_info {
  uint32_t state;        /* 0 queued, 1 running, 2 done, 3 failed */
  msv_status status;
  double submit_us;      /* server clock, microseconds */
  double admit_us;
  double done_us;
  uint32_t out_len;      /* elements in the output (copy_where: the count) */
  uint32_t reserved;
  uint64_t bytes;        /* device bytes the request held while it ran */
} msv_request_info;
msv_status msv_server_create(const msv_server_config* cfg, msv_server** out);
/* Waits for every admitted request, then frees everything. */
msv_status msv_server_destroy(msv_server* s);
msv_status msv_server_submit(msv_server* s, const msv_request* r, uint64_t* ticket);
/* Blocks until the request is done (or failed) and fills *info. */
msv_status msv_server_wait(msv_server* s, uint64_t ticket, msv_request_info* info);
/* Copies a finished request's outputs to host memory: `out` gets the main
 * output (scan / copy_where elements, the reduce value, sorted keys) and
 * `out2` (may be null) the sorted values of a sort. */
msv_status msv_server_result(msv_server* s, uint64_t ticket, void* out, uint64_t out_bytes,
                             void* out2, uint64_t out2_bytes);
/* Frees a finished request's outputs (returns its bytes to the budget). */
msv_status msv_server_release(msv_server* s, uint64_t ticket);
/* Waits until every submitted request is done. */
msv_status msv_server_drain(msv_server* s);
/* Server clock (microseconds since creation), the clock of the timestamps. */
double msv_server_now_us(msv_server* s);

#ifdef __cplusplus
}
#endif

#endif /* MSV_H */
