// Device-side building blocks shared by every gfx950 kernel of the library:
// element traits, the closed operator set, wave64 DPP scans, and the
// decoupled-lookback tile-status protocol.
//
// Wave64 scans use DPP row shifts inside each 16-lane row followed by the two
// row broadcasts (lane 15 -> rows 1/3, lane 31 -> rows 2/3). Lanes whose DPP
// source is outside the row are not written, so they keep the operator's
// identity passed as the `old` operand; this makes the same code correct for
// add, min, max and xor.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>

namespace msv {

constexpr int kWave = 64;

// Closed operator set (mirrors the built-in ReductionOp implementations a
// massively user writes: Add, Min, Max, Xor).
enum OpKind : int { OP_ADD = 0, OP_MIN = 1, OP_MAX = 2, OP_XOR = 3 };

template <typename T>
struct Bits;
template <>
struct Bits<uint32_t> { using U = uint32_t; };
template <>
struct Bits<int32_t> { using U = uint32_t; };
template <>
struct Bits<float> { using U = uint32_t; };
template <>
struct Bits<uint64_t> { using U = uint64_t; };
template <>
struct Bits<int64_t> { using U = uint64_t; };
template <>
struct Bits<double> { using U = uint64_t; };

template <typename To, typename From>
__device__ __host__ inline To bit_cast(From v) {
  static_assert(sizeof(To) == sizeof(From), "size");
  To out;
  __builtin_memcpy(&out, &v, sizeof(To));
  return out;
}

template <typename T>
struct IsFloat { static constexpr bool value = false; };
template <>
struct IsFloat<float> { static constexpr bool value = true; };
template <>
struct IsFloat<double> { static constexpr bool value = true; };

template <typename T>
struct Limits;
template <>
struct Limits<uint32_t> {
  __host__ __device__ static uint32_t lowest() { return 0u; }
  __host__ __device__ static uint32_t highest() { return 0xffffffffu; }
};
template <>
struct Limits<int32_t> {
  __host__ __device__ static int32_t lowest() { return INT32_MIN; }
  __host__ __device__ static int32_t highest() { return INT32_MAX; }
};
template <>
struct Limits<uint64_t> {
  __host__ __device__ static uint64_t lowest() { return 0ull; }
  __host__ __device__ static uint64_t highest() { return ~0ull; }
};
template <>
struct Limits<int64_t> {
  __host__ __device__ static int64_t lowest() { return INT64_MIN; }
  __host__ __device__ static int64_t highest() { return INT64_MAX; }
};
template <>
struct Limits<float> {
  __host__ __device__ static float lowest() { return -__builtin_huge_valf(); }
  __host__ __device__ static float highest() { return __builtin_huge_valf(); }
};
template <>
struct Limits<double> {
  __host__ __device__ static double lowest() { return -__builtin_huge_val(); }
  __host__ __device__ static double highest() { return __builtin_huge_val(); }
};

// Operator application. `a` is the earlier (left) operand. Integer addition
// wraps modulo 2^bits (signed types add through their unsigned twin). Float
// addition's identity is -0.0 so that padding lanes never change the sign of
// a zero result. Min/max keep the left operand on ties: min(a,b) = b<a ? b : a.
template <typename T, int OP>
struct Combine {
  __host__ __device__ static inline T apply(T a, T b) {
    if constexpr (OP == OP_ADD) {
      if constexpr (IsFloat<T>::value) {
        return a + b;
      } else {
        using U = typename Bits<T>::U;
        return (T)((U)a + (U)b);
      }
    } else if constexpr (OP == OP_MIN) {
      return (b < a) ? b : a;
    } else if constexpr (OP == OP_MAX) {
      return (a < b) ? b : a;
    } else {
      static_assert(!IsFloat<T>::value, "xor is integer-only");
      return a ^ b;
    }
  }
  __host__ __device__ static inline T identity() {
    if constexpr (OP == OP_ADD) {
      if constexpr (IsFloat<T>::value) {
        return (T)-0.0;
      } else {
        return (T)0;
      }
    } else if constexpr (OP == OP_MIN) {
      return Limits<T>::highest();
    } else if constexpr (OP == OP_MAX) {
      return Limits<T>::lowest();
    } else {
      return (T)0;
    }
  }
};

// ---------------------------------------------------------------------------
// DPP moves for 32- and 64-bit values.
// ---------------------------------------------------------------------------
enum DppCtrl : int {
  DPP_ROW_SHR1 = 0x111,
  DPP_ROW_SHR2 = 0x112,
  DPP_ROW_SHR4 = 0x114,
  DPP_ROW_SHR8 = 0x118,
  DPP_ROW_BCAST15 = 0x142,
  DPP_ROW_BCAST31 = 0x143,
};

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ inline uint32_t dpp_u32(uint32_t old, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, ROW_MASK, BANK_MASK,
                                               false);
}

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf, typename T>
__device__ inline T dpp_move(T old, T src) {
  if constexpr (sizeof(T) == 4) {
    return bit_cast<T>(dpp_u32<CTRL, ROW_MASK, BANK_MASK>(bit_cast<uint32_t>(old),
                                                          bit_cast<uint32_t>(src)));
  } else {
    static_assert(sizeof(T) == 8, "dpp_move: 4 or 8 byte values");
    uint64_t o = bit_cast<uint64_t>(old), s = bit_cast<uint64_t>(src);
    uint32_t lo = dpp_u32<CTRL, ROW_MASK, BANK_MASK>((uint32_t)o, (uint32_t)s);
    uint32_t hi = dpp_u32<CTRL, ROW_MASK, BANK_MASK>((uint32_t)(o >> 32), (uint32_t)(s >> 32));
    return bit_cast<T>(((uint64_t)hi << 32) | lo);
  }
}

template <typename T>
__device__ inline T readlane(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return bit_cast<T>((uint32_t)__builtin_amdgcn_readlane((int)bit_cast<uint32_t>(v), lane));
  } else {
    uint64_t u = bit_cast<uint64_t>(v);
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return bit_cast<T>(((uint64_t)hi << 32) | lo);
  }
}

// Lane l receives the value of lane l-1; lane 0 receives `fill`.
template <typename T>
__device__ inline T shift_up1(T v, T fill) {
  const int lane = (int)__lane_id();
  const int src = (lane == 0 ? 0 : lane - 1) << 2;
  T out;
  if constexpr (sizeof(T) == 4) {
    out = bit_cast<T>((uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)bit_cast<uint32_t>(v)));
  } else {
    uint64_t u = bit_cast<uint64_t>(v);
    uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)u);
    uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(u >> 32));
    out = bit_cast<T>(((uint64_t)hi << 32) | lo);
  }
  return lane == 0 ? fill : out;
}

// Lane l receives the value of lane l+1; lane 63 receives `fill`.
template <typename T>
__device__ inline T shift_down1(T v, T fill) {
  const int lane = (int)__lane_id();
  const int src = (lane == 63 ? 63 : lane + 1) << 2;
  T out;
  if constexpr (sizeof(T) == 4) {
    out = bit_cast<T>((uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)bit_cast<uint32_t>(v)));
  } else {
    uint64_t u = bit_cast<uint64_t>(v);
    uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)u);
    uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(u >> 32));
    out = bit_cast<T>(((uint64_t)hi << 32) | lo);
  }
  return lane == 63 ? fill : out;
}

// A non-temporal 16-byte load (input streamed once).
__device__ inline uint4 load16_nt(const void* p) {
  typedef unsigned int u4_t __attribute__((ext_vector_type(4)));
  const u4_t v = __builtin_nontemporal_load(reinterpret_cast<const u4_t*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Four consecutive elements from p + i: one or two 16-byte loads when `full`
// (the caller has checked alignment and range), else guarded scalar loads (0
// past n). NT: non-temporal loads, for a kernel that streams one input once
// (with two interleaved streams, input and stencil, they measured slower).
template <typename T, bool NT = false>
__device__ inline void load4(const T* __restrict__ p, size_t i, uint32_t n, bool full, T (&v)[4]) {
  if (full) {
    if constexpr (sizeof(T) == 4) {
      const uint4 q = NT ? load16_nt(p + i) : *reinterpret_cast<const uint4*>(p + i);
      v[0] = bit_cast<T>(q.x); v[1] = bit_cast<T>(q.y); v[2] = bit_cast<T>(q.z); v[3] = bit_cast<T>(q.w);
    } else {
      const uint4 q0 = NT ? load16_nt(p + i) : *reinterpret_cast<const uint4*>(p + i);
      const uint4 q1 = NT ? load16_nt(p + i + 2) : *reinterpret_cast<const uint4*>(p + i + 2);
      v[0] = bit_cast<T>(((uint64_t)q0.y << 32) | q0.x);
      v[1] = bit_cast<T>(((uint64_t)q0.w << 32) | q0.z);
      v[2] = bit_cast<T>(((uint64_t)q1.y << 32) | q1.x);
      v[3] = bit_cast<T>(((uint64_t)q1.w << 32) | q1.z);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = (i + k < n) ? p[i + k] : (T)0;
  }
}

// Kernels read their input through a "source": a pointer (the element at i is
// p[i]) or a lazy source, an object whose operator()(i) computes element i
// (inputs composed of counting / constant / mapped / permuted reads, compiled
// at run time). src_at and the non-pointer load4 serve the lazy form.
// Lazy sources derive from this empty tag (no effect on their layout); a
// pointer, restrict-qualified or not, is not one.
struct LazySource {};
template <typename Src>
constexpr bool is_lazy_src = __is_base_of(LazySource, Src);

// Whether a lazy source can be read as unchecked vectors over [lo, hi) (its
// slots in range and aligned); a pointer source always can (the caller checks
// its alignment). One test per tile keeps the tile's row loads branch-free.
template <typename Src>
__device__ inline bool src_tile_ok(const Src& s, size_t lo, size_t hi) {
  if constexpr (is_lazy_src<Src>) return s.tile_ok(lo, hi);
  else return true;
}

template <typename Src>
__device__ inline auto src_at(const Src& s, size_t i) {
  if constexpr (is_lazy_src<Src>) return s(i);
  else return s[i];
}

// A lazy source also evaluates K consecutive elements at once (s.load<K>),
// which turns its slot reads into vector loads where the range allows; used
// when all four elements are in range (`full`).
template <typename T, bool NT = false, typename Src, typename = std::enable_if_t<is_lazy_src<Src>>>
__device__ inline void load4(const Src& s, size_t i, uint32_t n, bool full, T (&v)[4]) {
  if (full) {  // the caller has checked src_tile_ok over the tile
    s.template load_fast<4>(i, v);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = (i + k < n) ? (T)s(i + k) : (T)0;
  }
}

// K consecutive elements of a device slot of n elements from index i: one
// 8- or 16-byte load when they are all in range and the address is aligned to
// their size, else guarded element loads (0 past n). Lazy read sources use it
// for their slot leaves.
template <int K, typename D>
__device__ inline void slot_read(const void* base, uint32_t n, size_t i, D (&o)[K]) {
  const D* p = (const D*)base;
  constexpr size_t B = K * sizeof(D);
  if constexpr (B == 16) {
    if (i + K <= n && ((uintptr_t)(p + i) & 15u) == 0) {
      const uint4 q = *reinterpret_cast<const uint4*>(p + i);
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if constexpr (sizeof(D) == 4) o[k] = bit_cast<D>(w[k]);
        else o[k] = bit_cast<D>(((uint64_t)w[2 * k + 1] << 32) | w[2 * k]);
      }
      return;
    }
  } else if constexpr (B == 8) {
    if (i + K <= n && ((uintptr_t)(p + i) & 7u) == 0) {
      const uint2 q = *reinterpret_cast<const uint2*>(p + i);
      if constexpr (sizeof(D) == 4) {
        o[0] = bit_cast<D>(q.x);
        o[1] = bit_cast<D>(q.y);
      } else {
        o[0] = bit_cast<D>(((uint64_t)q.y << 32) | q.x);
      }
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) o[k] = (i + k < n) ? p[i + k] : (D)0;
}

// The same without checks: the caller knows the K elements are in range and
// p + i is aligned to 16 bytes (16-byte groups) or to the element. The
// 16-byte loads are non-temporal: a lazy slot is streamed once per pass.
template <int K, typename D>
__device__ inline void slot_read_fast(const void* base, size_t i, D (&o)[K]) {
  typedef unsigned int u4_t __attribute__((ext_vector_type(4)));
  const D* p = (const D*)base + i;
  constexpr size_t B = K * sizeof(D);
  if constexpr (B % 16 == 0) {
#pragma unroll
    for (int g = 0; g < (int)(B / 16); ++g) {
      const u4_t v = __builtin_nontemporal_load(reinterpret_cast<const u4_t*>(p) + g);
      const uint4 q = make_uint4(v.x, v.y, v.z, v.w);
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
      constexpr int E = 16 / sizeof(D);
#pragma unroll
      for (int k = 0; k < E; ++k) {
        if constexpr (sizeof(D) == 4) o[g * E + k] = bit_cast<D>(w[k]);
        else o[g * E + k] = bit_cast<D>(((uint64_t)w[2 * k + 1] << 32) | w[2 * k]);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) o[k] = p[k];
  }
}

// The by-key family's built-in Equal: numeric == on the key dtype (bitwise
// for integers; for floats -0.0 equals +0.0 and a NaN equals nothing).
template <typename K>
__device__ inline bool key_equal(K a, K b) {
  return a == b;
}

// Inclusive wave64 scan (Hillis-Steele over DPP). Result lane l holds
// x_0 (+) ... (+) x_l.
template <typename T, int OP>
__device__ inline T wave_inclusive_scan(T x) {
  using C = Combine<T, OP>;
  const T id = C::identity();
  x = C::apply(dpp_move<DPP_ROW_SHR1>(id, x), x);
  x = C::apply(dpp_move<DPP_ROW_SHR2>(id, x), x);
  x = C::apply(dpp_move<DPP_ROW_SHR4>(id, x), x);
  x = C::apply(dpp_move<DPP_ROW_SHR8>(id, x), x);
  x = C::apply(dpp_move<DPP_ROW_BCAST15, 0xa>(id, x), x);
  x = C::apply(dpp_move<DPP_ROW_BCAST31, 0xc>(id, x), x);
  return x;
}

// Wave total, uniform across the wave.
template <typename T, int OP>
__device__ inline T wave_reduce(T x) {
  return readlane(wave_inclusive_scan<T, OP>(x), kWave - 1);
}

// Segmented (head-flag) wave scan: pair (flag, value), combine
// (f1,v1).(f2,v2) = (f1|f2, f2 ? v2 : v1(+)v2).
template <typename T, int OP>
__device__ inline void wave_inclusive_seg_scan(uint32_t& f, T& x) {
  using C = Combine<T, OP>;
  const T id = C::identity();
#define MSV_SEG_STEP(CTRL, RM)                                  \
  {                                                             \
    T pv = dpp_move<CTRL, RM>(id, x);                           \
    uint32_t pf = dpp_move<CTRL, RM>(0u, f);                    \
    x = f ? x : C::apply(pv, x);                                \
    f = f | pf;                                                 \
  }
  MSV_SEG_STEP(DPP_ROW_SHR1, 0xf)
  MSV_SEG_STEP(DPP_ROW_SHR2, 0xf)
  MSV_SEG_STEP(DPP_ROW_SHR4, 0xf)
  MSV_SEG_STEP(DPP_ROW_SHR8, 0xf)
  MSV_SEG_STEP(DPP_ROW_BCAST15, 0xa)
  MSV_SEG_STEP(DPP_ROW_BCAST31, 0xc)
#undef MSV_SEG_STEP
}

__device__ inline uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// ---------------------------------------------------------------------------
// Scoped atomics. Agent scope makes the accesses coherent across the eight
// XCDs (each with its own L2) of one GPU.
// ---------------------------------------------------------------------------
__device__ inline uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t add_agent(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Decoupled lookback.
//
// Tiles are numbered by blockIdx.x. Every launch carries a generation number
// chosen by the host (29 bits, never reused before the status array is
// cleared), so status words left by earlier launches are simply invalid and
// the array never needs clearing between launches.
//
// Status word (64 bit): [63:35] generation | [34] head bit | [33:32] flag
// (1 = aggregate, 2 = inclusive prefix) | [31:0] payload. 64-bit payloads use
// two words per tile (low, high half) that each carry the flag; a reader
// accepts them only when both are valid with equal flags. All status traffic
// is relaxed agent-scope atomics, coherent across the eight XCDs, so no
// release/acquire fence (and no L2 write-back) is needed.
//
// Forward progress never depends on the order in which workgroups are
// dispatched: a wave that has polled a predecessor for kSelfHelpPolls rounds
// without seeing its status computes that tile's aggregate itself from the
// input (self-help) and continues.
// ---------------------------------------------------------------------------
constexpr uint32_t kFlagAgg = 1;
constexpr uint32_t kFlagInc = 2;
constexpr uint32_t kGenMask = (1u << 29) - 1;
constexpr uint32_t kSelfHelpPolls = 1024;

__device__ inline uint64_t pack_status(uint32_t gen, uint32_t flag, uint32_t head,
                                       uint32_t payload) {
  return ((uint64_t)((gen << 3) | (head << 2) | flag) << 32) | payload;
}
__device__ inline uint32_t status_flag(uint64_t w, uint32_t gen) {
  const uint32_t hi = (uint32_t)(w >> 32);
  return ((hi >> 3) == gen) ? (hi & 3u) : 0u;
}

// Per-tile status for a payload of type T (4 or 8 bytes) plus an optional
// head bit used by segmented scans.
template <typename T>
struct TileStatus {
  static constexpr int kWords = sizeof(T) / 4;
  uint64_t* words;
  uint32_t gen;

  __device__ inline void publish(uint32_t tile, uint32_t flag, uint32_t head, T v) const {
    if constexpr (kWords == 1) {
      st_agent(words + tile, pack_status(gen, flag, head, bit_cast<uint32_t>(v)));
    } else {
      const uint64_t u = bit_cast<uint64_t>(v);
      st_agent(words + 2 * (size_t)tile, pack_status(gen, flag, head, (uint32_t)u));
      st_agent(words + 2 * (size_t)tile + 1, pack_status(gen, flag, head, (uint32_t)(u >> 32)));
    }
  }
  // One poll: returns 0 if the status of `tile` is not yet valid.
  __device__ inline uint32_t try_read(uint32_t tile, uint32_t& head, T& v) const {
    if constexpr (kWords == 1) {
      const uint64_t w = ld_agent(words + tile);
      const uint32_t f = status_flag(w, gen);
      head = (uint32_t)(w >> 34) & 1u;
      v = bit_cast<T>((uint32_t)w);
      return f;
    } else {
      const uint64_t lo = ld_agent(words + 2 * (size_t)tile);
      const uint64_t hi = ld_agent(words + 2 * (size_t)tile + 1);
      const uint32_t fl = status_flag(lo, gen), fh = status_flag(hi, gen);
      head = (uint32_t)(lo >> 34) & 1u;
      v = bit_cast<T>(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
      return (fl != 0 && fl == fh) ? fl : 0u;
    }
  }
};

// Lookback with self-help: `agg_of(q)` is called by the whole wave with a
// uniform tile index q and returns that tile's aggregate (uniform). It is used
// only for predecessors whose status stays invalid for kSelfHelpPolls polls.
template <typename T, int OP, typename AggFn>
__device__ inline T wave_lookback_sh(const TileStatus<T>& st, uint32_t tile, AggFn agg_of) {
  using C = Combine<T, OP>;
  const uint32_t lane = __lane_id();
  T prefix = C::identity();
  int64_t base = (int64_t)tile - 1;
  for (;;) {
    const int64_t p = base - (int64_t)lane;
    T v = C::identity();
    uint32_t head = 0;
    uint32_t flag = p >= 0 ? 0u : kFlagInc;
    for (uint32_t polls = 0;; ++polls) {
      if (flag == 0) flag = st.try_read((uint32_t)p, head, v);
      const uint64_t pending = ballot(flag == 0);
      if (!pending) break;
      if (polls >= kSelfHelpPolls) {
        for (uint64_t m = pending; m; m &= m - 1) {
          const uint32_t k = (uint32_t)__builtin_ctzll(m);
          const T a = agg_of((uint32_t)(base - (int64_t)k));
          if (lane == k) {
            v = a;
            flag = kFlagAgg;
          }
        }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    const uint64_t inc_mask = ballot(flag == kFlagInc);
    if (inc_mask) {
      const uint32_t first = (uint32_t)__builtin_ctzll(inc_mask);
      const T contrib = lane <= first ? v : C::identity();
      return C::apply(wave_reduce<T, OP>(contrib), prefix);
    }
    prefix = C::apply(wave_reduce<T, OP>(v), prefix);
    base -= kWave;
  }
}

// Segmented lookback with self-help: the walk stops at the nearest tile that
// is inclusive or that contains a segment head (its aggregate then already is
// the carry). A predecessor whose status stays invalid for kSelfHelpPolls
// polls is computed by the whole wave with `agg_of(q, head, value)` (uniform
// q), as in wave_lookback_sh, so no tile waits on a tile that is not resident.
// Segmented combination is not commutative: each window is folded in order
// from the farthest tile to the nearest.
template <typename T, int OP, typename AggFn>
__device__ inline void wave_lookback_seg_sh(const TileStatus<T>& st, uint32_t tile,
                                            uint32_t& head_out, T& carry, AggFn agg_of) {
  using C = Combine<T, OP>;
  const uint32_t lane = __lane_id();
  T prefix = C::identity();
  uint32_t any_head = 0;
  int64_t base = (int64_t)tile - 1;
  for (;;) {
    const int64_t p = base - (int64_t)lane;
    T v = C::identity();
    uint32_t head = 0;
    uint32_t flag = p >= 0 ? 0u : kFlagInc;
    for (uint32_t polls = 0;; ++polls) {
      if (flag == 0) flag = st.try_read((uint32_t)p, head, v);
      const uint64_t pending = ballot(flag == 0);
      if (!pending) break;
      if (polls >= kSelfHelpPolls) {
        for (uint64_t m = pending; m; m &= m - 1) {
          const uint32_t k = (uint32_t)__builtin_ctzll(m);
          uint32_t hf = 0;
          T a = C::identity();
          agg_of((uint32_t)(base - (int64_t)k), hf, a);
          if (lane == k) {
            v = a;
            head = hf;
            flag = kFlagAgg;
          }
        }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    const uint64_t stop_mask = ballot(flag == kFlagInc || head != 0);
    const uint64_t head_mask = ballot(head != 0);
    if (stop_mask) {
      const uint32_t first = (uint32_t)__builtin_ctzll(stop_mask);
      T acc = C::identity();
      for (int k = (int)first; k >= 0; --k) acc = C::apply(acc, readlane(v, k));
      any_head |= (head_mask & ((first == 63) ? ~0ull : ((2ull << first) - 1))) ? 1u : 0u;
      prefix = C::apply(acc, prefix);
      break;
    }
    T acc = C::identity();
    for (int k = kWave - 1; k >= 0; --k) acc = C::apply(acc, readlane(v, k));
    prefix = C::apply(acc, prefix);
    base -= kWave;
  }
  head_out = any_head;
  carry = prefix;
}

// Order-preserving unsigned encodings for radix ordering (integer order and
// IEEE total order), see massively core/radix.rs:13-118.
template <typename K>
__device__ inline typename Bits<K>::U radix_encode(K k) {
  using U = typename Bits<K>::U;
  if constexpr (IsFloat<K>::value) {
    const U raw = bit_cast<U>(k);
    const U sign = (U)1 << (sizeof(U) * 8 - 1);
    return (raw & sign) ? (U)~raw : (U)(raw ^ sign);
  } else if constexpr ((K)-1 < (K)0) {
    const U sign = (U)1 << (sizeof(U) * 8 - 1);
    return (U)k ^ sign;
  } else {
    return (U)k;
  }
}

// Comparison modes of the ordering primitives: radix total order, and the
// built-in comparators Less / Greater (under which -0.0 == +0.0).
enum OrderMode : int { ORD_TOTAL = 0, ORD_LESS = 1, ORD_GREATER = 2 };

// Unsigned word whose numeric order is the requested key order.
template <typename K, int MODE>
__device__ inline typename Bits<K>::U order_bits(K k) {
  using U = typename Bits<K>::U;
  constexpr U sign = (U)1 << (sizeof(U) * 8 - 1);
  U e;
  if constexpr (IsFloat<K>::value && MODE != ORD_TOTAL) {
    U raw = bit_cast<U>(k);
    if ((raw & (U)~sign) == 0) raw = 0;  // -0 == +0 under `<`
    e = (raw & sign) ? (U)~raw : (U)(raw ^ sign);
  } else {
    e = radix_encode<K>(k);
  }
  if constexpr (MODE == ORD_GREATER) e = (U)~e;
  return e;
}

// First index in [lo, hi) with a[i] >= v (hi if none), for a nondecreasing
// array, searched by one whole wave: each round samples 64 evenly spaced
// points and keeps the gap that holds the boundary, so a search over 2^20
// entries takes 4 dependent loads instead of 20.
__device__ inline uint32_t wave_lower_bound(const uint32_t* a, uint32_t lo, uint32_t hi,
                                            uint32_t v) {
  const uint32_t lane = __lane_id();
  while (hi - lo > (uint32_t)kWave) {
    const uint32_t step = (hi - lo + kWave - 1) / kWave;
    const uint32_t idx = lo + lane * step;
    const bool below = idx < hi && a[idx] < v;
    const uint32_t cnt = (uint32_t)__builtin_popcountll(ballot(below));
    // Points lo + k*step for k < cnt are below v; the answer lies after the
    // last of them and at or before the next sampled point.
    const uint32_t nlo = cnt ? lo + (cnt - 1) * step + 1 : lo;
    const uint32_t nhi = min(hi, lo + cnt * step);
    lo = nlo;
    hi = nhi;
  }
  const uint32_t idx = lo + lane;
  const bool below = idx < hi && a[idx] < v;
  return lo + (uint32_t)__builtin_popcountll(ballot(below));
}

}  // namespace msv
