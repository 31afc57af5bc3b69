// Built-in selection predicates (device code), shared by copy_where and the
// batched jobs.
#pragma once

#include "device_common.hpp"
#include "../../include/msv.h"

namespace msv {

// Built-in predicates (massively/src/core/predicate.rs:40-52 built-ins).
enum PredKind : uint32_t {
  PRED_NONZERO = 0, PRED_ZERO, PRED_LT, PRED_LE, PRED_GT, PRED_GE, PRED_BITS_ANY, PRED_BITS_EQ
};

template <typename S>
struct Pred {
  uint32_t kind;
  S a;
  typename Bits<S>::U ua, ub;
  __device__ inline bool operator()(S x) const {
    using U = typename Bits<S>::U;
    switch (kind) {
      case PRED_NONZERO: return x != (S)0;
      case PRED_ZERO: return x == (S)0;
      case PRED_LT: return x < a;
      case PRED_LE: return x <= a;
      case PRED_GT: return x > a;
      case PRED_GE: return x >= a;
      case PRED_BITS_ANY: return (bit_cast<U>(x) & ua) != 0;
      default: return (bit_cast<U>(x) & ua) == ub;
    }
  }
};

// Device form of a msv_predicate for stencil dtype S.
template <typename S>
inline Pred<S> make_pred(const msv_predicate* p) {
  Pred<S> pred{};
  if (p) {
    using U = typename Bits<S>::U;
    pred.kind = p->kind;
    pred.ua = (U)p->a;
    pred.ub = (U)p->b;
    pred.a = bit_cast<S>((U)p->a);
  } else {
    pred.kind = PRED_NONZERO;
  }
  return pred;
}

}  // namespace msv
