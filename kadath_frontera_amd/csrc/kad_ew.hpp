// The element-wise operation of the device programs (x: operand a, y: operand b, s: the
// instruction's scalar), shared by the element-wise kernels and by the spectral pass
// that applies one in its store phase, so both give the same bits.
#pragma once
#include <hip/hip_runtime.h>

namespace kad {

enum : int { EWOP_ADD = 0, EWOP_SUB = 1, EWOP_MUL = 2, EWOP_DIV = 3, EWOP_SCAL = 4, EWOP_ADDS = 5, EWOP_COPY = 6,
             EWOP_SDIV = 7 };

__device__ __forceinline__ double ew_op(int kind, double x, double y, double s) {
  switch (kind) {
    case EWOP_ADD: return x + y;
    case EWOP_SUB: return x - y;
    case EWOP_MUL: return x * y;
    case EWOP_DIV: return x / y;
    case EWOP_SCAL: return x * s;
    case EWOP_SDIV: return x / s;
    case EWOP_ADDS: return x + s;
    default: return x;
  }
}

// A chain of element-wise instructions applied to a value v as it is stored, each to
// the previous one's result: step k makes v = op(v, o[b][e]) (side 0) or op(o[b][e], v)
// (side 1), o its other operand (NULL: none; stride 0: one array for the whole batch).
// The spectral pass applies one to its result, and the chain kernel to the result of
// its first instruction; the intermediate arrays are never stored.
constexpr int kMaxChain = 3;
struct EwStep {
  int kind = 0;
  int side = 0;
  const double* o = nullptr;
  int64_t os = 0;
  double s = 0.0;
};
struct PostOp {
  int n = 0;  // steps (0: none)
  EwStep op[kMaxChain];
  double* dst = nullptr;
  int64_t ds = 0;
};

// The chain's other operands at element e of batch row b, all loads issued before any
// is used (one memory latency for the whole chain), then the steps.
__device__ __forceinline__ void post_load(const PostOp& p, int64_t b, int e, double (&o)[kMaxChain]) {
#pragma unroll
  for (int k = 0; k < kMaxChain; ++k) o[k] = (k < p.n && p.op[k].o) ? p.op[k].o[b * p.op[k].os + e] : 0.0;
}
__device__ __forceinline__ double post_ops(const PostOp& p, double v, const double (&o)[kMaxChain]) {
#pragma unroll
  for (int k = 0; k < kMaxChain; ++k)
    if (k < p.n) {
      const EwStep& st = p.op[k];
      v = st.side ? ew_op(st.kind, o[k], v, st.s) : ew_op(st.kind, v, o[k], st.s);
    }
  return v;
}
__device__ __forceinline__ double post_apply(const PostOp& p, int64_t b, int e, double v) {
  double o[kMaxChain];
  post_load(p, b, e, o);
  return post_ops(p, v, o);
}

}  // namespace kad
