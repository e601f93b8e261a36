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

// An element-wise instruction applied in the store phase of the spectral pass that
// produces one of its operands: out[b][e] = op(v, o[b][e]) (lin_b: op(o[b][e], v)),
// v the pass's result, o the other operand (NULL: none; stride 0: one array for the
// whole batch).
struct PostOp {
  int kind = -1;  // -1: none
  int lin_b = 0;
  const double* o = nullptr;
  int64_t os = 0;
  double s = 0.0;
  double* dst = nullptr;
  int64_t ds = 0;
};

}  // namespace kad
