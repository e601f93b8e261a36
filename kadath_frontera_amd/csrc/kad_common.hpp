// Shared device/host helpers for the Kadath MI355X Newton path.
// All device code here targets gfx950 (CDNA4): 64-lane wavefronts, MFMA, 160 KB LDS per CU.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#define KAD_OK 0
#define KAD_ERR_ARG (-1)
#define KAD_ERR_HIP (-2)
#define KAD_ERR_ALLOC (-3)
#define KAD_ERR_NOCONV (-4)

#define KAD_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      fprintf(stderr, "[kadath_amd] HIP error %s at %s:%d: %s\n", hipGetErrorName(_e),  \
              __FILE__, __LINE__, #expr);                                                \
      return KAD_ERR_HIP;                                                                \
    }                                                                                    \
  } while (0)

namespace kad {

// Precision selector for the dense factorisation.  Storage type of the factor
// and the MFMA operand type of the trailing update:
//   F64  : double storage, v_mfma_f64_16x16x4f64
//   F32  : float storage,  v_mfma_f32_16x16x4f32
//   BF16 : float storage,  operands rounded to bf16, v_mfma_f32_16x16x32_bf16, fp32 accumulate
enum Prec : int { F64 = 0, F32 = 1, BF16 = 2 };

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace kad
