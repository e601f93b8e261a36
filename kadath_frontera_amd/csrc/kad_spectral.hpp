// Launch interface of the batched 1-D / 3-D spectral kernels.
#pragma once
#include "kad_common.hpp"
#include "kad_ew.hpp"

namespace kad {

constexpr int MAX_PASSES = 6;

// A chain of 1-D passes over a 3-D array.  Pass p acts along axis[p] and maps
// that axis to nout[p] entries; line_mat[p] (device) gives, per line, the offset
// of a row-major nout x nin matrix in the pool (or -1: the line is zero).  Lines
// are numbered like the reference's per-dimension basis tables
// (src/Base_spectral/base_spectral.cpp:92-104): axis 0 by (j + d1*k), axis 1 by k,
// axis 2 by 0.
struct Passes {
  int npass = 0;
  int maxsize = 0;    // largest intermediate array (doubles)
  int pool_size = 0;  // doubles of the plan's matrix pool (staged in LDS)
  int axis[MAX_PASSES];
  int nout[MAX_PASSES];
  int nlines[MAX_PASSES];  // entries of line_mat[p] (staged in LDS)
  const int* line_mat[MAX_PASSES];
};

// post (may be NULL): element-wise instructions applied in the store phase; out is then
// not written (accumulate 0 and scale 1 only)
int apply3d_launch(const double* in, int64_t in_stride, double* out, int64_t out_stride, int nbatch, int d0, int d1,
                   int d2, const Passes& P, const double* pool, int accumulate, double scale, hipStream_t s,
                   const PostOp* post = nullptr);
int apply_lines_launch(const double* in, double* out, int64_t L, int nin, int nout, const double* M, hipStream_t s);

}  // namespace kad
