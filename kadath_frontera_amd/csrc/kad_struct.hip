// Structure maps of the Jacobian and of its factors: which tiles are exactly zero.
//
// The domain-major Jacobian is block-tridiagonal by domain (rows of an equation
// couple only the unknowns of its own and the adjacent domains: take_into_account,
// src/System_of_eqs/solver.cpp:181,203), and partial pivoting keeps the LU factors
// inside that band.  The refinement loop multiplies by J and solves with the factors
// many times per Newton iteration, so each is scanned once for its nonzero tiles and
// every later product reads only those.
//
//  * GEMV map: 256-row x 1024-column tiles of A; a product y = b - op(A) x runs one
//    workgroup per nonzero tile and a fixed-order reduction (bitwise reproducible).
//  * Sweep extents: per 128-column block of a triangular sweep, the rows its panel
//    update must touch (the rest of the column block is exactly zero).
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include <vector>
#include <algorithm>

namespace kad {


// map[chunk * nrb + rb] = 1 if A(rb tile, chunk tile) has a nonzero.
__global__ __launch_bounds__(256) void gemv_map_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                       int* __restrict__ map) {
  const int rb = blockIdx.x, chunk = blockIdx.y;
  const int nrb = gridDim.x;
  const int i = rb * GM_ROWS + threadIdx.x;
  const int c0 = chunk * GM_COLS, c1 = min(n, c0 + GM_COLS);
  int nz = 0;
  if (i < n) {
    const double* a = A + i;
    for (int c = c0; c < c1 && !nz; c += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (c + q < c1) ? a[(int64_t)(c + q) * lda] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) nz |= (v[q] != 0.0);
    }
  }
  nz = __syncthreads_or(nz);
  if (threadIdx.x == 0) map[(int64_t)chunk * nrb + rb] = nz ? 1 : 0;
}

// part[chunk * n + i] = sum_{c in chunk} A(i, c) x(c) over the listed tiles only.
__global__ __launch_bounds__(256) void gemv_tile_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                        const int* __restrict__ list, const double* __restrict__ x,
                                                        double* __restrict__ part) {
  const int rb = list[2 * blockIdx.x], chunk = list[2 * blockIdx.x + 1];
  const int i = rb * GM_ROWS + threadIdx.x;
  if (i >= n) return;
  const int c0 = chunk * GM_COLS, c1 = min(n, c0 + GM_COLS);
  const double* a = A + i;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int c = c0;
  for (; c + 8 <= c1; c += 8) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = a[(int64_t)(c + q) * lda];
    s0 += v[0] * x[c] + v[4] * x[c + 4];
    s1 += v[1] * x[c + 1] + v[5] * x[c + 5];
    s2 += v[2] * x[c + 2] + v[6] * x[c + 6];
    s3 += v[3] * x[c + 3] + v[7] * x[c + 7];
  }
  for (; c < c1; ++c) s0 += a[(int64_t)c * lda] * x[c];
  part[(int64_t)chunk * n + i] = (s0 + s1) + (s2 + s3);
}

// y = b - sum over the nonzero chunks of row i's tile row (b may be null: y = sum)
__global__ void gemv_tile_reduce_kernel(const double* __restrict__ part, const int* __restrict__ map, int nrb,
                                        int nchunk, int n, const double* __restrict__ b, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int rb = i / GM_ROWS;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c)
    if (map[(int64_t)c * nrb + rb]) s += part[(int64_t)c * n + i];
  y[i] = b ? b[i] - s : s;
}

// y = b - A^T x: one workgroup per column, over the nonzero row tiles of its chunk.
__global__ __launch_bounds__(256) void gemv_t_tile_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                          const int* __restrict__ map, int nrb,
                                                          const double* __restrict__ x, const double* __restrict__ b,
                                                          double* __restrict__ y) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  const int chunk = c / GM_COLS;
  const double* a = A + (int64_t)c * lda;
  double s = 0.0;
  for (int rb = 0; rb < nrb; ++rb) {
    if (!map[(int64_t)chunk * nrb + rb]) continue;
    const int i = rb * GM_ROWS + threadIdx.x;
    if (i < n) s += a[i] * x[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[c] = b ? b[c] - red[0] : red[0];
}

// Scan A for its nonzero GEMV tiles (device map + compact host-built list).
int GemvMap::build(const double* A_, int64_t lda_, int n_, hipStream_t s) { return adopt(A_, lda_, n_, nullptr, s); }

int GemvMap::adopt(const double* A_, int64_t lda_, int n_, const int* dmap, hipStream_t s) {
  A = A_;
  lda = lda_;
  n = n_;
  nrb = (int)cdiv(n, GM_ROWS);
  nchunk = (int)cdiv(n, GM_COLS);
  const size_t cells = (size_t)nrb * nchunk;
  if (cells > cap) {
    if (map) KAD_HIP_CHECK(hipFree(map));
    if (list) KAD_HIP_CHECK(hipFree(list));
    map = list = nullptr;
    cap = 0;
    KAD_HIP_CHECK(hipMalloc((void**)&map, cells * sizeof(int)));
    KAD_HIP_CHECK(hipMalloc((void**)&list, 2 * cells * sizeof(int)));
    cap = cells;
  }
  if (dmap) {
    KAD_HIP_CHECK(hipMemcpyAsync(map, dmap, cells * sizeof(int), hipMemcpyDeviceToDevice, s));
  } else {
    gemv_map_kernel<<<dim3((unsigned)nrb, (unsigned)nchunk), 256, 0, s>>>(A, lda, n, map);
    KAD_HIP_CHECK(hipGetLastError());
  }
  std::vector<int> h(cells), l;
  KAD_HIP_CHECK(hipMemcpyAsync(h.data(), map, cells * sizeof(int), hipMemcpyDeviceToHost, s));
  KAD_HIP_CHECK(hipStreamSynchronize(s));
  l.reserve(2 * cells);
  for (int c = 0; c < nchunk; ++c)
    for (int rb = 0; rb < nrb; ++rb)
      if (h[(size_t)c * nrb + rb]) {
        l.push_back(rb);
        l.push_back(c);
      }
  nact = (int)(l.size() / 2);
  if (nact) KAD_HIP_CHECK(hipMemcpyAsync(list, l.data(), l.size() * sizeof(int), hipMemcpyHostToDevice, s));
  KAD_HIP_CHECK(hipStreamSynchronize(s));  // the host list is released on return
  return KAD_OK;
}

// y = b - op(A) x (or op(A) x) over the nonzero tiles of a map built for this A.
// part: nchunk * n doubles.
int GemvMap::apply(const double* x, const double* b, double* y, int trans, double* part, hipStream_t s) const {
  if (!trans) {
    if (nact) gemv_tile_kernel<<<(unsigned)nact, 256, 0, s>>>(A, lda, n, list, x, part);
    gemv_tile_reduce_kernel<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(part, map, nrb, nchunk, n, b, y);
  } else {
    gemv_t_tile_kernel<<<(unsigned)n, 256, 0, s>>>(A, lda, n, map, nrb, x, b, y);
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

GemvMap::~GemvMap() {
  if (map) (void)hipFree(map);
  if (list) (void)hipFree(list);
}

}  // namespace kad
