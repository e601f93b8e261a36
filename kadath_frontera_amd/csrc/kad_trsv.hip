// Triangular solves with the LU factors (the solve half of pdgesv_,
// src_par/do_newton.cpp:317-320; Matrice::solve, src/Matrice/matrice.cpp:221-250).
//
// op(A) x = b with P A = L U.  trans = 0: L y = P b, U x = y; trans = 1:
// U^T z = b, L^T w = z, x = P^T w.  Either way the first sweep is a lower
// ("forward") triangle Tf and the second an upper ("backward") one Tb.
//
// The factors are reused by every refinement step of a Newton iteration, so a
// set-up pass builds, per block step, a contiguous fp64 matrix
//   W_k = [ inv(T_kk) | -inv(T_kk) T_kp ]        (TBS x 2 TBS, column-major)
// where p is the block solved just before k.  A sweep is then one launch per
// block: a few "chain" workgroups compute sol_k = W_k [rhs_k ; sol_p] (a
// coalesced GEMV, the only dependent work), the other workgroups subtract the
// columns of block p from the rows not yet reached.
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include <vector>
#include <algorithm>

namespace kad {

constexpr int TBS = 128;                 // diagonal block
constexpr int WCOLS = 2 * TBS;           // columns of W_k
constexpr int CHAIN_WG = 4;              // chain workgroups per step
constexpr int CHAIN_ROWS = TBS / CHAIN_WG;
constexpr int PANEL_ROWS = 64;           // rows per panel workgroup (4 column groups of 32)

// Entry (r, c) of the sweep's triangle: the stored factor, or its transpose.
template <typename T, bool TR>
__device__ __forceinline__ double tri_at(const T* F, int64_t lda, int r, int c) {
  return TR ? (double)F[c + (int64_t)r * lda] : (double)F[r + (int64_t)c * lda];
}

// Inverse of diagonal block blockIdx.x of the factor's unit-lower (uplo 0) or
// upper (uplo 1) triangle, by columns as LAPACK dtrti2; written transposed when
// tr_out (the inverse of the transposed block).  Out: first TBS columns of W_k.
template <typename T>
__global__ __launch_bounds__(256) void diag_inverse_kernel(const T* __restrict__ F, int64_t lda, int n, int uplo,
                                                           int tr_out, double* __restrict__ W, int64_t sF = 0,
                                                           int64_t sW = 0) {
  F += (int64_t)blockIdx.y * sF;  // factors blockIdx.y of a batch
  W += (int64_t)blockIdx.y * sW;
  __shared__ double M[TBS][TBS + 1];
  __shared__ double col[TBS];
  const int blk = blockIdx.x;
  const int r0 = blk * TBS;
  const int nb = min(TBS, n - r0);
  const int t = threadIdx.x;
  for (int e = t; e < TBS * TBS; e += 256) {
    const int i = e % TBS, c = e / TBS;
    double v = 0.0;
    if (i < nb && c < nb) {
      const bool in_tri = uplo == 0 ? (i > c) : (i <= c);
      if (in_tri) v = (double)F[(r0 + i) + (int64_t)(r0 + c) * lda];
      if (uplo == 0 && i == c) v = 1.0;
    }
    M[i][c] = v;
  }
  __syncthreads();
  if (uplo == 1) {
    // column j: M[j][j] <- 1/M[j][j]; M[0:j, j] <- -M[j][j] * (inv(M[0:j, 0:j]) M[0:j, j])
    for (int j = 0; j < nb; ++j) {
      if (t < j) col[t] = M[t][j];
      __syncthreads();
      const double ajj = 1.0 / M[j][j];
      if (t < j) {
        double s = 0.0;
        for (int k = t; k < j; ++k) s += M[t][k] * col[k];
        M[t][j] = -ajj * s;
      }
      __syncthreads();
      if (t == 0) M[j][j] = ajj;
      __syncthreads();
    }
  } else {
    // column j (descending): M[j+1:, j] <- -(inv(M[j+1:, j+1:]) M[j+1:, j])
    for (int j = nb - 2; j >= 0; --j) {
      if (t > j && t < nb) col[t] = M[t][j];
      __syncthreads();
      if (t > j && t < nb) {
        double s = 0.0;
        for (int k = j + 1; k <= t; ++k) s += M[t][k] * col[k];
        M[t][j] = -s;
      }
      __syncthreads();
    }
  }
  double* out = W + (int64_t)blk * TBS * WCOLS;
  for (int e = t; e < TBS * TBS; e += 256) {
    const int i = e % TBS, c = e / TBS;
    out[e] = tr_out ? M[c][i] : M[i][c];
  }
}

// Coupling half of W_k: -inv(T_kk) T_kp, with T accessed through tri_at and
// p = k - 1 (forward) or k + 1 (backward).  One workgroup per k with a p.
template <typename T, bool TR, bool FWD>
__global__ __launch_bounds__(256) void couple_kernel(const T* __restrict__ F, int64_t lda, int n, int nblk,
                                                     double* __restrict__ W, int64_t sF = 0, int64_t sW = 0) {
  F += (int64_t)blockIdx.y * sF;  // factors blockIdx.y of a batch
  W += (int64_t)blockIdx.y * sW;
  __shared__ double B[TBS][TBS + 1];
  const int k = FWD ? blockIdx.x + 1 : blockIdx.x;
  const int p = FWD ? k - 1 : k + 1;
  if (k >= nblk || p < 0 || p >= nblk) return;
  const int r0 = k * TBS, c0 = p * TBS;
  const int nr = min(TBS, n - r0), nc = min(TBS, n - c0);
  const int t = threadIdx.x;
  for (int e = t; e < TBS * TBS; e += 256) {
    const int i = e % TBS, c = e / TBS;
    B[i][c] = (i < nr && c < nc) ? tri_at<T, TR>(F, lda, r0 + i, c0 + c) : 0.0;
  }
  __syncthreads();
  double* Wk = W + (int64_t)k * TBS * WCOLS;
  const int j = t % TBS;            // output column
  const int i0 = (t / TBS) * 64;    // 64 output rows per thread
  double acc[64];
#pragma unroll
  for (int i = 0; i < 64; ++i) acc[i] = 0.0;
  for (int q = 0; q < TBS; ++q) {
    const double b = B[q][j];
    const double* dq = Wk + (int64_t)q * TBS + i0;  // column q of inv(T_kk), same for all lanes
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] += dq[i] * b;
  }
  double* G = Wk + (int64_t)(TBS + j) * TBS + i0;
#pragma unroll
  for (int i = 0; i < 64; ++i) G[i] = -acc[i];
}

// One block step.  cur: the block whose solution is final (outside [0, nblk) at
// the first step); next = cur +- 1 is solved by the chain workgroups:
//   sol_next = W_next [rhs_next ; sol_cur]
// and workgroups >= CHAIN_WG subtract T(rows, cur) sol_cur from the rows beyond next.
template <typename T, bool TR, bool FWD>
__global__ __launch_bounds__(256) void sweep_step_kernel(const T* __restrict__ F, int64_t lda, int n, int nblk,
                                                         const double* __restrict__ W, int cur, int rlo, int rhi,
                                                         double* __restrict__ rhs, double* __restrict__ sol,
                                                         int64_t sF, int64_t sW, int64_t sV) {
  // system blockIdx.y of a batch: its factors, tables and vectors
  F += (int64_t)blockIdx.y * sF;
  W += (int64_t)blockIdx.y * sW;
  rhs += (int64_t)blockIdx.y * sV;
  sol += (int64_t)blockIdx.y * sV;
  __shared__ double vin[WCOLS];
  __shared__ double part[4][PANEL_ROWS];
  __shared__ double cpart[8][CHAIN_ROWS];
  const int t = threadIdx.x;
  const bool have_cur = cur >= 0 && cur < nblk;
  const int c0 = cur * TBS;
  const int ncb = have_cur ? min(TBS, n - c0) : 0;
  const int next = FWD ? cur + 1 : cur - 1;
  if (blockIdx.x < CHAIN_WG) {
    if (next < 0 || next >= nblk) return;
    const int r0 = next * TBS;
    const int nrb = min(TBS, n - r0);
    if (t < TBS) vin[t] = (t < nrb) ? rhs[r0 + t] : 0.0;
    else vin[t] = (t - TBS < ncb) ? sol[c0 + t - TBS] : 0.0;
    __syncthreads();
    const int i = t % CHAIN_ROWS;           // row within this workgroup's slice
    const int g = t / CHAIN_ROWS;           // 8 groups of 32 columns
    const int row = blockIdx.x * CHAIN_ROWS + i;
    const int qe = have_cur ? WCOLS : TBS;
    const double* Wn = W + (int64_t)next * TBS * WCOLS + row;
    double s0 = 0.0, s1 = 0.0;
    if (g * 32 < qe) {
#pragma unroll
      for (int q = 0; q < 32; q += 2) {
        s0 += Wn[(int64_t)(g * 32 + q) * TBS] * vin[g * 32 + q];
        s1 += Wn[(int64_t)(g * 32 + q + 1) * TBS] * vin[g * 32 + q + 1];
      }
    }
    cpart[g][i] = s0 + s1;
    __syncthreads();
    if (t < CHAIN_ROWS && row < nrb) {
      double s = 0.0;
#pragma unroll
      for (int gg = 0; gg < 8; ++gg) s += cpart[gg][t];
      sol[r0 + row] = s;
    }
    return;
  }
  if (!have_cur) return;
  if (t < TBS) vin[t] = (t < ncb) ? sol[c0 + t] : 0.0;
  __syncthreads();
  const int i = t % PANEL_ROWS;
  const int g = t / PANEL_ROWS;  // 4 groups of 32 columns
  // rows [rlo, rhi): beyond the next block, within the nonzero extent of block cur
  const int wb = blockIdx.x - CHAIN_WG;
  const int r = rlo + wb * PANEL_ROWS + i;
  const int rend = rhi;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (r < rend) {
    const int cb = g * 32;
    if (cb + 32 <= ncb) {
      double v[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) v[q] = tri_at<T, TR>(F, lda, r, c0 + cb + q);
#pragma unroll
      for (int q = 0; q < 32; q += 4) {
        s0 += v[q] * vin[cb + q];
        s1 += v[q + 1] * vin[cb + q + 1];
        s2 += v[q + 2] * vin[cb + q + 2];
        s3 += v[q + 3] * vin[cb + q + 3];
      }
    } else {
      for (int q = cb; q < ncb; ++q) s0 += tri_at<T, TR>(F, lda, r, c0 + q) * vin[q];
    }
  }
  part[g][i] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (g == 0 && r < rend) rhs[r] -= (part[0][i] + part[1][i]) + (part[2][i] + part[3][i]);
}

__global__ void gather_perm_kernel(const double* __restrict__ src, const int* __restrict__ perm,
                                   double* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
__global__ void scatter_perm_kernel(const double* __restrict__ src, const int* __restrict__ perm,
                                    double* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}
__global__ void copy_vec_kernel(const double* __restrict__ src, double* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Nonzero flags of the off-diagonal 128 x 128 tiles of a sweep's triangle:
// flag[b * nblk + rc] = 1 if T(row block rc, column block b) has a nonzero, for rc
// beyond b in the sweep's direction (0 elsewhere).  T(r, c) is the stored factor's
// (r, c) entry, or its (c, r) entry for the transposed sweeps.
template <typename T, bool TR, bool FWD>
__global__ __launch_bounds__(256) void tile_flag_kernel(const T* __restrict__ F, int64_t lda, int n, int nblk,
                                                        int* __restrict__ flag) {
  const int b = blockIdx.x, rc = blockIdx.y;
  int nz = 0;
  if (FWD ? rc > b : rc < b) {
    // stored tile: rows of block (TR ? b : rc), columns of block (TR ? rc : b)
    const int r0 = (TR ? b : rc) * TBS, c0 = (TR ? rc : b) * TBS;
    const int nr = min(TBS, n - r0), nc = min(TBS, n - c0);
    const int i = threadIdx.x % TBS;
    for (int c = threadIdx.x / TBS; c < nc && !nz; c += 2)
      if (i < nr) nz |= (F[(r0 + i) + (int64_t)(c0 + c) * lda] != (T)0);
  }
  nz = __syncthreads_or(nz);
  if (threadIdx.x == 0) flag[(int64_t)b * nblk + rc] = nz ? 1 : 0;
}

// Doubles of the W tables of one factorisation (forward then backward sweep).
size_t trsv_table_doubles(int n) { return (size_t)2 * cdiv(n, TBS) * TBS * WCOLS; }
// Ints of the tile flags of one sweep.
size_t trsv_flag_ints(int n) { return (size_t)cdiv(n, TBS) * cdiv(n, TBS); }

// Row extents of the panel updates of both sweeps from the tile flags: for column
// block b, the forward sweep's update ends at row fhi[b] and the backward sweep's
// starts at row blo[b] (rows outside are exactly zero in that column block).
template <typename T, bool TR>
static int sweep_extents(const T* LU, int64_t lda, int n, int* flag, TrsvExtents& ext, hipStream_t s) {
  const int nblk = (int)cdiv(n, TBS);
  const size_t cells = (size_t)nblk * nblk;
  std::vector<int> h(cells);
  ext.fhi.assign(nblk, 0);
  ext.blo.assign(nblk, 0);
  tile_flag_kernel<T, TR, true><<<dim3(nblk, nblk), 256, 0, s>>>(LU, lda, n, nblk, flag);
  KAD_HIP_CHECK(hipGetLastError());
  KAD_HIP_CHECK(hipMemcpyAsync(h.data(), flag, cells * sizeof(int), hipMemcpyDeviceToHost, s));
  KAD_HIP_CHECK(hipStreamSynchronize(s));
  for (int b = 0; b < nblk; ++b) {
    int last = -1;
    for (int rc = b + 1; rc < nblk; ++rc)
      if (h[(size_t)b * nblk + rc]) last = rc;
    ext.fhi[b] = last < 0 ? 0 : std::min(n, (last + 1) * TBS);
  }
  tile_flag_kernel<T, TR, false><<<dim3(nblk, nblk), 256, 0, s>>>(LU, lda, n, nblk, flag);
  KAD_HIP_CHECK(hipGetLastError());
  KAD_HIP_CHECK(hipMemcpyAsync(h.data(), flag, cells * sizeof(int), hipMemcpyDeviceToHost, s));
  KAD_HIP_CHECK(hipStreamSynchronize(s));
  for (int b = 0; b < nblk; ++b) {
    int first = b;
    for (int rc = 0; rc < b; ++rc)
      if (h[(size_t)b * nblk + rc]) {
        first = rc;
        break;
      }
    ext.blo[b] = first * TBS;
  }
  return KAD_OK;
}

// Build the W tables and the row extents of the sweeps of op(A) (trans as in getrs_fast).
template <typename T>
int trsv_prepare(const T* LU, int64_t lda, int n, int trans, double* W, int* flag, TrsvExtents& ext,
                 hipStream_t s) {
  const int nblk = (int)cdiv(n, TBS);
  double* Wf = W;
  double* Wb = W + (size_t)nblk * TBS * WCOLS;
  // forward triangle L (trans 0) or U^T (trans 1): stored lower part, or upper part transposed;
  // backward triangle U or L^T: the same tiles as the forward ones of the other factor.
  // Without a flag buffer the extents stay empty and the sweeps cover every row.
  if (flag) {
    int rc = trans ? sweep_extents<T, true>(LU, lda, n, flag, ext, s) : sweep_extents<T, false>(LU, lda, n, flag, ext, s);
    if (rc) return rc;
  }
  // forward triangle: L (trans 0) or U^T (trans 1); backward: U or L^T
  diag_inverse_kernel<T><<<nblk, 256, 0, s>>>(LU, lda, n, trans ? 1 : 0, trans, Wf);
  diag_inverse_kernel<T><<<nblk, 256, 0, s>>>(LU, lda, n, trans ? 0 : 1, trans, Wb);
  if (nblk > 1) {
    if (!trans) {
      couple_kernel<T, false, true><<<nblk - 1, 256, 0, s>>>(LU, lda, n, nblk, Wf);
      couple_kernel<T, false, false><<<nblk - 1, 256, 0, s>>>(LU, lda, n, nblk, Wb);
    } else {
      couple_kernel<T, true, true><<<nblk - 1, 256, 0, s>>>(LU, lda, n, nblk, Wf);
      couple_kernel<T, true, false><<<nblk - 1, 256, 0, s>>>(LU, lda, n, nblk, Wb);
    }
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// nbat > 1: nbat systems at LU + z sF, W + z sW, rhs / sol + z sV.
template <typename T, bool TR, bool FWD>
static void sweep(const T* LU, int64_t lda, int n, int nblk, const double* W, const TrsvExtents& ext, double* rhs,
                  double* sol, hipStream_t s, int nbat = 1, int64_t sF = 0, int64_t sW = 0, int64_t sV = 0) {
  if (FWD) {
    for (int cur = -1; cur < nblk - 1; ++cur) {
      // panel rows: beyond block cur + 1, up to the column block's last nonzero row tile
      const int rlo = (cur + 2) * TBS;
      const int rhi = cur < 0 ? 0 : ext.fhi.empty() ? n : std::min(n, ext.fhi[cur]);
      const int rows = rhi - rlo;
      const unsigned grid = CHAIN_WG + ((cur >= 0 && rows > 0) ? (unsigned)cdiv(rows, PANEL_ROWS) : 0u);
      sweep_step_kernel<T, TR, true><<<dim3(grid, nbat), 256, 0, s>>>(LU, lda, n, nblk, W, cur, rlo, rhi, rhs, sol,
                                                                     sF, sW, sV);
    }
  } else {
    for (int cur = nblk; cur >= 1; --cur) {
      // panel rows: from the column block's first nonzero row tile, before block cur - 1
      const int rhi = (cur - 1) * TBS;
      const int rlo = cur >= nblk ? rhi : ext.blo.empty() ? 0 : ext.blo[cur];
      const int rows = rhi - rlo;
      const unsigned grid = CHAIN_WG + ((cur < nblk && rows > 0) ? (unsigned)cdiv(rows, PANEL_ROWS) : 0u);
      sweep_step_kernel<T, TR, false><<<dim3(grid, nbat), 256, 0, s>>>(LU, lda, n, nblk, W, cur, rlo, rhi, rhs, sol,
                                                                      sF, sW, sV);
    }
  }
}

// b <- op(A)^{-1} b.  work: 3 n doubles.  W from trsv_prepare with the same trans.
template <typename T>
int getrs_fast(const T* LU, int64_t lda, int n, const int* perm, const double* W, const TrsvExtents& ext, double* b,
               int trans, double* work, hipStream_t s) {
  const int nblk = (int)cdiv(n, TBS);
  const double* Wf = W;
  const double* Wb = W + (size_t)nblk * TBS * WCOLS;
  double* x = work;
  double* y = work + n;
  const unsigned g = (unsigned)cdiv(n, 256);
  if (!trans) {
    gather_perm_kernel<<<g, 256, 0, s>>>(b, perm, x, n);
    sweep<T, false, true>(LU, lda, n, nblk, Wf, ext, x, y, s);   // L y = P b
    sweep<T, false, false>(LU, lda, n, nblk, Wb, ext, y, b, s);  // U x = y
  } else {
    double* z = work + 2 * (size_t)n;
    copy_vec_kernel<<<g, 256, 0, s>>>(b, x, n);
    sweep<T, true, true>(LU, lda, n, nblk, Wf, ext, x, y, s);   // U^T z = b
    sweep<T, true, false>(LU, lda, n, nblk, Wb, ext, y, z, s);  // L^T w = z
    scatter_perm_kernel<<<g, 256, 0, s>>>(z, perm, b, n);  // x = P^T w
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// Batched: the W tables of nbat fp64 factorisations (LU + z sA, order n, trans 0, no
// row extents) at W + z trsv_table_doubles(n).
int trsv_prepare_batched(const double* LU, int64_t lda, int n, int nbat, int64_t sA, double* W, hipStream_t s) {
  const int nblk = (int)cdiv(n, TBS);
  const int64_t sW = (int64_t)trsv_table_doubles(n);
  double* Wf = W;
  double* Wb = W + (size_t)nblk * TBS * WCOLS;
  diag_inverse_kernel<double><<<dim3(nblk, nbat), 256, 0, s>>>(LU, lda, n, 0, 0, Wf, sA, sW);
  diag_inverse_kernel<double><<<dim3(nblk, nbat), 256, 0, s>>>(LU, lda, n, 1, 0, Wb, sA, sW);
  if (nblk > 1) {
    couple_kernel<double, false, true><<<dim3(nblk - 1, nbat), 256, 0, s>>>(LU, lda, n, nblk, Wf, sA, sW);
    couple_kernel<double, false, false><<<dim3(nblk - 1, nbat), 256, 0, s>>>(LU, lda, n, nblk, Wb, sA, sW);
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// x_z <- U_z^{-1} L_z^{-1} x_z for nbat systems whose right-hand sides x (nbat x n, at
// x + z n) are already in pivoted order; y: nbat x n doubles of work.
int getrs_permuted_batched(const double* LU, int64_t lda, int n, int nbat, int64_t sA, const double* W, double* x,
                           double* y, hipStream_t s) {
  const int nblk = (int)cdiv(n, TBS);
  const int64_t sW = (int64_t)trsv_table_doubles(n);
  const double* Wf = W;
  const double* Wb = W + (size_t)nblk * TBS * WCOLS;
  const TrsvExtents none;
  sweep<double, false, true>(LU, lda, n, nblk, Wf, none, x, y, s, nbat, sA, sW, n);   // L y = x
  sweep<double, false, false>(LU, lda, n, nblk, Wb, none, y, x, s, nbat, sA, sW, n);  // U x = y
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

template int trsv_prepare<double>(const double*, int64_t, int, int, double*, int*, TrsvExtents&, hipStream_t);
template int trsv_prepare<float>(const float*, int64_t, int, int, double*, int*, TrsvExtents&, hipStream_t);
template int getrs_fast<double>(const double*, int64_t, int, const int*, const double*, const TrsvExtents&, double*,
                                int, double*, hipStream_t);
template int getrs_fast<float>(const float*, int64_t, int, const int*, const double*, const TrsvExtents&, double*,
                               int, double*, hipStream_t);

}  // namespace kad
