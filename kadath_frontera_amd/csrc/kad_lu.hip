// Dense LU with partial pivoting + solves + mixed-precision refinement, single MI355X.
//
// Replaces the ScaLAPACK solve of System_of_eqs::do_newton
// (reference src_par/do_newton.cpp:317-320, pdgesv_) and the single-process
// LAPACK path (src/Matrice/matrice.cpp:192-250, dgetrf_/dgetrs_).
//
// Layout: the n x n Jacobian is column-major with leading dimension lda (the
// reference packs J_in column-major before pdgemr2d/pdgesv, do_newton.cpp:178-185).
// Pivots are 0-based global row indices, LAPACK semantics: at step i row i was
// interchanged with row ipiv[i] (ipiv[i] >= i).
//
// Factorisation: right-looking, outer panel width NB; each panel is factored by
// a recursive split (LAPACK getrf2 style) whose base is a per-column kernel that
// performs exact partial pivoting (first index of max |a|, as idamax) and whose
// internal updates are MFMA GEMMs.  The trailing update A22 -= L21 U12 is the
// MFMA GEMM of kad_gemm.hip in the factor precision (f64, f32, or bf16 operands
// with f32 accumulation).  U12 = L11^{-1} A12 is a blocked substitution whose
// updates are GEMMs and whose 64-row leaves use explicit inverses.  Trailing-update
// tile pairs whose L21 row tile or U12 column tile is exactly zero are skipped
// (structural zeros of a block-sparse Jacobian); the result is unchanged.
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include <vector>
#include <cmath>
#include <algorithm>

namespace kad {

int gemm_launch(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                int64_t ldc, hipStream_t s);
int gemm_launch_masked(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                       void* C, int64_t ldc, const int* rmask, const int* cmask, hipStream_t s);
int gemm_launch_sel(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                    void* C, int64_t ldc, const TileSel& sel, hipStream_t s);
int gemm_launch_batched(int prec, int op, int m, int n, int k, const void* A, int64_t lda, int64_t sa, const void* B,
                        int64_t ldb, int64_t sb, void* C, int64_t ldc, int64_t sc, int nbat, hipStream_t s);
constexpr int GOP_SUB = 0, GOP_SET = 1;

constexpr int PANEL_BASE = 32;  // columns factored by the per-column kernel

// An argument error of the factorisation or its solves, with its reason on stderr.
inline int arg_error(const char* why) {
  std::fprintf(stderr, "[kadath_amd] %s\n", why);
  return KAD_ERR_ARG;
}

// Block-wide (max |a|, first index) reduction.
template <typename T>
__device__ void block_argmax(double v, int idx, double* sv, int* si, double& bv, int& bi) {
  const int tid = threadIdx.x;
  sv[tid] = v;
  si[tid] = idx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double ov = sv[tid + s];
      const int oi = si[tid + s];
      if (ov > sv[tid] || (ov == sv[tid] && oi < si[tid] && oi >= 0) || si[tid] < 0) {
        if (oi >= 0) {
          sv[tid] = ov;
          si[tid] = oi;
        }
      }
    }
    __syncthreads();
  }
  bv = sv[0];
  bi = si[0];
}

// Row interchanges of pivots [k1, k2) applied to columns [c0, c1): one thread per column.
template <typename T>
__global__ void laswp_kernel(T* A, int64_t lda, int c0, int c1, const int* __restrict__ ipiv, int k1, int k2,
                             int64_t sA, int sP) {
  A += (int64_t)blockIdx.y * sA;  // matrix blockIdx.y of a batch
  ipiv += (int64_t)blockIdx.y * sP;
  const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  T* col = A + (int64_t)c * lda;
  for (int i = k1; i < k2; ++i) {
    const int p = ipiv[i];
    if (p != i) {
      const T t = col[i];
      col[i] = col[p];
      col[p] = t;
    }
  }
}

// Inverse of a unit-lower nb x nb block (nb <= 64), written as a full square
// (zeros above, ones on the diagonal) into Linv (ld ldo).  One workgroup of 256:
// X = I, then for j = 0..nb-1 every (i > j, c <= j) entry does
// X[i][c] -= L[i][j] X[j][c]  (row-oriented forward substitution, one barrier per j).
template <typename T>
__global__ __launch_bounds__(256) void trtri_small_kernel(const T* L, int64_t lda, int nb, T* Linv, int64_t ldo) {
  __shared__ double Ls[64][65];
  __shared__ double X[64][65];
  const int t = threadIdx.x;
  for (int e = t; e < nb * nb; e += 256) {
    const int i = e % nb, c = e / nb;
    Ls[i][c] = (double)L[i + (int64_t)c * lda];
    X[i][c] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < nb - 1; ++j) {
    // rows i in (j, nb), columns c in [0, j]
    const int nrow = nb - 1 - j, ncol = j + 1;
    for (int e = t; e < nrow * ncol; e += 256) {
      const int i = j + 1 + e % nrow, c = e / nrow;
      X[i][c] -= Ls[i][j] * X[j][c];
    }
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += 256) {
    const int i = e % nb, c = e / nb;
    Linv[i + (int64_t)c * ldo] = (T)X[i][c];
  }
}

template <typename T>
__global__ void copy2d_kernel(const T* src, int64_t lds, T* dst, int64_t ldd, int m, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i < m && c < n) dst[i + (int64_t)c * ldd] = src[i + (int64_t)c * lds];
}

template <typename T>
__global__ void negate_copy_kernel(const T* src, int64_t lds, T* dst, int64_t ldd, int m, int n, T scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i < m && c < n) dst[i + (int64_t)c * ldd] = scale * src[i + (int64_t)c * lds];
}

template <typename T>
__global__ void zero_upper_kernel(T* Lv, int64_t ld, int nb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i < nb && c < nb && i < c) Lv[i + (int64_t)c * ld] = T(0);
}

template <typename T> constexpr int prec_of() { return sizeof(T) == 8 ? F64 : F32; }

// Recursive inverse of unit-lower L (nb x nb) into Linv (ld ldo), full square.
template <typename T>
int trtri_rec(const T* L, int64_t lda, int nb, T* Linv, int64_t ldo, LuWork& w) {
  if (nb <= 64) {
    trtri_small_kernel<T><<<1, 256, 0, w.stream>>>(L, lda, nb, Linv, ldo);
    KAD_HIP_CHECK(hipGetLastError());
    return KAD_OK;
  }
  const int h = nb / 2;
  int rc = trtri_rec<T>(L, lda, h, Linv, ldo, w);
  if (rc) return rc;
  rc = trtri_rec<T>(L + h + (int64_t)h * lda, lda, nb - h, Linv + h + (int64_t)h * ldo, ldo, w);
  if (rc) return rc;
  // T1 = L21 * Linv11  ((nb-h) x h) into tmp2
  T* t1 = (T*)w.tmp2;
  rc = gemm_launch(prec_of<T>(), GOP_SET, nb - h, h, h, L + h, lda, Linv, ldo, t1, nb - h, w.stream);
  if (rc) return rc;
  // Linv21 = -Linv22 * T1
  T* out21 = Linv + h;
  rc = gemm_launch(prec_of<T>(), GOP_SET, nb - h, h, nb - h, Linv + h + (int64_t)h * ldo, ldo, t1, nb - h, out21, ldo,
                   w.stream);
  if (rc) return rc;
  dim3 g((unsigned)cdiv(nb - h, 256), h);
  negate_copy_kernel<T><<<g, 256, 0, w.stream>>>(out21, ldo, out21, ldo, nb - h, h, T(-1));
  dim3 gz((unsigned)cdiv(nb, 256), nb);
  zero_upper_kernel<T><<<gz, 256, 0, w.stream>>>(Linv, ldo, nb);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// Structural-zero masks of the trailing update: per 128-row tile of the L panel
// (rows x ncols) and per 128-column tile of the U row panel (nrows x cols): 1 if the
// tile holds any entry != 0 (NaN counts as nonzero), else 0.
template <typename T>
__global__ __launch_bounds__(256) void row_tile_mask_kernel(const T* __restrict__ A, int64_t lda, int nrows,
                                                            int ncols, int* __restrict__ mask, int nzrows = -1) {
  // rows at or past nzrows (when given) are known to be zero: those tiles are not read
  if (nzrows >= 0 && (int)blockIdx.x * 128 >= nzrows) {
    if (threadIdx.x == 0) mask[blockIdx.x] = 0;
    return;
  }
  const int r = blockIdx.x * 128 + (threadIdx.x & 127);
  int any = 0;
  if (r < nrows)
    for (int c = threadIdx.x >> 7; c < ncols; c += 2) any |= (A[r + (int64_t)c * lda] != T(0));
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) mask[blockIdx.x] = any;
}
// Last nonzero row (or -1) of each 128-column tile within each 2048-row chunk:
// out[chunk * ntc + tile].  One workgroup per (tile, chunk); rows coalesced.
template <typename T>
__global__ __launch_bounds__(256) void col_extent_kernel(const T* __restrict__ A, int64_t lda, int n,
                                                         int* __restrict__ out) {
  __shared__ int red[256];
  const int c0 = blockIdx.x * 128, r0 = blockIdx.y * 2048;
  const int ce = min(c0 + 128, n), re = min(r0 + 2048, n);
  int last = -1;
  for (int c = c0; c < ce; ++c)
    for (int r = r0 + threadIdx.x; r < re; r += 256)
      if (A[r + (int64_t)c * lda] != T(0)) last = max(last, r);
  red[threadIdx.x] = last;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Compacted lists of the active row / column tiles: list[0..) rows, list[stride..)
// columns, list[2 stride] and list[2 stride + 1] their counts.  One thread (a few
// hundred tiles).
__global__ void tile_lists_kernel(const int* __restrict__ rmask, int nr, const int* __restrict__ cmask, int nc,
                                  int* __restrict__ list, int stride) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int a = 0, b = 0;
  for (int i = 0; i < nr; ++i)
    if (rmask[i]) list[a++] = i;
  for (int j = 0; j < nc; ++j)
    if (cmask[j]) list[stride + b++] = j;
  list[2 * stride] = a;
  list[2 * stride + 1] = b;
}

template <typename T>
__global__ __launch_bounds__(256) void col_tile_mask_kernel(const T* __restrict__ A, int64_t lda, int nrows,
                                                            int ncols, int* __restrict__ mask) {
  const int c0 = blockIdx.x * 128;
  const int ce = min(c0 + 128, ncols);
  int any = 0;
  for (int c = c0; c < ce; ++c)
    for (int r = threadIdx.x; r < nrows; r += 256) any |= (A[r + (int64_t)c * lda] != T(0));
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) mask[blockIdx.x] = any;
}

// A12 <- L^{-1} A12 with L unit lower (r x r).  Blocked substitution as LAPACK's
// dtrsm: X1 = L11^{-1} B1, B2 -= L21 X1, X2 = L22^{-1} B2, recursively; only the
// leaves (at most 64 rows) use an explicit inverse, whose error grows with the
// leaf's condition number.
constexpr int TRSM_LEAF = 64;

// L X = B in place for a unit-lower L of order 32 and B of 32 rows: one lane per
// column of B, 64 columns per wave.  The lane keeps the rows not yet final in a
// window y[0..31] of registers: at step k, y[0] = B(k) is final (stored), and every
// other entry is updated and shifted down, y[i-1] = y[i] - L(k+i, k) y[0] -- the
// indices into y are constants, so the step loop needs no unrolling and no dynamic
// register indexing; entries past the rows still open compute values that are
// never used.  L is staged in LDS once and read with broadcast loads.  The order of
// operations on every B(i) is the column-oriented substitution of LAPACK's xTRSM
// ('L', 'L', 'N', 'U').  cmask (optional): 128-column tiles of B that are
// structurally zero stay untouched; a batch's ZeroRhs skips the 64-column groups whose
// rows here are all zero.
constexpr int TRSM_FUSED = 32;

// Where a batch's right-hand sides are known to be zero (a batch of sparse right-hand
// sides such as unit vectors); the solves skip that work, whose results would be zero.
struct ZeroRhs {
  const int* cols = nullptr;   // device, nbat entries: system z is zero past column cols[z]
  const int* start = nullptr;  // device, nbat x ng: group g (columns 64g .. 64g+63) of system z
                               // is zero above row start[z ng + g] (lower solves only)
  int ng = 0;
  int row0 = 0;  // the sub-system's first row in the full system (rows of start)
};

template <typename T>
__global__ __launch_bounds__(64, 4) void trsm_leaf_kernel(const T* __restrict__ L, int64_t lda, T* __restrict__ B,
                                                       int64_t ldb, int ncols, const int* __restrict__ cmask,
                                                       int64_t sL, int64_t sB, ZeroRhs zr) {
  constexpr int R = TRSM_FUSED;
  L += (int64_t)blockIdx.y * sL;  // system blockIdx.y of a batch
  B += (int64_t)blockIdx.y * sB;
  __builtin_amdgcn_s_setprio(2);  // a serial chain on the critical path, beside the bulk GEMM
  __shared__ T Ls[R * R + R];  // column-major, Ls[i + R k] = L(i, k); R entries of padding read past the end
  const int c0 = blockIdx.x * 64;
  if (cmask && !cmask[c0 / 128]) return;  // uniform: the block lies in one 128-column tile
  if (zr.cols && c0 >= zr.cols[blockIdx.y]) return;
  if (zr.start && zr.start[(int)blockIdx.y * zr.ng + (int)blockIdx.x] >= zr.row0 + R) return;
  for (int e = threadIdx.x; e < R * R; e += 64) Ls[e] = L[(int64_t)(e % R) + (int64_t)(e / R) * lda];
  if (threadIdx.x < R) Ls[R * R + threadIdx.x] = T(0);
  const int c = c0 + threadIdx.x;
  const bool act = c < ncols;
  T* b = B + (int64_t)(act ? c : c0) * ldb;
  T y[R];
#pragma unroll
  for (int i = 0; i < R; ++i) y[i] = act ? b[i] : T(0);
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < R; ++k) {
    const T xk = y[0];
    if (act) b[k] = xk;
    const T* lk = Ls + k + R * k;  // lk[i] = L(k + i, k); i + k < 2 R stays inside Ls
#pragma unroll
    for (int i = 1; i < R; ++i) y[i - 1] = y[i] - lk[i] * xk;
  }
}

// C_z -= A_z B_z over a batch at strides (sa, sb, sc).  zr (may be NULL): B_z is zero
// past column zr->cols[z], and (with start) the rows of B, which are rows zr->row0 ..
// zr->row0 + k - 1 of the full system, are zero in the 64-column groups that start
// below them: the tiles of C that would only subtract zeros are skipped.
static int batched_sub(int prec, int m, int n, int k, const void* A, int64_t lda, int64_t sa, const void* B,
                       int64_t ldb, int64_t sb, void* C, int64_t ldc, int64_t sc, int nbat, const ZeroRhs* zr,
                       hipStream_t s) {
  if (!zr) return gemm_launch_batched(prec, GOP_SUB, m, n, k, A, lda, sa, B, ldb, sb, C, ldc, sc, nbat, s);
  TileSel sel;
  sel.nbat = nbat;
  sel.sa = sa;
  sel.sb = sb;
  sel.sc = sc;
  sel.zcols = zr->cols;
  sel.zstart = zr->start;
  sel.zng = zr->ng;
  sel.zthr = zr->row0 + k;
  return gemm_launch_sel(prec, GOP_SUB, m, n, k, A, lda, B, ldb, C, ldc, sel, s);
}

// nbat > 1: nbat systems at L + z sL, A12 + z sB (no masks; r a multiple of 32, so
// every leaf is a 32-row kernel).  zr (may be NULL): where the batch's right-hand sides
// are zero; the work on those parts, which stay zero, is skipped.
template <typename T>
int trsm_left_unit_lower(const T* L, int64_t lda, int r, T* A12, int64_t ld12, int ncols, LuWork& w, int gemm_prec,
                         const int* cmask = nullptr, int nbat = 1, int64_t sL = 0, int64_t sB = 0,
                         const ZeroRhs* zr = nullptr) {
  if (ncols <= 0 || r <= 0) return KAD_OK;
  if (nbat > 1 && cmask) return arg_error("trsm_left_unit_lower: masks with a batch");
  if (zr && (nbat < 2 || (zr->start && zr->ng < cdiv(ncols, 64))))
    return arg_error("trsm_left_unit_lower: zero right-hand-side parts need a batch and a group per 64 columns");
  if (r > TRSM_LEAF || r == 2 * TRSM_FUSED) {
    const int h = std::max(32, (r / 2) / 32 * 32);
    int rc = trsm_left_unit_lower<T>(L, lda, h, A12, ld12, ncols, w, gemm_prec, cmask, nbat, sL, sB, zr);
    if (rc) return rc;
    const int gp = gemm_prec == BF16 ? F32 : gemm_prec;
    rc = nbat > 1 ? batched_sub(gp, r - h, ncols, h, L + h, lda, sL, A12, ld12, sB, A12 + h, ld12, sB, nbat, zr,
                                w.stream)
                  : gemm_launch_masked(gp, GOP_SUB, r - h, ncols, h, L + h, lda, A12, ld12, A12 + h, ld12, nullptr,
                                       cmask, w.stream);
    if (rc) return rc;
    ZeroRhs z2;
    if (zr) {
      z2 = *zr;
      z2.row0 += h;
    }
    return trsm_left_unit_lower<T>(L + h + (int64_t)h * lda, lda, r - h, A12 + h, ld12, ncols, w, gemm_prec, cmask,
                                   nbat, sL, sB, zr ? &z2 : nullptr);
  }
  if (r == TRSM_FUSED) {
    trsm_leaf_kernel<T><<<dim3((unsigned)cdiv(ncols, 64), nbat), 64, 0, w.stream>>>(L, lda, A12, ld12, ncols, cmask,
                                                                                  sL, sB, zr ? *zr : ZeroRhs());
    KAD_HIP_CHECK(hipGetLastError());
    return KAD_OK;
  }
  // leaves of other orders use an explicit inverse (one system)
  if (nbat > 1) return arg_error("trsm_left_unit_lower: a batch needs orders that are multiples of 32");
  T* linv = (T*)w.tinv;
  // Column chunks so that the scratch holds r x chunk; with more than one chunk they
  // are multiples of 128 (the mask granularity), else the masks are not used.
  const int64_t cap = w.tmp ? (int64_t)(w.tmp_bytes / sizeof(T)) / r : 0;
  if (!linv || cap < 1) return KAD_ERR_ALLOC;
  int chunk = (int)std::min<int64_t>(cap, ncols);
  if (chunk < ncols && chunk >= 128) chunk = chunk / 128 * 128;
  if (chunk < ncols && chunk % 128 != 0) cmask = nullptr;
  int rc = trtri_rec<T>(L, lda, r, linv, r, w);
  if (rc) return rc;
  for (int c = 0; c < ncols; c += chunk) {
    const int nc = std::min(chunk, ncols - c);
    T* tmp = (T*)w.tmp;
    rc = gemm_launch_masked(gemm_prec == BF16 ? F32 : gemm_prec, GOP_SET, r, nc, r, linv, r, A12 + (int64_t)c * ld12,
                            ld12, tmp, r, nullptr, cmask ? cmask + c / 128 : nullptr, w.stream);
    if (rc) return rc;
    dim3 g((unsigned)cdiv(r, 256), nc);
    copy2d_kernel<T><<<g, 256, 0, w.stream>>>(tmp, r, A12 + (int64_t)c * ld12, ld12, r, nc);
    KAD_HIP_CHECK(hipGetLastError());
  }
  return KAD_OK;
}

// U X = B in place for a non-unit upper U of order 32 and B of 32 rows: the leaf of
// trsm_left_upper.  The rows are taken in reverse (U'(i, k) = U(31 - i, 31 - k) is
// lower), so the lane's register window works as in trsm_leaf_kernel; each step
// divides by the diagonal, then updates, as LAPACK's xTRSM ('L', 'U', 'N', 'N').
template <typename T>
__global__ __launch_bounds__(64, 4) void trsm_leaf_upper_kernel(const T* __restrict__ U, int64_t ldu,
                                                             T* __restrict__ B, int64_t ldb, int ncols, int64_t sU,
                                                             int64_t sB, const int* __restrict__ zcols) {
  constexpr int R = TRSM_FUSED;
  __builtin_amdgcn_s_setprio(2);
  if (zcols && (int)blockIdx.x * 64 >= zcols[blockIdx.y]) return;  // only zero columns
  U += (int64_t)blockIdx.y * sU;  // system blockIdx.y of a batch
  B += (int64_t)blockIdx.y * sB;
  __shared__ T Us[R * R + R];  // Us[i + R k] = U(R-1-i, R-1-k); R entries of padding read past the end
  for (int e = threadIdx.x; e < R * R; e += 64) {
    const int i = e % R, k = e / R;
    Us[e] = U[(int64_t)(R - 1 - i) + (int64_t)(R - 1 - k) * ldu];
  }
  if (threadIdx.x < R) Us[R * R + threadIdx.x] = T(0);
  const int c0 = blockIdx.x * 64;
  const int c = c0 + threadIdx.x;
  const bool act = c < ncols;
  T* b = B + (int64_t)(act ? c : c0) * ldb;
  T y[R];
#pragma unroll
  for (int i = 0; i < R; ++i) y[i] = act ? b[R - 1 - i] : T(0);
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < R; ++k) {
    const T* uk = Us + k + R * k;  // uk[i] = U'(k + i, k)
    const T xk = y[0] / uk[0];
    if (act) b[R - 1 - k] = xk;
#pragma unroll
    for (int i = 1; i < R; ++i) y[i - 1] = y[i] - uk[i] * xk;
  }
}

// B <- U^{-1} B with U non-unit upper (r x r, r a multiple of 32): X2 = U22^{-1} B2,
// B1 -= U12 X2, X1 = U11^{-1} B1, recursively down to 32-row leaves.  nbat systems at
// U + z sU, B + z sB; zcols (device, nbat entries, may be NULL): B_z is zero past
// column zcols[z], and those columns are skipped.
template <typename T>
int trsm_left_upper(const T* U, int64_t ldu, int r, T* B, int64_t ldb, int ncols, LuWork& w, int gemm_prec, int nbat,
                    int64_t sU, int64_t sB, const int* zcols = nullptr) {
  if (ncols <= 0 || r <= 0) return KAD_OK;
  if (r % TRSM_FUSED != 0) return arg_error("trsm_left_upper: order not a multiple of 32");
  if (r == TRSM_FUSED) {
    trsm_leaf_upper_kernel<T><<<dim3((unsigned)cdiv(ncols, 64), nbat), 64, 0, w.stream>>>(U, ldu, B, ldb, ncols, sU,
                                                                                        sB, zcols);
    KAD_HIP_CHECK(hipGetLastError());
    return KAD_OK;
  }
  const int h = std::max(32, (r / 2) / 32 * 32);
  int rc =
      trsm_left_upper<T>(U + h + (int64_t)h * ldu, ldu, r - h, B + h, ldb, ncols, w, gemm_prec, nbat, sU, sB, zcols);
  if (rc) return rc;
  ZeroRhs zc;
  zc.cols = zcols;
  rc = batched_sub(gemm_prec, h, ncols, r - h, U + (int64_t)h * ldu, ldu, sU, B + h, ldb, sB, B, ldb, sB, nbat,
                   zcols ? &zc : nullptr, w.stream);
  if (rc) return rc;
  return trsm_left_upper<T>(U, ldu, h, B, ldb, ncols, w, gemm_prec, nbat, sU, sB, zcols);
}

// B <- A^{-1} B for nbat factored matrices of order n (n a multiple of 32) whose
// right-hand sides are already in pivoted row order (B_z rows permuted by perm_z):
// the unit-lower, then the upper triangular solve, in FP64.  zcols (device, nbat
// entries, may be NULL): B_z is zero past its first zcols[z] columns, which are skipped;
// zstart (device, nbat x zng, may be NULL, needs zcols): columns 64g .. 64g+63 of B_z are
// zero above row zstart[z zng + g], which the unit-lower solve skips.
int lu_solve_permuted_batched(const double* LU, int64_t lda, int n, int nbat, int64_t sA, double* B, int64_t ldb,
                              int64_t sB, int nrhs, LuWork& w, const int* zcols, const int* zstart, int zng) {
  if (n % TRSM_FUSED != 0 || nbat < 1) return arg_error("lu_solve_permuted_batched: order or batch");
  if (zstart && !zcols) return arg_error("lu_solve_permuted_batched: row starts without column counts");
  ZeroRhs zr;
  zr.cols = zcols;
  zr.start = zstart;
  zr.ng = zng;
  int rc = trsm_left_unit_lower<double>(LU, lda, n, B, ldb, nrhs, w, F64, nullptr, nbat, sA, sB,
                                        zcols && nbat > 1 ? &zr : nullptr);
  if (rc) return rc;
  return trsm_left_upper<double>(LU, lda, n, B, ldb, nrhs, w, F64, nbat, sA, sB, zcols);
}

// Row interchanges of pivots [k1, k2) (at most 512) applied to COLS columns per
// workgroup through LDS.  Slots 0 .. nsw-1 hold rows k1 .. k2-1; slot nsw + j holds the
// pivot row of step j when that row lies below k2 and step j is its first occurrence.
// All slots are loaded in parallel, the interchange sequence is replayed in LDS (one
// thread per column, LAPACK order), and every touched row is written back: the
// result equals the sequential interchanges.
template <typename T, int NSW, int COLS>
__global__ __launch_bounds__(256) void laswp_lds_kernel(T* __restrict__ A, int64_t lda, int c0, int c1,
                                                        const int* __restrict__ ipiv, int k1, int k2, int64_t sA,
                                                        int sP) {
  A += (int64_t)blockIdx.y * sA;  // matrix blockIdx.y of a batch
  ipiv += (int64_t)blockIdx.y * sP;
  __shared__ double vals[2 * NSW * COLS];  // [slot][column]
  __shared__ int slot_of[NSW];
  __shared__ int prow[NSW];                // row of slot nsw + j, or -1
  __shared__ int piv[NSW];                 // ipiv[k1 .. k2), staged once (the searches below read it)
  const int nsw = k2 - k1;                 // <= NSW (checked by the launcher)
  const int t = threadIdx.x;
  const int cb = c0 + blockIdx.x * COLS;
  const int nc = min(COLS, c1 - cb);
  for (int j = t; j < nsw; j += 256) piv[j] = ipiv[k1 + j];
  __syncthreads();
  for (int j = t; j < nsw; j += 256) {
    const int p = piv[j];
    if (p < k2) {
      slot_of[j] = p - k1;
      prow[j] = -1;
    } else {
      int first = j;
      for (int q = 0; q < j; ++q)
        if (piv[q] == p) {
          first = q;
          break;
        }
      slot_of[j] = nsw + first;
      prow[j] = (first == j) ? p : -1;
    }
  }
  __syncthreads();
  // load: rows k1 .. k2-1 (coalesced along rows) and the first-occurrence pivot rows
  for (int e = t; e < nsw * COLS; e += 256) {
    const int r = e % nsw, c = e / nsw;
    if (c < nc) vals[r * COLS + c] = (double)A[(k1 + r) + (int64_t)(cb + c) * lda];
  }
  for (int e = t; e < nsw * COLS; e += 256) {
    const int j = e % nsw, c = e / nsw;
    const int p = prow[j];
    if (c < nc && p >= 0) vals[(nsw + j) * COLS + c] = (double)A[p + (int64_t)(cb + c) * lda];
  }
  __syncthreads();
  if (t < nc)
    for (int j = 0; j < nsw; ++j) {
      const int b = slot_of[j];
      if (b != j) {
        const double x = vals[j * COLS + t];
        vals[j * COLS + t] = vals[b * COLS + t];
        vals[b * COLS + t] = x;
      }
    }
  __syncthreads();
  for (int e = t; e < nsw * COLS; e += 256) {
    const int r = e % nsw, c = e / nsw;
    if (c < nc) A[(k1 + r) + (int64_t)(cb + c) * lda] = (T)vals[r * COLS + c];
  }
  for (int e = t; e < nsw * COLS; e += 256) {
    const int j = e % nsw, c = e / nsw;
    const int p = prow[j];
    if (c < nc && p >= 0) A[p + (int64_t)(cb + c) * lda] = (T)vals[(nsw + j) * COLS + c];
  }
}

template <typename T>
int laswp(T* A, int64_t lda, int c0, int c1, const int* ipiv, int k1, int k2, LuWork& w,
          const LuBatch& bt = LuBatch()) {
  if (c1 <= c0 || k2 <= k1) return KAD_OK;
  const int nsw = k2 - k1;
  const unsigned nz = (unsigned)bt.nbat;
  // 8 (4) columns per workgroup: ~33 KB of LDS, so the workgroups fit beside the
  // concurrent update's GEMM workgroups instead of waiting for a whole CU
  if (nsw <= 256) {
    laswp_lds_kernel<T, 256, 8><<<dim3((unsigned)cdiv(c1 - c0, 8), nz), 256, 0, w.stream>>>(A, lda, c0, c1, ipiv, k1,
                                                                                           k2, bt.sA, bt.sP);
  } else if (nsw <= 512) {
    laswp_lds_kernel<T, 512, 4><<<dim3((unsigned)cdiv(c1 - c0, 4), nz), 256, 0, w.stream>>>(A, lda, c0, c1, ipiv, k1,
                                                                                           k2, bt.sA, bt.sP);
  } else {
    laswp_kernel<T><<<dim3((unsigned)cdiv(c1 - c0, 256), nz), 256, 0, w.stream>>>(A, lda, c0, c1, ipiv, k1, k2, bt.sA,
                                                                                bt.sP);
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// ---------------------------------------------------------------------------
// Base panel without atomics or device-scope fences.  Inside the panel rows are
// never moved: a logical -> physical row map (double-buffered per step) tracks
// the interchanges, and every block re-reduces the previous step's per-block
// maxima, so all communication between blocks goes through kernel boundaries.
// The interchange sequence is LAPACK's (first index of max |a| among the
// logical rows >= j); the panel rows are permuted physically once at the end.
// ---------------------------------------------------------------------------
constexpr int FF_ROWS = 64;  // logical rows per block (x 4 column groups = 256 threads)

template <typename T>
__device__ void ff_reduce_partials(const double* __restrict__ pv, const int* __restrict__ pi, int np, double* sv,
                                   int* si, double& bv, int& bi) {
  double v = -1.0;
  int id = -1;
  for (int b = threadIdx.x; b < np; b += blockDim.x) {
    const double ov = pv[b];
    const int oi = pi[b];
    if (oi >= 0 && (id < 0 || ov > v || (ov == v && oi < id))) {
      v = ov;
      id = oi;
    }
  }
  block_argmax<T>(v, id, sv, si, bv, bi);
}

// Partial maxima of column c0 over logical rows [c0, n) (identity map).
template <typename T>
__global__ __launch_bounds__(256) void ff_first_kernel(const T* __restrict__ A, int64_t lda, int n, int c0,
                                                       int* __restrict__ map_out, double* __restrict__ pv,
                                                       int* __restrict__ pi) {
  __shared__ double sv[256];
  __shared__ int si[256];
  const int i = c0 + blockIdx.x * 256 + threadIdx.x;
  double v = -1.0;
  int id = -1;
  if (i < n) {
    map_out[i] = i;
    v = fabs((double)A[i + (int64_t)c0 * lda]);
    id = i;
  }
  double bv;
  int bi;
  block_argmax<T>(v, id, sv, si, bv, bi);
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = bv;
    pi[blockIdx.x] = bi;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void ff_step_kernel(T* __restrict__ A, int64_t lda, int n, int c0, int j, int c1,
                                                      const int* __restrict__ map_in, int* __restrict__ map_out,
                                                      const double* __restrict__ pv_in, const int* __restrict__ pi_in,
                                                      int np_in, double* __restrict__ pv_out, int* __restrict__ pi_out,
                                                      int* __restrict__ ipiv, int* __restrict__ info) {
  __shared__ double sv[256];
  __shared__ int si[256];
  __shared__ T urow[PANEL_BASE];
  __shared__ T lcol[FF_ROWS];
  __shared__ int pshared;
  double bv;
  int p;
  ff_reduce_partials<T>(pv_in, pi_in, np_in, sv, si, bv, p);
  if (p < 0) p = j;
  const int phys_piv = map_in[p];  // physical row that becomes logical row j
  const int phys_j = map_in[j];    // physical row that moves to logical row p
  const int w = c1 - j;
  for (int c = threadIdx.x; c < w; c += blockDim.x) urow[c] = A[phys_piv + (int64_t)(j + c) * lda];
  if (blockIdx.x == 0) {
    // logical rows [c0, j) are final: carry them into the other map buffer
    for (int k = threadIdx.x; k < j - c0; k += blockDim.x) map_out[c0 + k] = map_in[c0 + k];
    if (threadIdx.x == 0) {
      ipiv[j] = p;
      map_out[j] = phys_piv;
      if (bv == 0.0 && *info == 0) *info = j + 1;
    }
  }
  const int tx = threadIdx.x % FF_ROWS, ty = threadIdx.x / FF_ROWS;
  const int i = j + 1 + blockIdx.x * FF_ROWS + tx;  // logical row
  int phys = -1;
  if (i < n) phys = (i == p) ? phys_j : map_in[i];
  __syncthreads();
  const T piv = urow[0];
  if (ty == 0) {
    T l = T(0);
    if (i < n) {
      l = A[phys + (int64_t)j * lda];
      if (piv != T(0)) l = l / piv;
      A[phys + (int64_t)j * lda] = l;
      map_out[i] = phys;
    }
    lcol[tx] = l;
  }
  __syncthreads();
  double v = -1.0;
  int id = -1;
  if (i < n) {
    const T l = lcol[tx];
    T* ai = A + phys;
    for (int c = 1 + ty; c < w; c += 4) {
      T x = ai[(int64_t)(j + c) * lda];
      x -= l * urow[c];
      ai[(int64_t)(j + c) * lda] = x;
      if (c == 1) {
        v = fabs((double)x);
        id = i;
      }
    }
  }
  if (j + 1 >= c1) return;  // uniform: no partials after the last column
  int bi;
  block_argmax<T>(v, id, sv, si, bv, bi);
  if (threadIdx.x == 0) {
    pv_out[blockIdx.x] = bv;
    pi_out[blockIdx.x] = bi;
  }
  (void)pshared;
}

// Physical permutation of the panel rows: tmp[i - c0][c] = A[map[i]][c], then back.
template <typename T>
__global__ void ff_gather_kernel(const T* __restrict__ A, int64_t lda, int n, int c0, int c1,
                                 const int* __restrict__ map, T* __restrict__ tmp) {
  const int i = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int c = c0 + blockIdx.y;
  if (i < n && c < c1) tmp[(i - c0) + (int64_t)(c - c0) * (n - c0)] = A[map[i] + (int64_t)c * lda];
}
template <typename T>
__global__ void ff_scatter_kernel(T* __restrict__ A, int64_t lda, int n, int c0, int c1, const T* __restrict__ tmp) {
  const int i = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int c = c0 + blockIdx.y;
  if (i < n && c < c1) A[i + (int64_t)c * lda] = tmp[(i - c0) + (int64_t)(c - c0) * (n - c0)];
}

// ---------------------------------------------------------------------------
// Persistent base panel: one launch factors columns [c0, c1) (c1 - c0 <= W) of rows
// [c0, n).  Workgroup g owns logical rows [c0 + g rpw, c0 + (g + 1) rpw) of the block,
// one per thread, and keeps them in registers for the whole launch; interchanges move
// row contents between owners.  Per column j:
//   1. every workgroup publishes its pivot candidate (logical row and the candidate's
//      row of the block); the owner of row j publishes row j;
//   2. every workgroup reads all candidates with their rows in one round trip, each
//      word flagged with its column (no grid barrier: a word is current when its flag
//      is), reduces them itself -- LAPACK's choice, the first index of max |a| -- and
//      swaps, scales and updates its rows with the per-column kernel's operations.
// At the end the interchanges are applied to the panel's other columns [k0, c0) and
// [c1, k1) (exact: those columns are not read in between).
// The exchange's cost grows with the number of workgroups, so each one holds 256 rows,
// and it fits beside a GEMM workgroup on one CU (the trailing update runs
// concurrently): 144 VGPRs (FP64) and 19 KB of LDS, 4 waves.
// The grid must be resident at once: the host launches at most one workgroup per
// CU, and nothing else on the device waits on this kernel.  Every wait is bounded: a
// workgroup that reaches the spin limit raises the abort flag (sync[2]), every waiter
// leaves on it, and the host reports the failure.
// ---------------------------------------------------------------------------
constexpr int FFP_MAX_GRID = 64;
constexpr int FFP_THREADS = 256;

// Exchange slots.  Every 64-bit slot holds one 32-bit data word and, in its high half,
// the flag of the column that wrote it (launch sequence x 32 + column).  Slots are
// read and written as agent-scope relaxed atomics (sc1: written through to, and read
// from, the level the XCDs share), so a reader that sees the flag it expects holds
// that column's word: each slot validates itself, no slot is ordered against another,
// and no device-scope fence (a write-back and invalidate of the XCD's whole L2, the
// concurrent GEMM's working set) or arrival counter is needed.  A column's slots are
// one record per workgroup (its candidate's logical row, for FP64 a pad word that
// keeps the values 8-byte aligned, then the W values of that row, one word each for
// FP32 and two for FP64) and, after the G records, the W values of row j.  Every
// record and row j carry all W values (those of columns past the panel's end are
// never read).  A column's stride covers whole 256-slot strips, so every thread reads
// NU slots unconditionally and keeps those below G records + row j.
template <typename T, int W>
struct FfpSlots {
  static constexpr int VW = (int)(sizeof(T) / 4);        // words per value
  static constexpr int VOFF = VW;                        // first value word of a record
  static constexpr int REC = VOFF + VW * W;              // one workgroup's record
  static constexpr int USED = FFP_MAX_GRID * REC + VW * W;  // G <= FFP_MAX_GRID records, row j
  static constexpr int NU = (USED + 255) / 256;          // slots per thread (256 threads)
  static constexpr int COL = NU * 256;                   // one column's stride
  // where logical slot e (strip e / 256, thread e % 256) lies: wave-major, so one
  // thread's NU slots are 512 bytes apart (immediate offsets from one or two bases)
  // while each wave's load of a strip stays 64 consecutive slots
  static __device__ __forceinline__ int pos(int e) {
    const int t = e & 255;
    return (t >> 6) * (NU * 64) + (e >> 8) * 64 + (t & 63);
  }
};

__device__ __forceinline__ void slot_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long slot_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ unsigned word_of(T v, int k) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    return k ? (unsigned)(b >> 32) : (unsigned)b;
  } else {
    return __builtin_bit_cast(unsigned, v);
  }
}

__device__ __forceinline__ void argmax_take(double& v, int& i, double ov, int oi) {
  if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) {
    v = ov;
    i = oi;
  }
}

// One step of a wave-level argmax: combine with the lane that DPP control CTRL reads
// (a permutation inside each 16-lane row, so every source lane is valid).
template <int CTRL>
__device__ __forceinline__ void argmax_dpp(double& v, int& i) {
  const long long b = __double_as_longlong(v);
  const int hi = (int)(b >> 32), lo = (int)b;
  const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const int oi = __builtin_amdgcn_update_dpp(i, i, CTRL, 0xF, 0xF, false);
  argmax_take(v, i, __longlong_as_double(((long long)ohi << 32) | (unsigned)olo), oi);
}

// Max |a| with the smallest index among ties (invalid entries carry index -1) over
// one wave; the result reaches every lane.  Quad swaps, then half-row and row mirrors
// (every lane of a 16-lane row ends with the row's best), then the four rows' results
// are read from lanes 0, 16, 32, 48.  The selection is a total order, so the result
// does not depend on the reduction order.
__device__ __forceinline__ void argmax_wave(double& v, int& i) {
  argmax_dpp<0xB1>(v, i);   // quad_perm [1,0,3,2]
  argmax_dpp<0x4E>(v, i);   // quad_perm [2,3,0,1]
  argmax_dpp<0x141>(v, i);  // row_half_mirror
  argmax_dpp<0x140>(v, i);  // row_mirror
  const long long b = __double_as_longlong(v);
  const int hi = (int)(b >> 32), lo = (int)b;
  double wv = -1.0;
  int wi = -1;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int rhi = __builtin_amdgcn_readlane(hi, 16 * r), rlo = __builtin_amdgcn_readlane(lo, 16 * r);
    argmax_take(wv, wi, __longlong_as_double(((long long)rhi << 32) | (unsigned)rlo),
                __builtin_amdgcn_readlane(i, 16 * r));
  }
  v = wv;
  i = wi;
}

// The same over the 256-thread block, through 4 entries of LDS (sv / si) and one
// barrier.  Every thread of the block must call it, and the caller must pass another
// barrier before the next call writes sv / si.
__device__ __forceinline__ void argmax_256(double& v, int& i, double* sv, int* si) {
  argmax_wave(v, i);
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = v;
    si[threadIdx.x >> 6] = i;
  }
  __syncthreads();
  v = -1.0;
  i = -1;
#pragma unroll
  for (int q = 0; q < FFP_THREADS / 64; ++q) argmax_take(v, i, sv[q], si[q]);
}

template <typename T, int W>
__global__ __launch_bounds__(FFP_THREADS) void ffp_kernel(T* __restrict__ A, int64_t lda, int n, int c0, int c1,
                                                          int k0, int k1, int rpw,
                                                          unsigned long long* __restrict__ xbuf, unsigned flag0,
                                                          int* __restrict__ ipiv, int* __restrict__ info,
                                                          int* __restrict__ sync, int spin_limit, int64_t sA, int sP,
                                                          int64_t sX) {
  static_assert(FFP_MAX_GRID <= 64, "one candidate per lane of a wave");
  static_assert(FFP_THREADS == 256, "256-slot strips");
  static_assert((W & (W - 1)) == 0, "rotation mask");
  using S = FfpSlots<T, W>;
  constexpr int NU = S::NU;
  // issue priority over the GEMM waves that share the SIMDs: every column waits for
  // the slowest workgroup, and the GEMM's MFMA chains leave issue slots to spare
  __builtin_amdgcn_s_setprio(3);
  // this column's slots, data words only: G records, then row j
  __shared__ __align__(16) unsigned xim[S::COL];
  __shared__ double sv[FFP_THREADS / 64];
  __shared__ int si[FFP_THREADS / 64];
  __shared__ int s_piv[W];
  __shared__ int s_rows[2 * W];
  __shared__ int s_src[2 * W];
  __shared__ int s_nrows;
  __shared__ __align__(16) T stage[2][W];  // the candidate row and row j, for their waves
  // matrix blockIdx.y of a batch: its own rows, interchanges, exchange slots and info
  // word; the workgroups of one matrix (blockIdx.x) exchange only among themselves
  A += (int64_t)blockIdx.y * sA;
  ipiv += (int64_t)blockIdx.y * sP;
  xbuf += (int64_t)blockIdx.y * sX;
  info += blockIdx.y;
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const T* const jrow = reinterpret_cast<const T*>(xim + G * S::REC);
  const int r0 = c0 + g * rpw;
  const int nr = max(0, min(rpw, n - r0));
  const int w = c1 - c0;
  // this thread's row of the block, in registers for the whole launch (one row per
  // thread: the host launches at most FFP_THREADS rows per workgroup); columns past
  // the panel's end hold zeros
  // (rotated by one position per column: at column jj, row[q] holds column
  // (q + jj) mod W, so the active column is row[0] and every index into the array is
  // a compile-time constant)
  const int r = tid;
  const bool own = r < nr;
  T row[W];
#pragma unroll
  for (int c = 0; c < W; ++c) row[c] = own && c < w ? A[(int64_t)(r0 + r) + (int64_t)(c0 + c) * lda] : T(0);
  // local candidate of the first column; later columns' come out of the update
  double v = own ? fabs((double)row[0]) : -1.0;
  int id = own ? r0 + r : -1;
  bool live = true;  // uniform: false once a wait was abandoned
  int jfail = w;     // first column not factored (after an abandoned wait)
  for (int jj = 0; jj < w; ++jj) {
    const int j = c0 + jj;
    // 1. the workgroup's candidate among its logical rows >= j, then publish its
    //    record (logical row -1: none) and, from row j's owner, row j
    argmax_256(v, id, sv, si);
    const unsigned flag = flag0 + (unsigned)jj;
    const unsigned long long tag = (unsigned long long)flag << 32;
    unsigned long long* const col = xbuf + (size_t)jj * S::COL;
    // the owning lane stages the row in LDS, then its wave stores the row's words
    // (one slot per lane) without a workgroup barrier
    const int lane = tid & 63;
    const bool cand_wave = id >= 0 && (tid >> 6) == ((id - r0) >> 6);
    const bool rowj_wave = j >= r0 && j < r0 + nr && (tid >> 6) == ((j - r0) >> 6);
    if (cand_wave && tid == id - r0) {
#pragma unroll
      for (int q = 0; q < W; ++q) stage[0][(q + jj) & (W - 1)] = row[q];
    }
    if (rowj_wave && tid == j - r0) {
#pragma unroll
      for (int q = 0; q < W; ++q) stage[1][(q + jj) & (W - 1)] = row[q];
    }
    if (cand_wave || rowj_wave) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    static_assert(S::VW * W <= 64, "one word per lane");
    if (cand_wave) {
      if (lane < S::VOFF) slot_store(col + S::pos(g * S::REC + lane), tag | (lane == 0 ? (unsigned)id : 0u));
      if (lane < S::VW * W)
        slot_store(col + S::pos(g * S::REC + S::VOFF + lane),
                   tag | reinterpret_cast<const unsigned*>(stage[0])[lane]);
    } else if (id < 0 && tid < S::REC) {
      slot_store(col + S::pos(g * S::REC + tid), tag | (tid == 0 ? (unsigned)id : 0u));
    }
    if (rowj_wave && lane < S::VW * W)
      slot_store(col + S::pos(G * S::REC + lane), tag | reinterpret_cast<const unsigned*>(stage[1])[lane]);
    // 2. every record and row j, in one round trip: a thread loads its NU slots
    //    (strip u, logical slot u * 256 + tid) at once, then reads again those below the used
    //    slots until they carry this column's flag.  Bounded: at the spin limit the
    //    workgroup raises the abort flag, and every waiter leaves on it; the host then
    //    reports the failure.
    const int nused = G * S::REC + S::VW * W;
    const int nv = min(NU, max(0, (nused - tid + 255) >> 8));  // strips below nused
    const unsigned valid = nv >= 32 ? ~0u : (1u << nv) - 1u;
    const unsigned long long* const colt = col + (tid >> 6) * (NU * 64) + lane;  // strip u: + 64 u
    // every word goes to the LDS image as it arrives; only the pending mask is carried
    unsigned pend = 0;
    {
      unsigned long long xs[NU];
#pragma unroll
      for (int u = 0; u < NU; ++u) xs[u] = slot_load(colt + u * 64);
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        xim[u * 256 + tid] = (unsigned)xs[u];
        pend |= (unsigned)((unsigned)(xs[u] >> 32) != flag) << u;
      }
    }
    pend &= valid;
    int bad = 0;
    for (int it = 0; pend; ++it) {
      if (it >= spin_limit) {
        __hip_atomic_store(&sync[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bad = 1;
        break;
      }
      if ((it & 255) == 255 && __hip_atomic_load(&sync[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        bad = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      // again the groups of 4 strips that hold a pending slot
#pragma unroll
      for (int u0 = 0; u0 < NU; u0 += 4) {
        if ((pend >> u0) & 15u) {
          constexpr int GU = 4;
          unsigned long long y[GU];
#pragma unroll
          for (int q = 0; q < GU; ++q)
            if (u0 + q < NU) y[q] = slot_load(colt + (u0 + q) * 64);
          unsigned nb = 0;
#pragma unroll
          for (int q = 0; q < GU; ++q)
            if (u0 + q < NU) {
              xim[(u0 + q) * 256 + tid] = (unsigned)y[q];
              nb |= (unsigned)((unsigned)(y[q] >> 32) != flag) << q;
            }
          pend = (pend & ~(15u << u0)) | ((nb << u0) & valid);
        }
      }
    }
    if (__syncthreads_or(bad)) {  // uniform; the host sees sync[2] and reports the failure
      live = false;
      jfail = jj;
      break;
    }
    // 3. the pivot: LAPACK's choice, the first logical row of max |a| (a candidate's
    //    |a| is its row's entry in this column); every wave reduces the G <= 64
    //    candidates itself
    v = -1.0;
    id = -1;
    if (lane < G && (int)xim[lane * S::REC] >= 0) {
      v = fabs((double)reinterpret_cast<const T*>(xim + lane * S::REC + S::VOFF)[jj]);
      id = (int)xim[lane * S::REC];
    }
    argmax_wave(v, id);
    int p = id;
    if (p < j || p >= n) {
      // only reachable with corrupt candidates: keep every index in range and fail
      p = j;
      if (tid == 0) __hip_atomic_store(&sync[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const T* urow = reinterpret_cast<const T*>(xim + ((p - c0) / rpw) * S::REC + S::VOFF);  // p's owner's
    if (tid == 0) s_piv[jj] = p;
    if (g == 0 && tid == 0) {
      ipiv[j] = p;
      if (v == 0.0 && *info == 0) *info = j + 1;
    }
    // 4. the interchange: row j's owner takes the pivot row, row p's owner row j
    //    (registers of one thread each: no barrier)
    if (p != j) {
      if (j >= r0 && j < r0 + nr && tid == j - r0) {
#pragma unroll
        for (int q = 0; q < W; ++q) row[q] = urow[(q + jj) & (W - 1)];
      } else if (p >= r0 && p < r0 + nr && tid == p - r0) {
#pragma unroll
        for (int q = 0; q < W; ++q) row[q] = jrow[(q + jj) & (W - 1)];
      }
    }
    // 5. rows below j: the multiplier (column jj: row[0]), the rank-1 update of
    //    columns jj < c < W (row[q], q = c - jj: no wrap; columns past the panel's end
    //    are never written back) with the pivot row, and the candidate of the next
    //    column (row[1]); then every row rotates by one position
    const bool below = own && r0 + r > j;
    const T* const uj = urow + jj;  // uj[q]: the pivot row's column jj + q
    const T piv = uj[0];
    v = -1.0;
    id = -1;
    if (below) {
      T l = row[0];
      if (piv != T(0)) l = l / piv;
      row[0] = l;
#pragma unroll
      for (int q = 1; q < W; ++q)
        if (q < W - jj) row[q] = row[q] - l * uj[q];
      if (jj + 1 < w) {
        v = fabs((double)row[1]);
        id = r0 + r;
      }
    }
    {
      const T t = row[0];
#pragma unroll
      for (int q = 0; q < W - 1; ++q) row[q] = row[q + 1];
      row[W - 1] = t;
    }
  }
  if (!live) {
    // the columns not factored get identity interchanges: later kernels index rows
    // with ipiv, which must stay within the matrix (the factors are reported invalid)
    if (g == 0)
      for (int jj = jfail + tid; jj < w; jj += FFP_THREADS) ipiv[c0 + jj] = c0 + jj;
    return;
  }
  if (own) {  // after w columns, row[q] holds column (q + w) mod W
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const int c = (q + w) & (W - 1);
      if (c < w) A[(int64_t)(r0 + r) + (int64_t)(c0 + c) * lda] = row[q];
    }
  }
  __syncthreads();  // s_piv of the last column
  // the interchanges on the panel's other columns: the affected rows are rows
  // c0 .. c1-1 and the pivot rows (2 w entries, duplicates harmless: they carry the
  // same value); entry k's content comes from the row found by undoing the
  // interchanges in reverse order
  if (tid < 2 * w) {
    const int r = tid < w ? c0 + tid : s_piv[tid - w];
    int x = r;
    for (int jj = w - 1; jj >= 0; --jj) {
      const int a = c0 + jj, b = s_piv[jj];
      x = (x == a) ? b : ((x == b) ? a : x);
    }
    s_rows[tid] = r;
    s_src[tid] = x;
  }
  if (tid == 0) s_nrows = 2 * w;
  __syncthreads();
  // (column, affected row) pairs, columns split over the workgroups: every thread
  // gathers up to FFP_TAIL of them into registers, the workgroup synchronises, then
  // every thread scatters them -- one round trip per batch, and no row of a column is
  // written before all of that column's reads (a column stays in one workgroup)
  constexpr int FFP_TAIL = 16;
  const int nrw = s_nrows, nleft = c0 - k0, nother = (c0 - k0) + (k1 - c1);
  const int cpg = (int)cdiv(nother, G);  // columns of this workgroup: [g cpg, (g + 1) cpg)
  const int qa = min(nother, g * cpg), qe = min(nother, qa + cpg);
  const int npairs = (qe - qa) * nrw;
  const int step = ((FFP_TAIL * FFP_THREADS) / nrw) * nrw;  // whole columns per batch (nrw <= 2 W)
  for (int b0 = 0; b0 < npairs; b0 += step) {
    const int bend = min(npairs, b0 + step);
    T val[FFP_TAIL];
#pragma unroll
    for (int u = 0; u < FFP_TAIL; ++u) {
      const int e = b0 + u * FFP_THREADS + tid;
      if (e < bend) {
        const int q = qa + e / nrw, k = e - (e / nrw) * nrw;
        const int col = q < nleft ? k0 + q : c1 + (q - nleft);
        val[u] = A[(int64_t)s_src[k] + (int64_t)col * lda];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < FFP_TAIL; ++u) {
      const int e = b0 + u * FFP_THREADS + tid;
      if (e < bend) {
        const int q = qa + e / nrw, k = e - (e / nrw) * nrw;
        const int col = q < nleft ? k0 + q : c1 + (q - nleft);
        A[(int64_t)s_rows[k] + (int64_t)col * lda] = val[u];
      }
    }
    __syncthreads();
  }
}

// Exchange slots of the persistent base panel for one matrix (W columns, FP64
// records), in 64-bit words, and the bytes of the whole buffer (one region per
// matrix of a batch, LU_MAX_BATCH regions).
constexpr int64_t ffp_matrix_words() { return (int64_t)PANEL_BASE * FfpSlots<double, PANEL_BASE>::COL; }
size_t ffp_bytes() { return (size_t)LU_MAX_BATCH * ffp_matrix_words() * sizeof(unsigned long long); }

// Launch shape of the persistent base panel for m rows of element type T: grid and
// rows per workgroup, or false when it does not fit (the per-column kernels run).
template <typename T>
bool ffp_shape(int m, int ncu, int& grid, int& rpw) {
  // one row per thread, held in registers: a second row for a few threads would
  // double the latency of every column step for the whole workgroup
  grid = (int)cdiv(m, FFP_THREADS);
  if (m < 1 || grid > std::min(ncu, FFP_MAX_GRID)) return false;
  rpw = (int)cdiv(m, grid);
  grid = (int)cdiv(m, rpw);
  return true;
}

// One persistent base panel: columns [c0, c1) of rows [c0, n), interchanges applied
// across the panel's columns [k0, k1).
// A batch launches the workgroups of every matrix together (grid y); all of them must
// be resident at once, so a batch holds at most one workgroup per CU in total.
template <typename T>
int ffp_launch(T* A, int64_t lda, int n, int c0, int c1, int k0, int k1, int* ipiv, LuWork& w,
               const LuBatch& bt = LuBatch()) {
  int grid, rpw;
  if (!ffp_shape<T>(n - c0, w.ncu, grid, rpw) || rpw > FFP_THREADS || c1 - c0 > PANEL_BASE || c1 > n)
    return arg_error("ffp_launch: panel shape");
  if (bt.nbat < 1 || bt.nbat > LU_MAX_BATCH || grid * bt.nbat > w.ncu || (bt.nbat > 1 && !bt.info))
    return arg_error("ffp_launch: batch does not fit the CUs");
  // flags: PANEL_BASE per launch, never one the slots still hold (the buffer starts
  // zeroed and the sequence at 1; before the 32-bit flags wrap it is cleared again)
  if (++w.ffp_seq >= (1u << 26)) {
    KAD_HIP_CHECK(hipMemsetAsync(w.ffp, 0, ffp_bytes(), w.stream));
    w.ffp_seq = 1;
  }
  const unsigned flag0 = w.ffp_seq * (unsigned)PANEL_BASE;
  ffp_kernel<T, PANEL_BASE><<<dim3(grid, bt.nbat), FFP_THREADS, 0, w.stream>>>(
      A, lda, n, c0, c1, k0, k1, rpw, (unsigned long long*)w.ffp, flag0, ipiv, bt.info ? bt.info : w.info, w.ffp_sync,
      w.ffp_spin, bt.sA, bt.sP, ffp_matrix_words());
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// Base panel without row interchanges (static pivoting): the rows were put in the
// order of an earlier factorisation's pivots beforehand, so row j is the pivot of
// column j.  Columns [c0, c1) (at most W), rows [c0, n); matrix blockIdx.y of a batch.
// - Every workgroup factors the w x w diagonal block itself, row by row (threads
//   t < w hold its rows in registers; row j, final, goes to LDS before the rows below
//   it use it).  The operations and their order are the same in every workgroup, so
//   all hold the same factored block.
// - Its rows below the block then solve l U11 = a, one row per thread in registers,
//   with the same row operations.
// - The last workgroup to arrive (a counter, returned to zero by that workgroup) writes
//   the factored diagonal block: every other one has read the block by then, so none
//   reads a block that was already overwritten.  It also writes the identity
//   interchanges and the first zero pivot.
// - growth: the largest |multiplier| of the matrix, as the bits of a non-negative
//   double (an unsigned maximum orders them; inf and NaN above every finite value).
template <typename T, int W>
__global__ __launch_bounds__(256) void nopiv_panel_kernel(T* __restrict__ A, int64_t lda, int n, int c0, int c1,
                                                          int* __restrict__ ipiv, int* __restrict__ info,
                                                          unsigned* __restrict__ cnt,
                                                          unsigned long long* __restrict__ growth, int64_t sA,
                                                          int sP) {
  A += (int64_t)blockIdx.y * sA;
  ipiv += (int64_t)blockIdx.y * sP;
  info += blockIdx.y;
  cnt += blockIdx.y;
  growth += blockIdx.y;
  __shared__ T U[W][W + 1];  // the factored diagonal block, row j once it is final
  __shared__ unsigned long long s_g[256 / 64];
  __shared__ int s_last;
  const int w = c1 - c0, tid = threadIdx.x;
  // this thread's row below the block (loaded first: its latency overlaps the block)
  const int64_t i = (int64_t)c0 + (int64_t)blockIdx.x * 256 + tid;
  const bool mine = i >= c0 + w && i < n;
  T row[W];
#pragma unroll
  for (int c = 0; c < W; ++c) row[c] = mine && c < w ? A[i + (int64_t)(c0 + c) * lda] : T(0);
  // the diagonal block, row t per thread t < w
  T drow[W];
#pragma unroll
  for (int c = 0; c < W; ++c) drow[c] = tid < w && c < w ? A[(int64_t)(c0 + tid) + (int64_t)(c0 + c) * lda] : T(0);
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (j < w) {
      if (tid == j) {
#pragma unroll
        for (int c = 0; c < W; ++c) U[j][c] = drow[c];
      }
      __syncthreads();
      if (tid > j && tid < w) {
        const T l = drow[j] / U[j][j];
        drow[j] = l;
#pragma unroll
        for (int c = j + 1; c < W; ++c) drow[c] -= l * U[j][c];
      }
    }
  }
  __syncthreads();
  // every thread has read its part of the diagonal block (its values are in LDS, behind
  // the barrier): count this workgroup in
  if (tid == 0)
    s_last = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
  unsigned long long g = 0;
  auto take = [&g](T l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(fabs((double)l));
    g = b > g ? b : g;
  };
  if (mine) {
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (j < w) {
        const T l = row[j] / U[j][j];
        row[j] = l;
        take(l);
#pragma unroll
        for (int c = j + 1; c < W; ++c) row[c] -= l * U[j][c];
      }
    }
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < w) A[i + (int64_t)(c0 + c) * lda] = row[c];
  }
  __syncthreads();  // s_last
  if (s_last) {
    if (tid < w) {
#pragma unroll
      for (int c = 0; c < W; ++c) {
        if (c < w) A[(int64_t)(c0 + tid) + (int64_t)(c0 + c) * lda] = U[tid][c];
        if (c < tid && c < w) take(U[tid][c]);
      }
      ipiv[c0 + tid] = c0 + tid;
    }
    if (tid == 0) {
      __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int j = 0; j < w; ++j)
        if (U[j][j] == T(0)) {
          if (*info == 0) *info = c0 + j + 1;
          break;
        }
    }
  }
  // the workgroup's largest multiplier, then the matrix's
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned long long h = __shfl_xor(g, o);
    g = h > g ? h : g;
  }
  if ((tid & 63) == 0) s_g[tid >> 6] = g;
  __syncthreads();
  if (tid == 0) {
    for (int q = 1; q < 256 / 64; ++q) g = s_g[q] > g ? s_g[q] : g;
    __hip_atomic_fetch_max(growth, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename T>
int nopiv_launch(T* A, int64_t lda, int n, int c0, int c1, int* ipiv, LuWork& w, const LuBatch& bt) {
  if (c1 - c0 > PANEL_BASE || c1 <= c0 || c1 > n) return arg_error("nopiv_launch: panel shape");
  if (bt.nbat < 1 || bt.nbat > LU_MAX_BATCH || !bt.growth || !w.bcnt) return arg_error("nopiv_launch: batch");
  const dim3 grid((unsigned)cdiv(n - c0, 256), (unsigned)bt.nbat);
  nopiv_panel_kernel<T, PANEL_BASE><<<grid, 256, 0, w.stream>>>(A, lda, n, c0, c1, ipiv, bt.info ? bt.info : w.info,
                                                                 w.bcnt, bt.growth, bt.sA, bt.sP);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// Base panel: columns [c0, c1) (c1 - c0 <= PANEL_BASE), rows [c0, n).
template <typename T>
int panel_base(T* A, int64_t lda, int n, int c0, int c1, int* ipiv, LuWork& w) {
  const int nb0 = (int)cdiv(n - c0, 256);
  const int nbs = (int)cdiv(n - c0, FF_ROWS);
  if (nb0 > w.red_cap || nbs > w.red_cap) return KAD_ERR_ARG;
  int* map[2] = {w.map, w.map + n};
  double* pv[2] = {w.red_val, w.red_val + w.red_cap};
  int* pi[2] = {w.red_idx, w.red_idx + w.red_cap};
  ff_first_kernel<T><<<nb0, 256, 0, w.stream>>>(A, lda, n, c0, map[0], pv[0], pi[0]);
  int np = nb0, cur = 0;
  for (int j = c0; j < c1; ++j) {
    const int rows = n - j - 1;
    const int nbk = rows > 0 ? (int)cdiv(rows, FF_ROWS) : 1;
    ff_step_kernel<T><<<nbk, 256, 0, w.stream>>>(A, lda, n, c0, j, c1, map[cur], map[cur ^ 1], pv[cur], pi[cur], np,
                                                 pv[cur ^ 1], pi[cur ^ 1], ipiv, w.info);
    np = nbk;
    cur ^= 1;
  }
  KAD_HIP_CHECK(hipGetLastError());
  // map[cur] now holds the final logical -> physical order of rows [c0, n)
  const size_t need = (size_t)(n - c0) * (c1 - c0) * sizeof(T);
  if (need > w.tmp2_bytes) return KAD_ERR_ARG;
  dim3 g((unsigned)cdiv(n - c0, 256), c1 - c0);
  ff_gather_kernel<T><<<g, 256, 0, w.stream>>>(A, lda, n, c0, c1, map[cur], (T*)w.tmp2);
  ff_scatter_kernel<T><<<g, 256, 0, w.stream>>>(A, lda, n, c0, c1, (const T*)w.tmp2);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// Recursive panel factorisation of columns [c0, c1), rows [c0, n), inside the panel
// [pk0, pk1).  fused: the base panels are the persistent kernel, which applies its
// interchanges across [pk0, pk1) itself, so the recursion swaps no rows.
// A batch (bt.nbat > 1) runs the persistent kernel only.
template <typename T>
int panel_rec(T* A, int64_t lda, int n, int c0, int c1, int* ipiv, LuWork& w, int pk0, int pk1, bool fused,
              const LuBatch& bt = LuBatch()) {
  const int wdt = c1 - c0;
  if (bt.nbat > 1 && !fused && !bt.nopiv) return arg_error("panel_rec: a batch needs the persistent panel");
  if (wdt <= PANEL_BASE) {
    if (bt.nopiv) return nopiv_launch<T>(A, lda, n, c0, c1, ipiv, w, bt);
    return fused ? ffp_launch<T>(A, lda, n, c0, c1, pk0, pk1, ipiv, w, bt) : panel_base<T>(A, lda, n, c0, c1, ipiv, w);
  }
  int h = wdt / 2;
  h = (h + 15) / 16 * 16;
  const int cm = c0 + h;
  int rc = panel_rec<T>(A, lda, n, c0, cm, ipiv, w, pk0, pk1, fused, bt);
  if (rc) return rc;
  if (!fused && (rc = laswp<T>(A, lda, cm, c1, ipiv, c0, cm, w))) return rc;
  rc = trsm_left_unit_lower<T>(A + c0 + (int64_t)c0 * lda, lda, h, A + c0 + (int64_t)cm * lda, lda, c1 - cm, w,
                               prec_of<T>(), nullptr, bt.nbat, bt.sA, bt.sA);
  if (rc) return rc;
  rc = gemm_launch_batched(prec_of<T>(), GOP_SUB, n - cm, c1 - cm, h, A + cm + (int64_t)c0 * lda, lda, bt.sA,
                           A + c0 + (int64_t)cm * lda, lda, bt.sA, A + cm + (int64_t)cm * lda, lda, bt.sA, bt.nbat,
                           w.stream);
  if (rc) return rc;
  rc = panel_rec<T>(A, lda, n, cm, c1, ipiv, w, pk0, pk1, fused, bt);
  if (rc) return rc;
  return fused ? KAD_OK : laswp<T>(A, lda, c0, cm, ipiv, cm, c1, w);
}

// dst[i + c ldt] = A[map[i] + c lda] for i < nrows, c < ncols (row gather of a column block).
template <typename T>
__global__ void gather_rows_kernel(const T* __restrict__ A, int64_t lda, const int* __restrict__ map, int nrows,
                                   T* __restrict__ dst, int64_t ldt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i < nrows) dst[i + (int64_t)c * ldt] = A[map[i] + (int64_t)c * lda];
}

// The interchanges of later panels applied to the L columns of earlier panels, once,
// after the factorisation (those columns are never read again while factoring, so
// deferring them is exact).  For panel p the interchanges of steps k1_p .. n-1,
// composed on the host, are one row gather per column of rows [k1_p, n).
template <typename T>
int apply_left_swaps(T* A, int64_t lda, int n, const int* ipiv, int nb, LuWork& w) {
  const int np = (int)cdiv(n, nb);
  if (np < 2) return KAD_OK;
  std::vector<int> hp(n);
  KAD_HIP_CHECK(hipMemcpyAsync(hp.data(), ipiv, n * sizeof(int), hipMemcpyDeviceToHost, w.stream));
  KAD_HIP_CHECK(hipStreamSynchronize(w.stream));
  // the interchanges index host memory below: a factorisation that did not complete
  // (an abandoned panel barrier) leaves entries unwritten
  for (int i = 0; i < n; ++i)
    if (hp[i] < i || hp[i] >= n) return KAD_ERR_HIP;
  size_t tot = 0;
  for (int p = 0; p + 1 < np; ++p) tot += (size_t)(n - (p + 1) * nb);
  if (tot > w.swap_cap) {
    if (w.swap_map) KAD_HIP_CHECK(hipFree(w.swap_map));
    w.swap_map = nullptr;
    w.swap_cap = 0;
    KAD_HIP_CHECK(hipMalloc((void**)&w.swap_map, tot * sizeof(int)));
    w.swap_cap = tot;
  }
  std::vector<int> maps(tot);
  std::vector<size_t> moff(np, 0);
  size_t o = 0;
  for (int p = 0; p + 1 < np; ++p) {
    const int k1 = (p + 1) * nb;
    moff[p] = o;
    int* idx = maps.data() + o;
    for (int i = 0; i < n - k1; ++i) idx[i] = k1 + i;
    for (int i = k1; i < n; ++i) {
      const int q = hp[i];
      if (q != i) std::swap(idx[i - k1], idx[q - k1]);
    }
    o += (size_t)(n - k1);
  }
  KAD_HIP_CHECK(hipMemcpyAsync(w.swap_map, maps.data(), tot * sizeof(int), hipMemcpyHostToDevice, w.stream));
  for (int p = 0; p + 1 < np; ++p) {
    const int k0 = p * nb, k1 = k0 + nb;
    const int nrows = n - k1;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(k1 - k0, w.tmp_bytes / sizeof(T) / nrows));
    for (int c = k0; c < k1; c += chunk) {
      const int nc = std::min(chunk, k1 - c);
      dim3 g((unsigned)cdiv(nrows, 256), nc);
      gather_rows_kernel<T><<<g, 256, 0, w.stream>>>(A + (int64_t)c * lda, lda, w.swap_map + moff[p], nrows,
                                                     (T*)w.tmp, nrows);
      copy2d_kernel<T><<<g, 256, 0, w.stream>>>((const T*)w.tmp, nrows, A + k1 + (int64_t)c * lda, lda, nrows, nc);
      KAD_HIP_CHECK(hipGetLastError());
    }
  }
  // the host map buffer must outlive its asynchronous upload
  KAD_HIP_CHECK(hipStreamSynchronize(w.stream));
  return KAD_OK;
}

// Full factorisation.  gemm_prec selects the trailing-update MFMA precision
// (F64 for double storage; F32 or BF16 for float storage).
//
// Look-ahead of depth one on two streams of the same priority.  For step p (panel
// [k0, k1), next panel [k1, k2), the one after [k2, k3)):
//   panel stream: interchanges, U12 and update of the next panel's columns
//     [k1, k2), then the next panel's factorisation and its L21 row-tile mask;
//   main stream, part A: interchanges, U12 and update of [k2, k3) only, then the
//     signal the panel stream waits for at step p+1;
//   main stream, part B: interchanges, U12 and update of [k3, n) -- the bulk of the
//     step, which runs beside the next panel instead of ahead of it.
// Events (ping-pong by step parity so a new record never overtakes a pending
// wait): ev_panel "panel p and its mask are done", ev_next "columns of the panel
// after next are up to date".  The two streams work on disjoint columns, each with
// its own triangular-inverse / GEMM scratch.  The L21 row-tile masks rotate over
// three buffers: part B of step p still reads panel p's mask while the panel stream
// writes panel p+2's.
//
// A batch (bt.nbat > 1) factors bt.nbat matrices of order n together: every launch
// covers all of them, so the panel chain is that of one matrix.  The matrices are
// taken as dense (no structural-zero masks or row extents).
template <typename T>
int getrf(T* A, int64_t lda, int n, int* ipiv, int nb, int gemm_prec, LuWork& w, const LuBatch& bt) {
  // both streams' scratch must exist before anything is launched
  if (!w.stream2 || !w.tinv || !w.tinv_m || !w.tmp || !w.tmp_m || w.tmp_m_bytes < w.tmp_bytes || !w.tmp2 ||
      !w.mask_r || !w.mask_r2 || !w.mask_r3 || !w.ext)
    return KAD_ERR_ALLOC;
  LuWork wp = w;  // panel stream: tinv, tmp, tmp2 and the panel reductions
  wp.stream = w.stream2;
  // the persistent panel's launch sequence advances in wp; it outlives this call on
  // every return path (reused flags would let stale exchange slots look current)
  struct SeqBack {
    LuWork& to;
    const LuWork& from;
    ~SeqBack() { to.ffp_seq = from.ffp_seq; }
  } seq_back{w, wp};
  LuWork wm = w;  // main stream: its own triangular inverse and GEMM scratch
  wm.tinv = w.tinv_m;
  wm.tmp = w.tmp_m;
  wm.tmp_bytes = w.tmp_m_bytes;
  int* mask_rows[3] = {w.mask_r, w.mask_r2, w.mask_r3};
  const int stride = (int)cdiv(n, 128) + 1;
  // Row extent of every 128-column tile of the input (one pass, one readback).
  // A panel's rows at or below R are zero in its columns and stay zero: the
  // interchanges and the update of an earlier panel q touch rows above R_q only,
  // so R_p = max(initial extent of p's columns, R_q for q < p) bounds them.
  // a batch, or static pivoting: no masks, no row extents, the persistent panel or the
  // static-pivot panel (no interchanges at all)
  const bool bat = bt.nbat > 1 || bt.nopiv;
  const bool piv = !bt.nopiv;
  if (bt.nbat < 1 || bt.nbat > LU_MAX_BATCH) return arg_error("getrf: batch size");
  if (bt.nopiv) {
    if (!bt.growth) return arg_error("getrf: static pivoting needs growth words");
    KAD_HIP_CHECK(hipMemsetAsync(bt.growth, 0, bt.nbat * sizeof(unsigned long long), w.stream));
  }
  const int ntc = (int)cdiv(n, 128), nrc = (int)cdiv(n, 2048);
  std::vector<int> ext(ntc, bat ? n - 1 : -1);
  if (!bat) {
    col_extent_kernel<T><<<dim3(ntc, nrc), 256, 0, w.stream>>>(A, lda, n, w.ext);
    KAD_HIP_CHECK(hipGetLastError());
    std::vector<int> hext((size_t)ntc * nrc);
    KAD_HIP_CHECK(hipMemcpyAsync(hext.data(), w.ext, hext.size() * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    KAD_HIP_CHECK(hipStreamSynchronize(w.stream));
    for (int r = 0; r < nrc; ++r)
      for (int t = 0; t < ntc; ++t) ext[t] = std::max(ext[t], hext[(size_t)r * ntc + t]);
  }
  int r_prev = 0;
  // factor panel [k0, k1) on the panel stream (rows [k0, R)), then its L21 mask
  auto panel = [&](int k0, int k1, int* mask) -> int {
    int R = std::max(k1, r_prev);
    for (int t = k0 / 128; t <= (k1 - 1) / 128; ++t) R = std::max(R, ext[t] + 1);
    R = std::min(R, n);
    r_prev = R;
    int fg, frpw;
    const bool fused = bt.nopiv || (wp.ffp && !wp.ffp_off && ffp_shape<T>(R - k0, wp.ncu, fg, frpw));
    int rc = panel_rec<T>(A, lda, R, k0, k1, ipiv, wp, k0, k1, fused, bt);
    if (rc) return rc;
    if (k1 < n && mask) {
      // the panel's rows at or below R are zero in its columns
      row_tile_mask_kernel<T><<<(unsigned)cdiv(n - k1, 128), 256, 0, wp.stream>>>(A + k1 + (int64_t)k0 * lda, lda,
                                                                                 n - k1, k1 - k0, mask, R - k1);
      KAD_HIP_CHECK(hipGetLastError());
    }
    return KAD_OK;
  };
  // the caller's work on the main stream precedes the first panel
  KAD_HIP_CHECK(hipEventRecord(w.ev_start, w.stream));
  KAD_HIP_CHECK(hipStreamWaitEvent(wp.stream, w.ev_start, 0));
  int rc = panel(0, std::min(nb, n), bat ? nullptr : mask_rows[0]);
  if (rc) return rc;
  KAD_HIP_CHECK(hipEventRecord(w.ev_panel[0], wp.stream));
  for (int p = 0, k0 = 0; k0 < n; ++p, k0 += nb) {
    const int k1 = std::min(k0 + nb, n);
    const int k2 = std::min(k1 + nb, n);
    const int k3 = std::min(k2 + nb, n);
    const int par = p & 1;
    const int* mrows = bat ? nullptr : mask_rows[p % 3];
    const T* L21 = A + k1 + (int64_t)k0 * lda;
    const T* L11 = A + k0 + (int64_t)k0 * lda;
    // ---- main stream, part A: columns [k2, k3)
    KAD_HIP_CHECK(hipStreamWaitEvent(w.stream, w.ev_panel[par], 0));
    if (k1 >= n) break;
    if (k2 < n && bat) {
      if (piv && (rc = laswp<T>(A, lda, k2, k3, ipiv, k0, k1, wm, bt))) return rc;
      rc = trsm_left_unit_lower<T>(L11, lda, k1 - k0, A + k0 + (int64_t)k2 * lda, lda, k3 - k2, wm, gemm_prec,
                                   nullptr, bt.nbat, bt.sA, bt.sA);
      if (rc) return rc;
      rc = gemm_launch_batched(gemm_prec, GOP_SUB, n - k1, k3 - k2, k1 - k0, L21, lda, bt.sA,
                               A + k0 + (int64_t)k2 * lda, lda, bt.sA, A + k1 + (int64_t)k2 * lda, lda, bt.sA,
                               bt.nbat, w.stream);
      if (rc) return rc;
    } else if (k2 < n) {
      rc = laswp<T>(A, lda, k2, k3, ipiv, k0, k1, wm);
      if (rc) return rc;
      // structural zeros of A12 (column tiles; U12 = L11^{-1} A12 keeps them)
      col_tile_mask_kernel<T><<<(unsigned)cdiv(k3 - k2, 128), 256, 0, w.stream>>>(A + k0 + (int64_t)k2 * lda, lda,
                                                                                 k1 - k0, k3 - k2, w.mask_c);
      KAD_HIP_CHECK(hipGetLastError());
      rc = trsm_left_unit_lower<T>(L11, lda, k1 - k0, A + k0 + (int64_t)k2 * lda, lda, k3 - k2, wm, gemm_prec,
                                   w.mask_c);
      if (rc) return rc;
      rc = gemm_launch_masked(gemm_prec, GOP_SUB, n - k1, k3 - k2, k1 - k0, L21, lda, A + k0 + (int64_t)k2 * lda, lda,
                              A + k1 + (int64_t)k2 * lda, lda, mrows, w.mask_c, w.stream);
      if (rc) return rc;
    }
    KAD_HIP_CHECK(hipEventRecord(w.ev_next[par], w.stream));
    // ---- main stream, part B: columns [k3, n), beside the next panel
    if (k3 < n && bat) {
      if (piv && (rc = laswp<T>(A, lda, k3, n, ipiv, k0, k1, wm, bt))) return rc;
      rc = trsm_left_unit_lower<T>(L11, lda, k1 - k0, A + k0 + (int64_t)k3 * lda, lda, n - k3, wm, gemm_prec,
                                   nullptr, bt.nbat, bt.sA, bt.sA);
      if (rc) return rc;
      rc = gemm_launch_batched(gemm_prec, GOP_SUB, n - k1, n - k3, k1 - k0, L21, lda, bt.sA,
                               A + k0 + (int64_t)k3 * lda, lda, bt.sA, A + k1 + (int64_t)k3 * lda, lda, bt.sA,
                               bt.nbat, w.stream);
      if (rc) return rc;
    } else if (k3 < n) {
      rc = laswp<T>(A, lda, k3, n, ipiv, k0, k1, wm);
      if (rc) return rc;
      col_tile_mask_kernel<T><<<(unsigned)cdiv(n - k3, 128), 256, 0, w.stream>>>(A + k0 + (int64_t)k3 * lda, lda,
                                                                                k1 - k0, n - k3, w.mask_c2);
      KAD_HIP_CHECK(hipGetLastError());
      rc = trsm_left_unit_lower<T>(L11, lda, k1 - k0, A + k0 + (int64_t)k3 * lda, lda, n - k3, wm, gemm_prec,
                                   w.mask_c2);
      if (rc) return rc;
      // active tile pairs of the update, compacted on the grid (all XCDs busy)
      tile_lists_kernel<<<1, 64, 0, w.stream>>>(mrows, (int)cdiv(n - k1, 128), w.mask_c2, (int)cdiv(n - k3, 128),
                                                w.tile_list, stride);
      KAD_HIP_CHECK(hipGetLastError());
      TileSel sel;
      sel.rlist = w.tile_list;
      sel.clist = w.tile_list + stride;
      sel.cnt = w.tile_list + 2 * stride;
      rc = gemm_launch_sel(gemm_prec, GOP_SUB, n - k1, n - k3, k1 - k0, L21, lda, A + k0 + (int64_t)k3 * lda, lda,
                           A + k1 + (int64_t)k3 * lda, lda, sel, w.stream);
      if (rc) return rc;
    }
    // ---- panel stream: the next panel's columns [k1, k2), then the next panel
    if (p > 0) KAD_HIP_CHECK(hipStreamWaitEvent(wp.stream, w.ev_next[par ^ 1], 0));
    if (piv && (rc = laswp<T>(A, lda, k1, k2, ipiv, k0, k1, wp, bt))) return rc;
    rc = trsm_left_unit_lower<T>(L11, lda, k1 - k0, A + k0 + (int64_t)k1 * lda, lda, k2 - k1, wp, gemm_prec, nullptr,
                                 bt.nbat, bt.sA, bt.sA);
    if (rc) return rc;
    rc = bat ? gemm_launch_batched(gemm_prec, GOP_SUB, n - k1, k2 - k1, k1 - k0, L21, lda, bt.sA,
                                   A + k0 + (int64_t)k1 * lda, lda, bt.sA, A + k1 + (int64_t)k1 * lda, lda, bt.sA,
                                   bt.nbat, wp.stream)
             : gemm_launch_masked(gemm_prec, GOP_SUB, n - k1, k2 - k1, k1 - k0, L21, lda, A + k0 + (int64_t)k1 * lda,
                                  lda, A + k1 + (int64_t)k1 * lda, lda, mrows, nullptr, wp.stream);
    if (rc) return rc;
    rc = panel(k1, k2, bat ? nullptr : mask_rows[(p + 1) % 3]);
    if (rc) return rc;
    KAD_HIP_CHECK(hipEventRecord(w.ev_panel[par ^ 1], wp.stream));
  }
  for (int z = 0; z < bt.nbat && piv; ++z)
    if ((rc = apply_left_swaps<T>(A + z * bt.sA, lda, n, ipiv + (int64_t)z * bt.sP, nb, w))) return rc;
  return KAD_OK;
}

template int getrf<double>(double*, int64_t, int, int*, int, int, LuWork&, const LuBatch&);
template int getrf<float>(float*, int64_t, int, int*, int, int, LuWork&, const LuBatch&);

// perm from LAPACK-style ipiv (host side, O(n)).
// Returns false (perm unspecified) if some ipiv[i] is outside [i, n): interchanges read
// back from the device are checked before they index host memory.
bool ipiv_to_perm(const int* ipiv, int n, int* perm) {
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int i = 0; i < n; ++i) {
    const int p = ipiv[i];
    if (p < i || p >= n) return false;
    if (p != i) std::swap(perm[i], perm[p]);
  }
  return true;
}

// ---------------------------------------------------------------------------
// fp64 GEMV  y = b - A x  (or y = A x when b == nullptr), A column-major n x n.
// Two passes: partial sums per (row block, column chunk), then a fixed-order
// reduction, so results are bitwise reproducible.
// ---------------------------------------------------------------------------
constexpr int GEMV_CHUNK = 1024;

__global__ __launch_bounds__(256) void gemv_partial_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                           const double* __restrict__ x, double* __restrict__ part,
                                                           int trans) {
  const int chunk = blockIdx.y;
  const int c0 = chunk * GEMV_CHUNK, c1 = min(n, c0 + GEMV_CHUNK);
  if (!trans) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    const double* a = A + i;
    for (int c = c0; c < c1; ++c) s += a[(int64_t)c * lda] * x[c];
    part[(int64_t)chunk * n + i] = s;
  }
}

__global__ void gemv_reduce_kernel(const double* __restrict__ part, int nchunk, int n, const double* __restrict__ b,
                                   double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * n + i];
  y[i] = b ? b[i] - s : s;
}

// y = A^T x  (one workgroup per column)
__global__ __launch_bounds__(256) void gemv_t_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                     const double* __restrict__ x, const double* __restrict__ b,
                                                     double* __restrict__ y) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  const double* a = A + (int64_t)c * lda;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += a[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[c] = b ? b[c] - red[0] : red[0];
}

int gemv64(const double* A, int64_t lda, int n, const double* x, const double* b, double* y, int trans, double* part,
           hipStream_t s) {
  if (!trans) {
    const int nchunk = (int)cdiv(n, GEMV_CHUNK);
    dim3 g((unsigned)cdiv(n, 256), nchunk);
    gemv_partial_kernel<<<g, 256, 0, s>>>(A, lda, n, x, part, 0);
    gemv_reduce_kernel<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(part, nchunk, n, b, y);
  } else {
    gemv_t_kernel<<<n, 256, 0, s>>>(A, lda, n, x, b, y);
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// ---------------------------------------------------------------------------
// Small vector kernels.
// ---------------------------------------------------------------------------
__global__ void axpy_kernel(int n, double a, const double* x, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
__global__ void cast_d2f_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t nel) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nel) b[i] = (float)a[i];
}
// Column-major n x n copy with leading dimensions, fp64 -> fp32.
__global__ void cast_mat_d2f_kernel(const double* __restrict__ a, int64_t lda, float* __restrict__ b, int64_t ldb,
                                    int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i < n) b[i + (int64_t)c * ldb] = (float)a[i + (int64_t)c * lda];
}
// Equilibration by powers of two (LAPACK dgeequb-style, exact in fp64):
// R_i = 2^-round(log2 max_j |a_ij|), then C_j = 2^-round(log2 max_i |R_i a_ij|).
__device__ inline double pow2_recip(double m) {
  if (!(m > 0.0) || !isfinite(m)) return 1.0;
  int e;
  frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
  return ldexp(1.0, 1 - e);
}
__global__ __launch_bounds__(256) void row_absmax_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                         double* __restrict__ R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = 0.0;
  for (int c = 0; c < n; ++c) m = fmax(m, fabs(A[i + (int64_t)c * lda]));
  R[i] = pow2_recip(m);
}
__global__ __launch_bounds__(256) void col_scale_kernel(double* __restrict__ A, int64_t lda, int n,
                                                        const double* __restrict__ R, double* __restrict__ C) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double* a = A + (int64_t)c * lda;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(R[i] * a[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const double cj = pow2_recip(red[0]);
  if (threadIdx.x == 0) C[c] = cj;
  for (int i = threadIdx.x; i < n; i += 256) a[i] = R[i] * a[i] * cj;
}
// Column scales of diag(R) A without modifying A.
__global__ __launch_bounds__(256) void col_scale_only_kernel(const double* __restrict__ A, int64_t lda, int n,
                                                             const double* __restrict__ R, double* __restrict__ C) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  const double* a = A + (int64_t)c * lda;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(R[i] * a[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) C[c] = pow2_recip(red[0]);
}
// B = (float)(R_i A_ij C_j)
__global__ void cast_scaled_kernel(const double* __restrict__ a, int64_t lda, float* __restrict__ b, int64_t ldb,
                                   int n, const double* __restrict__ R, const double* __restrict__ C) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i < n) b[i + (int64_t)c * ldb] = (float)(R[i] * a[i + (int64_t)c * lda] * C[c]);
}
__global__ void vec_scale_kernel(int n, const double* __restrict__ s, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= s[i];
}

// Block reduction for max |x| and for dot products (fixed order, two passes).
__global__ __launch_bounds__(256) void absmax_partial_kernel(const double* x, int n, double* part) {
  __shared__ double red[256];
  double m = 0.0;
  // NaN propagates (fmax would drop it): a NaN residual must never read as converged
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double a = fabs(x[i]);
    m = (a > m || a != a) ? a : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = red[threadIdx.x + s];
      if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// h[i] = <V_i, w> for i < k  (V: k vectors of length n, contiguous), one workgroup per i
__global__ __launch_bounds__(256) void multidot_kernel(const double* __restrict__ V, int n, int k,
                                                       const double* __restrict__ w, double* __restrict__ h) {
  __shared__ double red[256];
  const double* v = V + (int64_t)blockIdx.x * n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i] * w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) h[blockIdx.x] = red[0];
}
// w -= sum_i h[i] V_i   (or, with sub = 0, w = sum_i h[i] V_i)
__global__ void multiaxpy_kernel(const double* __restrict__ V, int n, int k, const double* __restrict__ h,
                                 double* __restrict__ w, int sub) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += h[i] * V[(int64_t)i * n + e];
  w[e] = sub ? w[e] - s : s;
}
__global__ __launch_bounds__(256) void dot_partial_kernel(const double* x, const double* y, int n, double* part) {
  __shared__ double red[256];
  double m = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m += x[i] * y[i];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace kad
