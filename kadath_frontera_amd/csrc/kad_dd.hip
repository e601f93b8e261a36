// Domain-decomposed direct solve of the Newton system, single MI355X.
//
// The reference factors the whole Jacobian with one pdgesv_ (src_par/do_newton.cpp:
// 317-320), whose panel chain runs over all n columns.  In a multi-domain spectral
// space the unknowns and the equations split by domain, and only the matching rows
// of an interface couple two domains (take_into_account, src/System_of_eqs/
// solver.cpp:181,203).  With each interface's value-matching rows counted in the
// inner domain and its derivative-matching rows in the outer one, every domain's
// diagonal block D_b is square: the domain's own operator with a Dirichlet-type outer
// and a Neumann-type inner interface condition.  Then
//   J = D + P C,   D = diag(D_b),   P = unit columns of the r coupling rows,
//   C = the entries of those rows outside their own block,
// and (Woodbury)  J^{-1} v = z - W (I + C W)^{-1} C z,  z = D^{-1} v,  W = D^{-1} P.
//
// Set-up (kad_dd_factor), all on the device:
//  * the coupling rows: a scan of J for entries outside each row's own block;
//  * the diagonal blocks, padded to one order nd (a multiple of 256; identity on the
//    padding), factored together by the batched LU: one panel chain of nd columns
//    instead of n, every launch covering all blocks;
//  * W: the unit right-hand sides placed directly in pivoted row order, then the
//    batched unit-lower and upper triangular solves (MFMA GEMM updates);
//  * C as a dense r x (blocks x nd) copy (zero in each row's own block), the
//    capacitance matrix I + C W by one GEMM per block, and its LU (order r).
// Solve (kad_dd_solve): flexible GMRES on the fp64 J, preconditioned by the Woodbury
// solve.  That solve is direct up to rounding, so one or two inner steps reach the
// fp64 floor; the products with J keep the residual exact whatever the blocks'
// conditioning.
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include "../../include/kadath_amd.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <vector>

namespace kad {

int gemm_launch(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                int64_t ldc, hipStream_t s);
int gemm_launch_batched(int prec, int op, int m, int n, int k, const void* A, int64_t lda, int64_t sa, const void* B,
                        int64_t ldb, int64_t sb, void* C, int64_t ldc, int64_t sc, int nbat, hipStream_t s);
int gemm_launch_sel(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                    void* C, int64_t ldc, const TileSel& sel, hipStream_t s);
int lu_solve_permuted_batched(const double* LU, int64_t lda, int n, int nbat, int64_t sA, double* B, int64_t ldb,
                              int64_t sB, int nrhs, LuWork& w, const int* zcols, const int* zstart, int zng);
int ir_loop(Ctx* c, int n, const double* A64, int64_t lda, const std::function<int(double*)>& precond,
            const double* b, double* x, double tol, int max_outer, int mode, int trans, int* iters, double* resid,
            bool map_ready = false);
size_t trsv_table_doubles(int n);
size_t trsv_flag_ints(int n);
int trsv_prepare_batched(const double* LU, int64_t lda, int n, int nbat, int64_t sA, double* W, hipStream_t s);
int getrs_permuted_batched(const double* LU, int64_t lda, int n, int nbat, int64_t sA, const double* W, double* x,
                           double* y, hipStream_t s);
template <typename T>
int trsv_prepare(const T* LU, int64_t lda, int n, int trans, double* W, int* flag, TrsvExtents& ext, hipStream_t s);
template <typename T>
int getrs_fast(const T* LU, int64_t lda, int n, const int* perm, const double* W, const TrsvExtents& ext, double* b,
               int trans, double* work, hipStream_t s);

constexpr int DD_COLS = 512;  // columns per workgroup of the coupling-row scan
static_assert(LU_MAX_BATCH <= 32, "the coupling-row scan keeps one bit per block");
constexpr int DD_CHUNK = 1024;  // columns per partial sum of the C products
constexpr int DD_KSPLIT = 8;    // k-chunks of each block's C_b W_b (when 8 divides the order)

static_assert(GM_ROWS == 256 && GM_COLS == 2 * DD_COLS, "a scan workgroup covers half a GEMV tile");

// Bit z of flag[i] is set when row i of J has an entry in block z's columns, z other
// than the row's own block (columns [by DD_COLS, +DD_COLS) of this workgroup; at most
// 32 blocks).  flag[i] != 0: row i is a coupling row.  gmap (may be NULL, zeroed
// before): the GEMV tile flags of GemvMap for the same J, gmap[chunk nrb + rb] = 1 when
// its 256-row, 1024-column tile has a nonzero (two workgroups per tile).
__global__ __launch_bounds__(256) void dd_flag_kernel(const double* __restrict__ J, int64_t ld, int n,
                                                      const int* __restrict__ rblk, const int* __restrict__ cblk,
                                                      int* __restrict__ flag, int* __restrict__ gmap) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * DD_COLS, c1 = min(n, c0 + DD_COLS);
  int nz = 0;
  if (i < n) {  // no early return: the whole workgroup reaches the barrier below
    const int b = rblk[i];
    unsigned hit = 0;
    for (int c = c0; c < c1; ++c) {
      const double v = J[i + (int64_t)c * ld];
      const int cb = cblk[c];
      nz |= v != 0.0;
      hit |= (cb != b && v != 0.0) ? 1u << cb : 0u;
    }
    if (hit) atomicOr(reinterpret_cast<unsigned*>(flag) + i, hit);
  }
  if (gmap) {  // the same for every thread of the launch
    nz = __syncthreads_or(nz);
    if (threadIdx.x == 0 && nz) atomicOr(gmap + (int64_t)(blockIdx.y / 2) * gridDim.x + blockIdx.x, 1);
  }
}

// D_z(i, j) = J(rowmap_z[li], colmap_z[j]), li = rperm_z[i] (rows in an earlier
// factorisation's pivot order) or i (rperm NULL); identity where either is padding.
__global__ __launch_bounds__(256) void dd_blocks_kernel(const double* __restrict__ J, int64_t ld,
                                                        const int* __restrict__ rowmap, const int* __restrict__ colmap,
                                                        const int* __restrict__ rperm, int nd, double* __restrict__ D) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y, z = blockIdx.z;
  if (i >= nd) return;
  const int li = rperm ? rperm[(int64_t)z * nd + i] : i;
  const int row = rowmap[(int64_t)z * nd + li], col = colmap[(int64_t)z * nd + j];
  D[(int64_t)z * nd * nd + i + (int64_t)j * nd] = (row >= 0 && col >= 0) ? J[row + (int64_t)col * ld]
                                                                         : (li == j ? 1.0 : 0.0);
}

// B(i, :) = A(perm[i], :), r x r column-major (ld r).
__global__ __launch_bounds__(256) void dd_rows_kernel(const double* __restrict__ A, const int* __restrict__ perm,
                                                      int r, double* __restrict__ B) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i < r) B[i + (int64_t)j * r] = A[perm[i] + (int64_t)j * r];
}

// W_z(k, j) = 1 at each unit right-hand side's pivoted position (W zeroed before).
__global__ void dd_units_kernel(const int64_t* __restrict__ pos, int r, double* __restrict__ W) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < r) W[pos[q]] = 1.0;
}

// Jc(q, z nd + j) = J(crow[q], colmap_z[j]) outside row q's own block, else 0.
__global__ __launch_bounds__(256) void dd_coupling_kernel(const double* __restrict__ J, int64_t ld,
                                                          const int* __restrict__ crow, const int* __restrict__ cz,
                                                          const int* __restrict__ colmap, int nd, int r,
                                                          double* __restrict__ Jc) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y, z = blockIdx.z;
  if (q >= r) return;
  const int col = colmap[(int64_t)z * nd + j];
  Jc[q + ((int64_t)z * nd + j) * r] = (col >= 0 && cz[q] != z) ? J[crow[q] + (int64_t)col * ld] : 0.0;
}

// C(i, j) = the sum of P_z(i, j) over z = 0 .. nz-1 in that order (P_z at P + z sp,
// r x kb, ld r): the k-chunks of a split product.
__global__ __launch_bounds__(256) void dd_ksum_kernel(const double* __restrict__ P, int r, int nz, int64_t sp,
                                                      double* __restrict__ C) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= r) return;
  const int64_t e = i + (int64_t)j * r;
  double acc = P[e];
  for (int z = 1; z < nz; ++z) acc += P[e + z * sp];
  C[e] = acc;
}

__global__ void dd_add_identity_kernel(double* __restrict__ A, int64_t ld, int r) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < r) A[q + (int64_t)q * ld] += 1.0;
}

// vb_z[i] = v[rowmap_z[perm_z[i]]]: the right-hand side of block z in pivoted order.
__global__ void dd_gather_kernel(const double* __restrict__ v, const int* __restrict__ rowmap,
                                 const int* __restrict__ perm, int nd, double* __restrict__ vb) {
  const int i = blockIdx.x * 256 + threadIdx.x, z = blockIdx.y;
  if (i >= nd) return;
  const int row = rowmap[(int64_t)z * nd + perm[(int64_t)z * nd + i]];
  vb[(int64_t)z * nd + i] = row >= 0 ? v[row] : 0.0;
}

// x[colmap_z[j]] = vb_z[j].
__global__ void dd_scatter_kernel(const double* __restrict__ vb, const int* __restrict__ colmap, int nd,
                                  double* __restrict__ x) {
  const int j = blockIdx.x * 256 + threadIdx.x, z = blockIdx.y;
  if (j >= nd) return;
  const int col = colmap[(int64_t)z * nd + j];
  if (col >= 0) x[col] = vb[(int64_t)z * nd + j];
}

// part[chunk r + q] = sum over the chunk's columns c of Jc(q, c) vb[c].
__global__ __launch_bounds__(256) void dd_cprod_kernel(const double* __restrict__ Jc, int r, int64_t ncol,
                                                       const double* __restrict__ vb, double* __restrict__ part) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= r) return;
  const int64_t c0 = (int64_t)blockIdx.y * DD_CHUNK, c1 = min(ncol, c0 + DD_CHUNK);
  double s0 = 0.0, s1 = 0.0;
  int64_t c = c0;
  for (; c + 2 <= c1; c += 2) {
    s0 += Jc[q + c * r] * vb[c];
    s1 += Jc[q + (c + 1) * r] * vb[c + 1];
  }
  if (c < c1) s0 += Jc[q + c * r] * vb[c];
  part[(int64_t)blockIdx.y * r + q] = s0 + s1;
}

// t[q] = sum of the partial sums, in a fixed order.
__global__ void dd_creduce_kernel(const double* __restrict__ part, int nchunk, int r, double* __restrict__ t) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= r) return;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += part[(int64_t)k * r + q];
  t[q] = s;
}

// yb[z kmax + j] = t[q] for coupling row q = q0_z + j of block z.
__global__ void dd_spread_kernel(const double* __restrict__ t, int r, const int* __restrict__ cz,
                                 const int* __restrict__ q0, int kmax, double* __restrict__ yb) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= r) return;
  const int z = cz[q];
  yb[(int64_t)z * kmax + (q - q0[z])] = t[q];
}

}  // namespace kad

using kad::Ctx;

struct kad_dd_s {
  Ctx* c = nullptr;
  // the coupling-row scan (memory-bound) runs here beside the block LU (compute-bound)
  hipStream_t scan_stream = nullptr;
  // GEMV tile flags of the last factored J (GemvMap layout), made by the coupling scan;
  // gmap_J: that J while the flags are unused (dd_solve_after_factor takes them once)
  int* gmap = nullptr;
  const double* gmap_J = nullptr;
  int64_t gmap_ld = 0;
  int n = 0, nblk = 0, nd = 0;
  std::vector<int> bsize;
  std::vector<int> h_rowmap, h_colmap;  // nblk x nd, -1 on padding
  std::vector<int> h_rblk;              // block of each row of J
  std::vector<int> h_rloc;              // local index of each row of J in its block
  int* rowmap = nullptr;
  int* colmap = nullptr;
  int* rblk = nullptr;
  int* cblk = nullptr;
  int* flag = nullptr;
  double* D = nullptr;
  int* ipiv = nullptr;
  int* perm = nullptr;
  // coupling (sized at the last factorisation)
  int r = 0, kmax = 0;
  std::vector<int> kb, q0;
  size_t cap_r = 0, cap_w = 0;  // capacities: coupling rows, W entries
  int* crow = nullptr;          // r rows of J, ordered by (block, local row)
  int* cz = nullptr;            // block of each coupling row
  int* dq0 = nullptr;           // first coupling index of each block, then kb (2 nblk entries)
  int* zst = nullptr;           // per block and 64-column group of W: first pivoted row of its units
  int* cmk = nullptr;           // per block and 128 coupling rows: any of them couples to the block
  double* kpart = nullptr;      // DD_KSPLIT partial products C_b W_b, r x kmax each
  int64_t* upos = nullptr;      // W offset of each unit right-hand side
  double* W = nullptr;
  double* Jc = nullptr;
  double* cap = nullptr;
  int* cap_ipiv = nullptr;
  int* cap_perm = nullptr;
  double* cap_tab = nullptr;
  int* cap_flag = nullptr;
  kad::TrsvExtents cap_ext;
  double* vb = nullptr;         // nblk x nd
  double* vy = nullptr;         // nblk x nd (work of the block sweeps)
  double* trs = nullptr;        // the block factors' sweep tables, nblk x trsv_table_doubles(nd)
  double* t = nullptr;          // r, then 3 r of capacitance-solve work
  double* yb = nullptr;         // nblk x kmax
  double* part = nullptr;       // partial sums of C vb
  double* cap2 = nullptr;       // r x r: the capacitance rows in the kept pivot order
  size_t cap2_r = 0;
  bool factored = false;
  double times[5] = {0, 0, 0, 0, 0};
  // the pivot orders of the last factorisation with partial pivoting (perm, cap_perm),
  // kept for factorisations with static pivoting while the coupling rows stay the same
  bool piv_valid = false;
  std::vector<int> h_perm;      // nblk x nd, that factorisation's block orders
  std::vector<int> h_crow;      // its coupling rows
  bool last_static = false;     // the last factorisation used the kept orders
  double last_growth = 0.0;     // its largest |multiplier| (0: partial pivoting)

  int dmalloc(void** p, size_t bytes) {
    KAD_HIP_CHECK(hipMalloc(p, bytes < 16 ? 16 : bytes));
    return KAD_OK;
  }
  void release_coupling() {
    void* bufs[] = {crow, cz, dq0, zst, cmk, kpart, upos, W, Jc, cap, cap_ipiv, cap_perm, cap_tab, cap_flag, t, yb, part, cap2};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    crow = cz = dq0 = zst = cmk = nullptr;
    upos = nullptr;
    W = Jc = cap = cap_tab = t = yb = part = cap2 = kpart = nullptr;
    cap_ipiv = cap_perm = cap_flag = nullptr;
    cap_r = cap_w = cap2_r = 0;
    piv_valid = false;  // cap_perm is gone
  }
  ~kad_dd_s() {
    release_coupling();
    void* bufs[] = {rowmap, colmap, rblk, cblk, flag, gmap, D, ipiv, perm, vb, vy, trs};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    if (scan_stream) (void)hipStreamDestroy(scan_stream);
  }
};

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Coupling buffers for r rows (kmax per block) of a system with nblk blocks of order nd.
int ensure_coupling(kad_dd* dd, int r, int kmax) {
  const size_t w = (size_t)dd->nblk * dd->nd * kmax;
  if ((size_t)r <= dd->cap_r && w <= dd->cap_w) return KAD_OK;
  dd->release_coupling();
  const size_t R = std::max(r, 1), ncol = (size_t)dd->nblk * dd->nd;
  const size_t nchunk = kad::cdiv((int64_t)ncol, kad::DD_CHUNK);
  int rc;
  if ((rc = dd->dmalloc((void**)&dd->crow, R * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->cz, R * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->dq0, 2 * (size_t)dd->nblk * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->zst, (size_t)dd->nblk * kad::cdiv(kmax, 64) * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->cmk, (size_t)dd->nblk * kad::cdiv((int64_t)R, 128) * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->kpart, (size_t)kad::DD_KSPLIT * R * std::max(kmax, 1) * sizeof(double))))
    return rc;
  if ((rc = dd->dmalloc((void**)&dd->upos, R * sizeof(int64_t)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->W, std::max<size_t>(w, 1) * sizeof(double)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->Jc, R * ncol * sizeof(double)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->cap, R * R * sizeof(double)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->cap_ipiv, R * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->cap_perm, R * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->cap_tab, kad::trsv_table_doubles((int)R) * sizeof(double)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->cap_flag, kad::trsv_flag_ints((int)R) * sizeof(int)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->t, 4 * R * sizeof(double)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->yb, (size_t)dd->nblk * std::max(kmax, 1) * sizeof(double)))) return rc;
  if ((rc = dd->dmalloc((void**)&dd->part, nchunk * R * sizeof(double)))) return rc;
  dd->cap_r = R;
  dd->cap_w = std::max<size_t>(w, 1);
  return KAD_OK;
}

// v <- J^{-1} v through the factored blocks and the capacitance matrix (v: rows of J
// in, unknowns out).
int dd_apply(kad_dd* dd, double* v) {
  Ctx* c = dd->c;
  hipStream_t s = c->stream;
  const int nd = dd->nd, nb = dd->nblk;
  const dim3 gv((unsigned)kad::cdiv(nd, 256), (unsigned)nb);
  kad::dd_gather_kernel<<<gv, 256, 0, s>>>(v, dd->rowmap, dd->perm, nd, dd->vb);
  KAD_HIP_CHECK(hipGetLastError());
  int rc = kad::getrs_permuted_batched(dd->D, nd, nd, nb, (int64_t)nd * nd, dd->trs, dd->vb, dd->vy, s);
  if (rc) return rc;
  if (dd->r > 0) {
    const int r = dd->r;
    const int64_t ncol = (int64_t)nb * nd;
    const unsigned nchunk = (unsigned)kad::cdiv(ncol, kad::DD_CHUNK);
    const unsigned gr = (unsigned)kad::cdiv(r, 256);
    kad::dd_cprod_kernel<<<dim3(gr, nchunk), 256, 0, s>>>(dd->Jc, r, ncol, dd->vb, dd->part);
    kad::dd_creduce_kernel<<<gr, 256, 0, s>>>(dd->part, (int)nchunk, r, dd->t);
    KAD_HIP_CHECK(hipGetLastError());
    if ((rc = kad::getrs_fast<double>(dd->cap, r, r, dd->cap_perm, dd->cap_tab, dd->cap_ext, dd->t, 0, dd->t + r, s)))
      return rc;
    KAD_HIP_CHECK(hipMemsetAsync(dd->yb, 0, (size_t)nb * dd->kmax * sizeof(double), s));
    kad::dd_spread_kernel<<<gr, 256, 0, s>>>(dd->t, r, dd->cz, dd->dq0, dd->kmax, dd->yb);
    KAD_HIP_CHECK(hipGetLastError());
    // z_b -= W_b y_b
    if ((rc = kad::gemm_launch_batched(kad::F64, 0, nd, 1, dd->kmax, dd->W, nd, (int64_t)nd * dd->kmax, dd->yb,
                                       dd->kmax, dd->kmax, dd->vb, nd, nd, nb, s)))
      return rc;
  }
  kad::dd_scatter_kernel<<<gv, 256, 0, s>>>(dd->vb, dd->colmap, nd, v);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

}  // namespace

extern "C" {

int kad_dd_create(kad_ctx* ctx, int n, int nblk, const int* blk_size, const int* rows, const int* cols, kad_dd** out) {
  if (!ctx || !out || n <= 0 || nblk < 1 || nblk > kad::LU_MAX_BATCH || !blk_size || !rows || !cols)
    return KAD_ERR_ARG;
  *out = nullptr;
  int64_t tot = 0, mx = 0;
  for (int b = 0; b < nblk; ++b) {
    if (blk_size[b] < 1) return KAD_ERR_ARG;
    tot += blk_size[b];
    mx = std::max<int64_t>(mx, blk_size[b]);
  }
  if (tot != n) return KAD_ERR_ARG;
  // every row and every column of J in exactly one block
  std::vector<int> seen_r(n, 0), seen_c(n, 0);
  for (int i = 0; i < n; ++i) {
    if (rows[i] < 0 || rows[i] >= n || cols[i] < 0 || cols[i] >= n) return KAD_ERR_ARG;
    if (seen_r[rows[i]]++ || seen_c[cols[i]]++) return KAD_ERR_ARG;
  }
  auto dd = new kad_dd_s();
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  dd->c = c;
  dd->n = n;
  dd->nblk = nblk;
  // padded block order: a multiple of the batched factorisation's panel width
  // (KAD_LU_NB when it is a multiple of 256, else 256)
  int q = 256;
  const char* e = std::getenv("KAD_LU_NB");
  if (e && std::atoi(e) > 0 && std::atoi(e) % 256 == 0) q = std::atoi(e);
  dd->nd = (int)(kad::cdiv(mx, q) * q);
  dd->bsize.assign(blk_size, blk_size + nblk);
  const int nd = dd->nd;
  dd->h_rowmap.assign((size_t)nblk * nd, -1);
  dd->h_colmap.assign((size_t)nblk * nd, -1);
  dd->h_rblk.assign(n, -1);
  dd->h_rloc.assign(n, -1);
  std::vector<int> hcb(n, -1);
  int64_t o = 0;
  for (int b = 0; b < nblk; ++b)
    for (int i = 0; i < blk_size[b]; ++i, ++o) {
      dd->h_rowmap[(size_t)b * nd + i] = rows[o];
      dd->h_colmap[(size_t)b * nd + i] = cols[o];
      dd->h_rblk[rows[o]] = b;
      dd->h_rloc[rows[o]] = i;
      hcb[cols[o]] = b;
    }
  int rc = KAD_OK;
  auto fail = [&](int e) {
    delete dd;
    return e;
  };
  if (hipSetDevice(c->device) != hipSuccess) return fail(KAD_ERR_HIP);
  if ((rc = dd->dmalloc((void**)&dd->rowmap, dd->h_rowmap.size() * sizeof(int)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->colmap, dd->h_colmap.size() * sizeof(int)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->rblk, (size_t)n * sizeof(int)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->cblk, (size_t)n * sizeof(int)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->flag, (size_t)n * sizeof(int)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->gmap,
                        (size_t)kad::cdiv(n, kad::GM_ROWS) * kad::cdiv(n, kad::GM_COLS) * sizeof(int))))
    return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->D, (size_t)nblk * nd * nd * sizeof(double)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->ipiv, (size_t)nblk * nd * sizeof(int)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->perm, (size_t)nblk * nd * sizeof(int)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->vb, (size_t)nblk * nd * sizeof(double)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->vy, (size_t)nblk * nd * sizeof(double)))) return fail(rc);
  if ((rc = dd->dmalloc((void**)&dd->trs, (size_t)nblk * kad::trsv_table_doubles(nd) * sizeof(double))))
    return fail(rc);
  if (hipMemcpy(dd->rowmap, dd->h_rowmap.data(), dd->h_rowmap.size() * sizeof(int), hipMemcpyHostToDevice) ||
      hipMemcpy(dd->colmap, dd->h_colmap.data(), dd->h_colmap.size() * sizeof(int), hipMemcpyHostToDevice) ||
      hipMemcpy(dd->rblk, dd->h_rblk.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice) ||
      hipMemcpy(dd->cblk, hcb.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice))
    return fail(KAD_ERR_HIP);
  *out = dd;
  return KAD_OK;
}

int kad_dd_destroy(kad_dd* dd) {
  delete dd;
  return KAD_OK;
}

}  // extern "C"

namespace {

// The factorisation of kad_dd_factor (stat false: partial pivoting) and of
// kad_dd_factor_static (stat true: the rows of every block and of the capacitance
// matrix in the pivot orders of the last factorisation with partial pivoting, factored
// without interchanges).  Static returns > 0 when the orders do not apply:
// kStaticNone no orders kept, kStaticBlocks a zero pivot or a multiplier above
// max_growth in a block, kStaticCoupling other coupling rows, kStaticCapacitance the
// same in the capacitance matrix.
constexpr int kStaticNone = 1, kStaticBlocks = 2, kStaticCoupling = 3, kStaticCapacitance = 4;
// stat: every row order kept (multipliers within max_growth, else kStatic*); otherwise
// partial pivoting in the blocks and, with cap_keep >= 1, the capacitance matrix in its
// kept order when the coupling rows are the same and its multipliers stay within
// cap_keep (I + C D^{-1} P does not depend on the blocks' pivot orders), else with
// partial pivoting too.
int dd_factor_impl(kad_dd* dd, const double* J, int64_t ldj, bool stat, double max_growth, double cap_keep = 0.0) {
  if (!dd || !J || ldj < dd->n) return KAD_ERR_ARG;
  Ctx* c = dd->c;
  hipStream_t s = c->stream;
  const int n = dd->n, nd = dd->nd, nb = dd->nblk;
  dd->factored = false;
  dd->gmap_J = nullptr;
  dd->last_static = stat;
  dd->last_growth = 0.0;
  if (stat && !dd->piv_valid) return kStaticNone;
  const bool had_orders = dd->piv_valid;
  if (!stat) dd->piv_valid = false;
  int rc;
  KAD_HIP_CHECK(hipStreamSynchronize(s));
  double t0 = now_s();
  kad::dd_blocks_kernel<<<dim3((unsigned)kad::cdiv(nd, 256), (unsigned)nd, (unsigned)nb), 256, 0, s>>>(
      J, ldj, dd->rowmap, dd->colmap, stat ? dd->perm : nullptr, nd, dd->D);
  KAD_HIP_CHECK(hipGetLastError());
  KAD_HIP_CHECK(hipStreamSynchronize(s));
  double t1 = now_s();
  // coupling rows, scanned on the second stream while the blocks are factored (a
  // memory-bound scan beside compute-bound updates; it only reads J, which is complete:
  // s was synchronised before the gather)
  if (!dd->scan_stream) KAD_HIP_CHECK(hipStreamCreateWithFlags(&dd->scan_stream, hipStreamNonBlocking));
  hipStream_t s2 = dd->scan_stream;
  struct ScanJoin {  // every return waits for the scan: the caller may reuse J at once
    hipStream_t st;
    ~ScanJoin() { (void)hipStreamSynchronize(st); }
  } scan_join{s2};
  const size_t gcells = (size_t)kad::cdiv(n, kad::GM_ROWS) * kad::cdiv(n, kad::GM_COLS);
  KAD_HIP_CHECK(hipMemsetAsync(dd->flag, 0, (size_t)n * sizeof(int), s2));
  KAD_HIP_CHECK(hipMemsetAsync(dd->gmap, 0, gcells * sizeof(int), s2));
  kad::dd_flag_kernel<<<dim3((unsigned)kad::cdiv(n, 256), (unsigned)kad::cdiv(n, kad::DD_COLS)), 256, 0, s2>>>(
      J, ldj, n, dd->rblk, dd->cblk, dd->flag, dd->gmap);
  KAD_HIP_CHECK(hipGetLastError());
  // the diagonal blocks, factored together
  std::vector<int> info(nb, 0);
  if (!stat) {
    rc = kad_lu_factor_batched(reinterpret_cast<kad_ctx*>(c), KAD_PREC_F64, nd, nb, dd->D, nd, (int64_t)nd * nd,
                               dd->ipiv, dd->perm, info.data());
    if (rc) return rc;
    for (int b = 0; b < nb; ++b)
      if (info[b] != 0) return info[b];  // a singular domain block (1-based column inside it)
  } else {
    std::vector<double> gr(nb, 0.0);
    rc = kad_lu_factor_batched_static(reinterpret_cast<kad_ctx*>(c), KAD_PREC_F64, nd, nb, dd->D, nd,
                                      (int64_t)nd * nd, dd->ipiv, info.data(), gr.data());
    if (rc) return rc;
    for (int b = 0; b < nb; ++b) {
      dd->last_growth = std::max(dd->last_growth, std::isnan(gr[b]) ? INFINITY : gr[b]);
      if (info[b] != 0 || !(gr[b] <= max_growth)) return kStaticBlocks;
    }
  }
  if ((rc = kad::trsv_prepare_batched(dd->D, nd, nd, nb, (int64_t)nd * nd, dd->trs, s))) return rc;
  KAD_HIP_CHECK(hipStreamSynchronize(s));
  double t2 = now_s();
  std::vector<int> hflag(n);
  KAD_HIP_CHECK(hipMemcpyAsync(hflag.data(), dd->flag, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s2));
  KAD_HIP_CHECK(hipStreamSynchronize(s2));
  // coupling rows by (block, local row); their unit right-hand sides in pivoted order
  std::vector<std::vector<int>> byblk(nb);
  for (int i = 0; i < n; ++i)
    if (hflag[i]) byblk[dd->h_rblk[i]].push_back(i);
  std::vector<int> kb(nb, 0), q0(nb, 0);
  int r = 0, kmax = 0;
  for (int b = 0; b < nb; ++b) {
    std::sort(byblk[b].begin(), byblk[b].end(),
              [&](int x, int y) { return dd->h_rloc[x] < dd->h_rloc[y]; });
    q0[b] = r;
    kb[b] = (int)byblk[b].size();
    r += kb[b];
    kmax = std::max(kmax, kb[b]);
  }
  std::vector<int> hcrow(r), hcz(r);
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < kb[b]; ++j) {
      hcrow[q0[b] + j] = byblk[b][j];
      hcz[q0[b] + j] = b;
    }
  // the kept capacitance order belongs to the same coupling rows
  if (stat && hcrow != dd->h_crow) return kStaticCoupling;
  dd->kb = kb;
  dd->q0 = q0;
  dd->r = r;
  dd->kmax = kmax;
  if (r > 0) {
    if ((rc = ensure_coupling(dd, r, kmax))) return rc;
    if (!stat) {
      dd->h_perm.resize((size_t)nb * nd);
      KAD_HIP_CHECK(hipMemcpyAsync(dd->h_perm.data(), dd->perm, dd->h_perm.size() * sizeof(int),
                                   hipMemcpyDeviceToHost, s));
      KAD_HIP_CHECK(hipStreamSynchronize(s));
    }
    std::vector<int64_t> hpos(r);
    std::vector<int> inv(nd);
    const int ng = (int)kad::cdiv(kmax, 64);
    std::vector<int> hst((size_t)nb * ng, nd);  // groups without units: nothing to solve
    for (int b = 0; b < nb; ++b) {
      for (int k = 0; k < nd; ++k) inv[dd->h_perm[(size_t)b * nd + k]] = k;
      for (int j = 0; j < kb[b]; ++j) {
        const int q = q0[b] + j, row = byblk[b][j];
        // e_{local row} in pivoted order sits at inv[local row]; column j of W_b
        const int p = inv[dd->h_rloc[row]];
        hpos[q] = (int64_t)b * nd * kmax + (int64_t)j * nd + p;
        int& st = hst[(size_t)b * ng + j / 64];
        st = std::min(st, p);
      }
    }
    KAD_HIP_CHECK(hipMemcpyAsync(dd->zst, hst.data(), hst.size() * sizeof(int), hipMemcpyHostToDevice, s));
    // the coupling rows with entries in block b's columns (those of the neighbouring
    // domains' interfaces), per 128-row tile of C_b: the other tiles of C_b W_b are zero
    const int nt = (int)kad::cdiv(r, 128);
    std::vector<int> hmk((size_t)nb * nt, 0);
    for (int q = 0; q < r; ++q) {
      const unsigned bits = (unsigned)hflag[hcrow[q]];
      for (int b = 0; b < nb; ++b)
        if ((bits >> b) & 1u) hmk[(size_t)b * nt + q / 128] = 1;
    }
    KAD_HIP_CHECK(hipMemcpyAsync(dd->cmk, hmk.data(), hmk.size() * sizeof(int), hipMemcpyHostToDevice, s));
    KAD_HIP_CHECK(hipMemcpyAsync(dd->crow, hcrow.data(), r * sizeof(int), hipMemcpyHostToDevice, s));
    KAD_HIP_CHECK(hipMemcpyAsync(dd->cz, hcz.data(), r * sizeof(int), hipMemcpyHostToDevice, s));
    KAD_HIP_CHECK(hipMemcpyAsync(dd->dq0, dd->q0.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
    KAD_HIP_CHECK(hipMemcpyAsync(dd->dq0 + nb, dd->kb.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
    KAD_HIP_CHECK(hipMemcpyAsync(dd->upos, hpos.data(), r * sizeof(int64_t), hipMemcpyHostToDevice, s));
    // W = D^{-1} P; block b has kb[b] right-hand sides, the solve skips the zero
    // columns past them and, in the unit-lower solve, the zero rows above each group's
    // first unit (forward substitution keeps them zero)
    KAD_HIP_CHECK(hipMemsetAsync(dd->W, 0, (size_t)nb * nd * kmax * sizeof(double), s));
    kad::dd_units_kernel<<<(unsigned)kad::cdiv(r, 256), 256, 0, s>>>(dd->upos, r, dd->W);
    KAD_HIP_CHECK(hipGetLastError());
    if ((rc = kad::lu_solve_permuted_batched(dd->D, nd, nd, nb, (int64_t)nd * nd, dd->W, nd, (int64_t)nd * kmax, kmax,
                                             c->lu, dd->dq0 + nb, dd->zst, ng)))
      return rc;
    KAD_HIP_CHECK(hipStreamSynchronize(s));
    double t3 = now_s();
    // C, then I + C W
    kad::dd_coupling_kernel<<<dim3((unsigned)kad::cdiv(r, 256), (unsigned)nd, (unsigned)nb), 256, 0, s>>>(
        J, ldj, dd->crow, dd->cz, dd->colmap, nd, r, dd->Jc);
    KAD_HIP_CHECK(hipGetLastError());
    // C_b W_b in DD_KSPLIT k-chunks, a batch of partial products summed in chunk order:
    // the rows that couple to one block are too few tiles to fill the chip over the
    // full order
    const int ks = nd % kad::DD_KSPLIT == 0 ? kad::DD_KSPLIT : 1, kc = nd / ks;
    for (int b = 0; b < nb; ++b) {
      const int kbb = dd->kb[b];
      if (kbb == 0) continue;
      kad::TileSel sel;
      sel.rmask = dd->cmk + (int64_t)b * nt;
      sel.nbat = ks;
      sel.sa = (int64_t)kc * r;
      sel.sb = kc;
      sel.sc = (int64_t)r * kbb;
      if ((rc = kad::gemm_launch_sel(kad::F64, 1, r, kbb, kc, dd->Jc + (int64_t)b * nd * r, r,
                                     dd->W + (int64_t)b * nd * kmax, nd, dd->kpart, r, sel, s)))
        return rc;
      kad::dd_ksum_kernel<<<dim3((unsigned)kad::cdiv(r, 256), (unsigned)kbb), 256, 0, s>>>(
          dd->kpart, r, ks, (int64_t)r * kbb, dd->cap + (int64_t)dd->q0[b] * r);
      KAD_HIP_CHECK(hipGetLastError());
    }
    kad::dd_add_identity_kernel<<<(unsigned)kad::cdiv(r, 256), 256, 0, s>>>(dd->cap, r, r);
    KAD_HIP_CHECK(hipGetLastError());
    KAD_HIP_CHECK(hipStreamSynchronize(s));
    double t4 = now_s();
    // the rows in the kept order (cap_perm), into the second buffer, factored there
    // without interchanges; cap keeps I + C W.  1: accepted (the buffers trade places),
    // 0: a zero pivot or a multiplier above `bound`, < 0 an error
    auto factor_kept = [&](double bound) -> int {
      if (dd->cap2_r < dd->cap_r) {  // the capacity of cap: the two buffers trade places
        if (dd->cap2) (void)hipFree(dd->cap2);
        dd->cap2 = nullptr;
        dd->cap2_r = 0;
        int e;
        if ((e = dd->dmalloc((void**)&dd->cap2, dd->cap_r * dd->cap_r * sizeof(double)))) return e;
        dd->cap2_r = dd->cap_r;
      }
      kad::dd_rows_kernel<<<dim3((unsigned)kad::cdiv(r, 256), (unsigned)r), 256, 0, s>>>(dd->cap, dd->cap_perm, r,
                                                                                         dd->cap2);
      KAD_HIP_CHECK(hipGetLastError());
      int cinfo = 0;
      double cg = 0.0;
      const int e = kad_lu_factor_batched_static(reinterpret_cast<kad_ctx*>(c), KAD_PREC_F64, r, 1, dd->cap2, r, 0,
                                                 dd->cap_ipiv, &cinfo, &cg);
      if (e) return e < 0 ? e : KAD_ERR_ARG;
      const double g = std::isnan(cg) ? INFINITY : cg;
      if (cinfo != 0 || !(g <= bound)) {
        if (stat) dd->last_growth = std::max(dd->last_growth, g);
        return 0;
      }
      dd->last_growth = std::max(dd->last_growth, g);
      std::swap(dd->cap, dd->cap2);  // the factors of the kept-order rows, for the solves
      return 1;
    };
    if (!stat) {
      int kept = 0;
      if (cap_keep >= 1.0 && had_orders && hcrow == dd->h_crow && (kept = factor_kept(cap_keep)) < 0) return kept;
      if (!kept) {
        rc = kad_lu_factor(reinterpret_cast<kad_ctx*>(c), KAD_PREC_F64, r, dd->cap, r, dd->cap_ipiv, dd->cap_perm,
                           0);
        if (rc < 0) return rc;
        if (rc > 0) return n + rc;  // singular capacitance matrix
      }
    } else {
      const int kept = factor_kept(max_growth);
      if (kept < 0) return kept;
      if (!kept) return kStaticCapacitance;
    }
    if ((rc = kad::trsv_prepare<double>(dd->cap, r, r, 0, dd->cap_tab, dd->cap_flag, dd->cap_ext, s))) return rc;
    KAD_HIP_CHECK(hipStreamSynchronize(s));
    double t5 = now_s();
    dd->times[2] = t3 - t2;
    dd->times[3] = t4 - t3;
    dd->times[4] = t5 - t4;
  } else {
    dd->times[2] = dd->times[3] = dd->times[4] = 0.0;
  }
  dd->times[0] = t1 - t0;
  dd->times[1] = t2 - t1;
  if (!stat) {
    dd->h_crow = hcrow;
    dd->piv_valid = true;
  }
  dd->gmap_J = J;
  dd->gmap_ld = ldj;
  dd->factored = true;
  return KAD_OK;
}

}  // namespace

extern "C" {

int kad_dd_factor(kad_dd* dd, const double* J, int64_t ldj) { return dd_factor_impl(dd, J, ldj, false, 0.0); }

int kad_dd_factor_static(kad_dd* dd, const double* J, int64_t ldj, double max_growth) {
  if (!(max_growth >= 1.0)) return KAD_ERR_ARG;
  return dd_factor_impl(dd, J, ldj, true, max_growth);
}

int kad_dd_factor_keep(kad_dd* dd, const double* J, int64_t ldj, double max_growth) {
  if (!(max_growth >= 1.0)) return KAD_ERR_ARG;
  return dd_factor_impl(dd, J, ldj, false, 0.0, max_growth);
}

int kad_dd_solve(kad_dd* dd, const double* J, int64_t ldj, const double* b, double* x, double tol, int max_outer,
                 int* iters, double* resid) {
  if (!dd || !J || !b || !x || ldj < dd->n || max_outer < 1) return KAD_ERR_ARG;
  if (!dd->factored) return KAD_ERR_ARG;
  auto pre = [dd](double* v) { return dd_apply(dd, v); };
  return kad::ir_loop(dd->c, dd->n, J, ldj, pre, b, x, tol, max_outer, 1, 0, iters, resid);
}

int kad_dd_apply(kad_dd* dd, double* v) {
  if (!dd || !v || !dd->factored) return KAD_ERR_ARG;
  const int rc = dd_apply(dd, v);
  if (rc) return rc;
  KAD_HIP_CHECK(hipStreamSynchronize(dd->c->stream));
  return KAD_OK;
}

int kad_dd_pivoting(const kad_dd* dd, int* static_orders, double* growth) {
  if (!dd) return KAD_ERR_ARG;
  if (static_orders) *static_orders = dd->last_static ? 1 : 0;
  if (growth) *growth = dd->last_growth;
  return KAD_OK;
}

int kad_dd_info(const kad_dd* dd, int* sizes, double* times) {
  if (!dd) return KAD_ERR_ARG;
  if (sizes) {
    sizes[0] = dd->nd;
    sizes[1] = dd->r;
    sizes[2] = dd->kmax;
    sizes[3] = dd->nblk;
  }
  if (times)
    for (int q = 0; q < 5; ++q) times[q] = dd->times[q];
  return KAD_OK;
}

}  // extern "C"

namespace kad {

// kad_dd_solve for a J whose entries have not changed since the factorisation of the
// same J that just succeeded: the refinement's GEMV takes the nonzero tiles that the
// factorisation's coupling-row scan recorded instead of reading J once more.  A caller
// that cannot promise that (J refilled in place between the two) calls kad_dd_solve.
int dd_solve_after_factor(kad_dd* dd, const double* J, int64_t ldj, const double* b, double* x, double tol,
                          int max_outer, int* iters, double* resid) {
  if (!dd || !J || !b || !x || ldj < dd->n || max_outer < 1) return KAD_ERR_ARG;
  if (!dd->factored) return KAD_ERR_ARG;
  bool ready = false;
  if (dd->gmap_J == J && dd->gmap_ld == ldj) {
    if (const int rc = dd->c->jmap.adopt(J, ldj, dd->n, dd->gmap, dd->c->stream)) return rc;
    ready = true;
  }
  dd->gmap_J = nullptr;  // the flags serve one solve
  auto pre = [dd](double* v) { return dd_apply(dd, v); };
  return ir_loop(dd->c, dd->n, J, ldj, pre, b, x, tol, max_outer, 1, 0, iters, resid, ready);
}

}  // namespace kad
