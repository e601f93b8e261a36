// Device-resident Newton step of the XCTS system: executes the traced programs
// (value program -> residual F; tangent programs -> Jacobian columns, batched
// over unit seeds), factors and solves, and updates the fields in place.
//
// Replaces System_of_eqs::do_newton (src_par/do_newton.cpp:65-430):
//   sec_member / check_equations   -> value program + |F|_inf
//   do_col_J over all columns      -> tangent programs, B columns per launch chain
//   pdgesv_                        -> kad LU (fp64) or low-precision LU + fp64 refinement
//   xx_to_vars + old - var         -> CSR update kernel on the coefficient arrays
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include "kad_spectral.hpp"
#include "host/xcts.hpp"
#include "host/parsed_system.hpp"
#include "../../include/kadath_amd.h"
#include <vector>
#include <chrono>
#include <algorithm>
#include <stdexcept>
#include <cstring>
#include <functional>

namespace kad {
template <typename T>
int getrf(T* A, int64_t lda, int n, int* ipiv, int nb, int gemm_prec, LuWork& w, const LuBatch& bt = LuBatch());
bool ipiv_to_perm(const int* ipiv, int n, int* perm);
// kad_dd_solve right after a factorisation of the same, unchanged J (kad_dd.hip)
int dd_solve_after_factor(kad_dd* dd, const double* J, int64_t ldj, const double* b, double* x, double tol,
                          int max_outer, int* iters, double* resid);
}  // namespace kad

using namespace kadhost;

namespace {

struct DevRow {
  int pos, gpos;
  double gw;
};

using kad::ew_op;
static_assert(EW_ADD == kad::EWOP_ADD && EW_SUB == kad::EWOP_SUB && EW_MUL == kad::EWOP_MUL &&
                  EW_DIV == kad::EWOP_DIV && EW_SCAL == kad::EWOP_SCAL && EW_ADDS == kad::EWOP_ADDS &&
                  EW_COPY == kad::EWOP_COPY && EW_SDIV == kad::EWOP_SDIV,
              "the tracer's element-wise kinds are the kernels'");

// The same over 16-byte-aligned rows of even length: every thread takes EW_U pairs of
// elements, their loads issued together (one element per thread left each wave
// waiting out one load's latency at a time).
constexpr int EW_U = 2;
__global__ __launch_bounds__(256) void ew2_kernel(int kind, double* __restrict__ dst, int64_t dstride,
                                                  const double* __restrict__ a, int64_t astride,
                                                  const double* __restrict__ bsrc, int64_t bstride, double s,
                                                  int size2, int nbatch) {
  for (int bb = blockIdx.y; bb < nbatch; bb += gridDim.y) {
    const double2* ar = reinterpret_cast<const double2*>(a + (int64_t)bb * astride);
    const double2* br = bsrc ? reinterpret_cast<const double2*>(bsrc + (int64_t)bb * bstride) : nullptr;
    double2* dr = reinterpret_cast<double2*>(dst + (int64_t)bb * dstride);
    const int e0 = (blockIdx.x * 256 * EW_U) + threadIdx.x;
    double2 x[EW_U], y[EW_U];
#pragma unroll
    for (int u = 0; u < EW_U; ++u) {
      const int e = e0 + u * 256;
      x[u] = e < size2 ? ar[e] : double2{0.0, 0.0};
      y[u] = (br && e < size2) ? br[e] : double2{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < EW_U; ++u) {
      const int e = e0 + u * 256;
      if (e < size2) dr[e] = double2{ew_op(kind, x[u].x, y[u].x, s), ew_op(kind, x[u].y, y[u].y, s)};
    }
  }
}

// An element-wise instruction followed by the ones it absorbed (kad::PostOp): the
// first instruction's result goes through the chain and only the last one is stored.
__global__ __launch_bounds__(256) void ewc_kernel(int kind, const double* __restrict__ a, int64_t astride,
                                                  const double* __restrict__ bsrc, int64_t bstride, double s,
                                                  kad::PostOp post, int size, int nbatch) {
  for (int bb = blockIdx.y; bb < nbatch; bb += gridDim.y) {
    const double* ar = a + (int64_t)bb * astride;
    const double* br = bsrc ? bsrc + (int64_t)bb * bstride : nullptr;
    double* dr = post.dst + (int64_t)bb * post.ds;
    const int e0 = blockIdx.x * 256 * EW_U + threadIdx.x;
    // every load of the EW_U elements first (operands of the head and of every step)
    double x[EW_U], y[EW_U], o[EW_U][kad::kMaxChain];
#pragma unroll
    for (int u = 0; u < EW_U; ++u) {
      const int e = e0 + u * 256;
      const bool in = e < size;
      x[u] = in ? ar[e] : 0.0;
      y[u] = (in && br) ? br[e] : 0.0;
      kad::post_load(post, bb, in ? e : 0, o[u]);
    }
#pragma unroll
    for (int u = 0; u < EW_U; ++u) {
      const int e = e0 + u * 256;
      if (e < size) dr[e] = kad::post_ops(post, ew_op(kind, x[u], y[u], s), o[u]);
    }
  }
}

// dst[b][e] = op(a[b or 0][e], b[b or 0][e], s)
// Element-wise instruction over a batch: blockIdx.y walks the batch rows (a stride of
// 0 broadcasts one array to every row), blockIdx.x / threadIdx.x the elements.
__global__ void ew_kernel(int kind, double* __restrict__ dst, int64_t dstride, const double* __restrict__ a,
                          int64_t astride, const double* __restrict__ bsrc, int64_t bstride, double s, int size,
                          int nbatch) {
  for (int bb = blockIdx.y; bb < nbatch; bb += gridDim.y) {
    const double* ar = a + (int64_t)bb * astride;
    const double* br = bsrc ? bsrc + (int64_t)bb * bstride : nullptr;
    double* dr = dst + (int64_t)bb * dstride;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < size; e += gridDim.x * blockDim.x) {
      const double x = ar[e];
      double r;
      switch (kind) {
        case EW_ADD: r = x + br[e]; break;
        case EW_SUB: r = x - br[e]; break;
        case EW_MUL: r = x * br[e]; break;
        case EW_DIV: r = x / br[e]; break;
        case EW_SCAL: r = x * s; break;
        case EW_SDIV: r = x / s; break;
        case EW_ADDS: r = x + s; break;
        default: r = x; break;
      }
      dr[e] = r;
    }
  }
}

// Inside rows: out[(row0 + r) + (col0 + b) ldo] (+)= sign (src[pos] - gw src[gpos])
template <typename T>
__global__ void export_in_kernel(const DevRow* __restrict__ rows, int nrows, const double* __restrict__ src,
                                 int64_t sstride, T* __restrict__ out, int row0, int64_t col0, int64_t ldo,
                                 double sign, int acc, int nbatch) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nbatch) return;
  const int64_t bb = t / nrows;
  const int r = (int)(t % nrows);
  const DevRow R = rows[r];
  const double* x = src + bb * sstride;
  double v = x[R.pos];
  if (R.gpos >= 0) v -= R.gw * x[R.gpos];
  T* o = out + (row0 + r) + (col0 + bb) * ldo;
  *o = acc ? (T)((double)*o + sign * v) : (T)(sign * v);
}

// Boundary rows: X(line) = sum_i bsign^i cf(i, line)  (val_boundary, domain_shell_systems.cpp:43-74)
template <typename T>
__global__ void export_bd_kernel(const DevRow* __restrict__ rows, int nrows, const double* __restrict__ src,
                                 int64_t sstride, int nr, int bsign, T* __restrict__ out, int row0,
                                 int64_t col0, int64_t ldo, double sign, int acc, int nbatch) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nbatch) return;
  const int64_t bb = t / nrows;
  const int r = (int)(t % nrows);
  const DevRow R = rows[r];
  const double* x = src + bb * sstride;
  double v = 0.0, g = 0.0, sg = 1.0;
  for (int i = 0; i < nr; ++i) {
    v += sg * x[i + (int64_t)nr * R.pos];
    if (R.gpos >= 0) g += sg * x[i + (int64_t)nr * R.gpos];
    sg *= bsign;
  }
  if (R.gpos >= 0) v -= R.gw * g;
  T* o = out + (row0 + r) + (col0 + bb) * ldo;
  *o = acc ? (T)((double)*o + sign * v) : (T)(sign * v);
}

// Transposed exports (the VJP of the row export): for every column b of the batch,
// out[col0 + b] += sign * sum_r w[row0 + r] row_r(b), where row_r is the inside
// (bd = 0) or boundary (bd = 1) row recipe above.  One workgroup per column, a
// fixed-order reduction, and instructions run in stream order: deterministic.
__global__ __launch_bounds__(256) void export_dot_kernel(const DevRow* __restrict__ rows, int nrows, int bd,
                                                         const double* __restrict__ src, int64_t sstride, int nr,
                                                         int bsign, const double* __restrict__ w, int row0,
                                                         double* __restrict__ out, int64_t col0, double sign) {
  __shared__ double red[256];
  const int bb = blockIdx.x;
  const double* x = src + (int64_t)bb * sstride;
  double acc = 0.0;
  for (int r = threadIdx.x; r < nrows; r += 256) {
    const DevRow R = rows[r];
    double v;
    if (!bd) {
      v = x[R.pos];
      if (R.gpos >= 0) v -= R.gw * x[R.gpos];
    } else {
      double g = 0.0, sg = 1.0;
      v = 0.0;
      for (int i = 0; i < nr; ++i) {
        v += sg * x[i + (int64_t)nr * R.pos];
        if (R.gpos >= 0) g += sg * x[i + (int64_t)nr * R.gpos];
        sg *= bsign;
      }
      if (R.gpos >= 0) v -= R.gw * g;
    }
    acc += w[row0 + r] * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col0 + bb] += sign * red[0];
}

// Unit seeds of unknowns u0 .. u0 + nbatch - 1 (affecte_tau_one_coef), zero elsewhere.
__global__ void seed_kernel(double* __restrict__ dst, int64_t dstride, int size, const int* __restrict__ spos,
                            const double* __restrict__ sval, const int* __restrict__ scnt, int u0, int nu,
                            int nbatch) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)size * nbatch) return;
  const int bb = (int)(t / size);
  const int e = (int)(t % size);
  const int u = u0 + bb;
  double v = 0.0;
  if (u < nu)
    for (int q = 0; q < scnt[u]; ++q)
      if (spos[3 * u + q] == e) v = sval[3 * u + q];
  dst[(int64_t)bb * dstride + e] = v;
}

// cf[q] += scale * sum_t w[t] X[off + col[t]]   (affecte_tau as a sparse map)
__global__ void csr_update_kernel(double* __restrict__ cf, const int* __restrict__ rowptr, const int* __restrict__ col,
                                  const double* __restrict__ w, const double* __restrict__ X, int off, int ncf,
                                  double scale) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncf) return;
  double s = 0.0;
  for (int t = rowptr[q]; t < rowptr[q + 1]; ++t) s += w[t] * X[off + col[t]];
  cf[q] += scale * s;
}

// max |x| over one workgroup.  A NaN propagates (fmax would drop it): a residual
// holding a NaN must read as NaN, so that do_newton never reports it as converged
// (the reference's max starts from data[0] and keeps a NaN there too).
__global__ void absmax_kernel(const double* x, int n, double* out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
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
  if (threadIdx.x == 0) *out = red[0];
}

// sum of squares (compute_f of the line search, do_newton.cpp:594-599) and signed maximum
// (Array max, array_math.cpp:308-314) of x[0..n), one workgroup; out[0] = sum, out[1] = max.
// A NaN reaches both results.
__global__ void sumsq_max_kernel(const double* x, int n, double* out) {
  __shared__ double rs[256], rm[256];
  double s = 0.0, m = n > 0 ? x[0] : 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double a = x[i];
    s += a * a;
    m = (a > m || a != a) ? a : m;
  }
  rs[threadIdx.x] = s;
  rm[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      rs[threadIdx.x] += rs[threadIdx.x + k];
      const double o = rm[threadIdx.x + k];
      if (o > rm[threadIdx.x] || o != o) rm[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = rs[0];
    out[1] = rm[0];
  }
}

// B = A^T for an n x n column-major matrix: 64 x 64 tiles through LDS (padded row), both
// sides coalesced.  Grid (cdiv(n, 64), cdiv(n, 64)), 256 threads.
__global__ __launch_bounds__(256) void transpose_kernel(const double* __restrict__ A, int64_t lda,
                                                        double* __restrict__ B, int64_t ldb, int n) {
  __shared__ double t[64][65];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int c = ty; c < 64; c += 4) {
    const int gr = r0 + tx, gc = c0 + c;
    t[c][tx] = (gr < n && gc < n) ? A[gr + (int64_t)gc * lda] : 0.0;
  }
  __syncthreads();
  // B(c0 + tx, r0 + r) = A(r0 + r, c0 + tx)
  for (int r = ty; r < 64; r += 4) {
    const int gr = c0 + tx, gc = r0 + r;
    if (gr < n && gc < n) B[gr + (int64_t)gc * ldb] = t[tx][r];
  }
}

// Surface integral of one quantity (Domain_compact::integ, domain_compact.cpp:89-121):
// cf holds the coefficients of r r sin(theta) so; value on the boundary x = bsign of
// phi coefficient 0, theta weights pi (COS_EVEN, j = 0) or 2/(2j+1) (SIN_ODD), times
// 2 pi and fac.  One workgroup.
__global__ void surface_integral_kernel(const double* __restrict__ cf, int nr, int nt, int bt0, int bsign,
                                        double fac, double* __restrict__ out) {
  __shared__ double red[64];
  const int j = threadIdx.x;
  double v = 0.0;
  if (j < nt && (bt0 == COS_EVEN || bt0 == SIN_ODD)) {
    double b = 0.0, sg = 1.0;
    for (int i = 0; i < nr; ++i) {
      b += sg * cf[i + (int64_t)nr * j];
      sg *= bsign;
    }
    if (bt0 == COS_EVEN) v = (j == 0) ? M_PI * b : 0.0;
    else v = 2.0 / (2.0 * j + 1.0) * b;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < 64; ++q) t += red[q];
    *out = t * 2.0 * M_PI * fac;
  }
}

// Per column b: fac * 2 pi * sum_j w_j sum_i cf_b(i, j, k = 0) (the integ weights of
// surface_integral_kernel at the outer boundary) into out[b + q ldo].
__global__ void export_int_kernel(const double* __restrict__ src, int64_t sstride, int nr, int nt, int bt0,
                                  double fac, double* __restrict__ out, int q, int64_t col0, int64_t ldo,
                                  int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  const double* x = src + (int64_t)b * sstride;
  double tot = 0.0;
  if (bt0 == COS_EVEN || bt0 == SIN_ODD) {
    for (int j = 0; j < nt; ++j) {
      const double w = bt0 == COS_EVEN ? (j == 0 ? M_PI : 0.0) : 2.0 / (2.0 * j + 1.0);
      if (w == 0.0) continue;
      double v = 0.0;
      for (int i = 0; i < nr; ++i) v += x[i + (int64_t)nr * j];
      tot += w * v;
    }
  }
  out[col0 + b + (int64_t)q * ldo] = tot * 2.0 * M_PI * fac;
}

// Integral-equation row (Domain::integ at x = bsign of the coefficients of r r sin(theta)
// so, theta weights of phi coefficient 0 (pi for COS_EVEN j = 0, 2/(2j+1) for SIN_ODD),
// times 2 pi and fac): value programs write out[row] = v + c0 (or add v); tangent
// programs write J(row, col0 + b) (column-major, ld ldo, always added: J is cleared);
// with w, the transposed export adds w[row] v to out[col0 + b].  src == nullptr: the
// constant alone.
template <typename T>
__global__ void export_gl_kernel(const double* __restrict__ src, int64_t sstride, int nr, int nt, int bt0,
                                 int bsign, double fac, double c0, int acc, int pt, T* __restrict__ out, int row,
                                 int64_t col0, int64_t ldo, const double* __restrict__ w, int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  double tot = 0.0;
  if (src && pt) {
    // point value of a constant field (a parameter): entry 0 of its array, times fac
    const double val = src[(int64_t)b * sstride] * fac;
    if (w) {
      ((double*)out)[col0 + b] += w[row] * val;
      return;
    }
    T* o = out + row + (col0 + b) * ldo;
    *o = acc ? (T)((double)*o + val + c0) : (T)(val + c0);
    return;
  }
  if (src && (bt0 == COS_EVEN || bt0 == SIN_ODD)) {
    const double* x = src + (int64_t)b * sstride;
    for (int j = 0; j < nt; ++j) {
      const double wj = bt0 == COS_EVEN ? (j == 0 ? M_PI : 0.0) : 2.0 / (2.0 * j + 1.0);
      if (wj == 0.0) continue;
      double v = 0.0, sg = 1.0;
      for (int i = 0; i < nr; ++i) {
        v += sg * x[i + (int64_t)nr * j];
        sg *= bsign;
      }
      tot += wj * v;
    }
  }
  const double val = tot * 2.0 * M_PI * fac;
  if (w) {
    ((double*)out)[col0 + b] += w[row] * val;
    return;
  }
  T* o = out + row + (col0 + b) * ldo;
  *o = acc ? (T)((double)*o + val + c0) : (T)(val + c0);
}

// Double unknowns' fields (constant fields of the value arena): v[e] += scale * X[k].
__global__ void double_update_kernel(double* __restrict__ v, int size, const double* __restrict__ X, int k,
                                     double scale) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < size) v[e] += scale * X[k];
}

struct DevPlanX {
  kad::Passes P;
  double* pool = nullptr;
};

struct DevProgram {
  std::vector<Instr> code;       // slots of tangent operands remapped to physical slots
  int nphys = 0;
  int seed_tab = -1;
  int field = 0, dom = 0, off = 0, cnt = 0;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct kad_system_impl {
  kad::Ctx* ctx = nullptr;
  SystemBase* sys = nullptr;
  XctsSystem* xsys = nullptr;  // the same object when the system is the hand-traced XCTS one
  int B = 256;
  int64_t stride = 0;          // doubles per slot per batch element
  double* val_arena = nullptr;  // nval_slots * stride
  double* tan_arena = nullptr;  // nphys_max * B * stride
  double* F = nullptr;          // n_rows
  double* scratch = nullptr;    // small
  std::vector<DevPlanX> plans;
  std::vector<DevRow*> row_tabs;
  std::vector<int*> seed_pos;
  std::vector<double*> seed_val;
  std::vector<int*> seed_cnt;
  std::vector<int*> csr_rowptr, csr_col;
  std::vector<double*> csr_w;
  std::vector<DevProgram> tprogs;
  std::vector<DevProgram> obs_tprogs;  // d(observables)/d(unknowns)
  std::vector<DevProgram> par_progs;   // dF/dp of each parameter
  std::vector<DevProgram> glob_tprogs;      // dQ/du of the global quantities, per layout entry
  std::vector<DevProgram> glob_par_progs;   // dQ/dp at fixed fields, per parameter
  double* dQ = nullptr;                // 3 x n, row q contiguous
  double* dFdp = nullptr;              // n per parameter
  int nphys_max = 0;
  // dense linear algebra buffers (lazily allocated)
  double* J = nullptr;
  float* Jlo = nullptr;
  int* ipiv = nullptr;
  int* perm = nullptr;
  double* X = nullptr;
  double* Fsave = nullptr;
  double* Jt = nullptr;   // Gauss-Newton: J^T and J^T J
  double* JtJ = nullptr;
  double* held = nullptr;   // line search: the unknowns' storage at the start of the search
  size_t held_cap = 0;
  double* Rs = nullptr;  // equilibration scales
  double* Cs = nullptr;
  std::vector<void*> allocs;
  double last_times[4] = {0, 0, 0, 0};
  int last_iters[2] = {0, 0};
  double last_resid = 0;
  int last_fallback = 0;  // the last low-precision step fell back to an fp64 factorisation
  int last_factored = 0;  // the last step factored J (0: it reused earlier low-precision factors)
  // low-precision solve options (kad_system_set_option)
  double forcing = 0.0;   // refinement target max(1e-13, forcing * min(1, ||F||_inf)); 0: 1e-13
  int reuse_inner = 0;    // reuse the low-precision factors while the previous refinement took
                          // at most this many inner steps (0: factor every step)
  bool lo_valid = false;  // Jlo / perm / Rs / Cs hold usable low-precision factors
  int lo_prec = -1;
  int prev_inner = 0;     // inner steps of the previous refinement
  double* Jfac = nullptr;  // solver 5: fp64 factors of an earlier Jacobian (J and Jfac alternate)
  bool fac_valid = false;  // Jfac / perm hold usable fp64 factors
  kad_dd* dd = nullptr;    // solvers 6-7: the domain-decomposed factors (their own copy of the blocks)
  bool dd_valid = false;
  // solvers 6-7: factor with the pivot orders of the solve's last factorisation with
  // partial pivoting (static pivoting, kad_dd_factor_static) once there is one
  bool dd_static = true;
  bool dd_orders = false;  // this solve has such a factorisation (option 2 forgets it)
  // Jacobian fill on several streams (kad_system_set_option 3): the tangent programs are
  // dealt to the streams by estimated cost, each stream with its own tangent arena;
  // every program writes only its own columns of J, so the streams never touch the
  // same memory.  Stream 0 is the context's stream with tan_arena.
  bool fuse_ew = true;  // apply marked element-wise instructions in the preceding pass (option 5)
  static constexpr int kFillMax = 4;
  int nfill = 3;  // 17x13x12 x 5 domains, batch 4096: 0.241 s on one stream, 0.207 / 0.201 / 0.210 on 2 / 3 / 4
  hipStream_t fstream[kFillMax] = {};
  double* ftan[kFillMax] = {};
  hipEvent_t fev[kFillMax] = {};

  ~kad_system_impl() {
    if (dd) kad_dd_destroy(dd);
    for (int i = 0; i < kFillMax; ++i) {
      if (fstream[i]) (void)hipStreamDestroy(fstream[i]);
      if (fev[i]) (void)hipEventDestroy(fev[i]);
    }
    for (void* p : allocs) (void)hipFree(p);
    delete sys;
  }
  int dmalloc(void** p, size_t bytes) {
    KAD_HIP_CHECK(hipMalloc(p, bytes < 16 ? 16 : bytes));
    allocs.push_back(*p);
    return KAD_OK;
  }
  template <typename T>
  int upload(const std::vector<T>& h, T** d) {
    int rc = dmalloc((void**)d, h.size() * sizeof(T));
    if (rc) return rc;
    if (!h.empty()) KAD_HIP_CHECK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return KAD_OK;
  }
  double* slot_ptr(const Opnd& o, double* tan) const {
    return o.space == SP_VAL ? val_arena + (int64_t)o.slot * stride : tan + (int64_t)o.slot * B * stride;
  }
  int64_t slot_stride(const Opnd& o) const { return o.space == SP_VAL ? 0 : stride; }
  int setup();
  // out: fp64 (dtype 0) or fp32 (dtype 1) column-major rows x columns with leading
  // dimension ldo, columns from col0; with w != nullptr the exports are transposed
  // instead (out: fp64 per column, += w . row).  st / tan: the stream and tangent arena
  // (default: the context's stream and tan_arena).
  int run(const std::vector<Instr>& code, int nb, int u0, int seed_tab, void* out, int64_t col0, int64_t ldo,
          int dtype = 0, const double* w = nullptr, hipStream_t st = nullptr, double* tan = nullptr);
  // every tangent program's columns of J (fp64, ld ldj), after the value pass and the
  // double unknowns' columns, on nfill streams; the context's stream waits for all
  int fill_columns(double* Jd, int64_t ldj);
  // the side streams (and, with arenas, their tangent arenas) and events for ns streams
  int ensure_streams(int ns, bool arenas);
  // the value program split by domain over ns streams (value_pass); false: not splittable
  bool split_values(int ns);
  int vsplit_ns = 0;
  bool vsplit_ok = false;
  std::vector<std::vector<Instr>> vparts;
  std::vector<Instr> vexports;
  bool vals_current = false;  // the value arena holds the value pass of the current fields
  int ensure_values() { return vals_current ? KAD_OK : value_pass(); }
  int value_pass();
  int jacobian(double* Jd, int64_t ldj);
  int update(const double* Xd, double scale);
  // columns [c0, c1) of the double unknowns (their parameter programs) into J, whose
  // column c0 sits at Jout + (c0 - col_base) ld
  int double_columns(void* Jout, int64_t ld, int c0, int c1, int col_base, int dtype = 0) {
    for (int k = c0; k < c1 && k < sys->ndouble; ++k) {
      const DevProgram& pp = par_progs[sys->dvar_par[k]];
      int rc = run(pp.code, 1, 0, pp.seed_tab, Jout, (int64_t)k - col_base, ld, dtype);
      if (rc) return rc;
    }
    return KAD_OK;
  }
};

// Chains of element-wise instructions that run inside the instruction before them: an
// instruction (a spectral pass or an element-wise one) whose result feeds only the
// element-wise instruction right after it absorbs that instruction, which then applies
// to the value as it is stored, and so on up to kad::kMaxChain of them.  The
// intermediate arrays never go to HBM (one launch and two array transfers fewer per
// absorbed instruction).  Conditions, on the virtual slots (before allocate_slots):
// tangent destinations (the slots are the program's own), the result is the next
// instruction's a or b but not both, it is read nowhere else, and the next
// instruction covers the whole array.  code[t].fuse = the number absorbed.
// A pass read by a single element-wise instruction a few places later moves down to just
// before it (the slots are written once, so the instructions in between read neither
// its result nor anything it overwrites), unless the instruction before that reader
// already feeds it.
static void pull_passes_to_readers(std::vector<Instr>& code, int nvirt) {
  constexpr int kReach = 3;
  auto reader_of = [&](int slot, size_t from) -> long {
    long at = -1;
    for (size_t u = from; u < code.size(); ++u) {
      const Instr& in = code[u];
      const bool ra = in.a.space == SP_TAN && in.a.slot == slot, rb = in.b.slot >= 0 && in.b.space == SP_TAN && in.b.slot == slot;
      if (ra || rb) {
        if (at >= 0) return -1;  // read twice
        at = (long)u;
      }
    }
    return at;
  };
  (void)nvirt;
  for (size_t t = 0; t + 2 < code.size(); ++t) {
    const Instr& l = code[t];
    if (l.kind != I_LIN || l.dst.space != SP_TAN || l.dst.slot < 0) continue;
    const long u = reader_of(l.dst.slot, t + 1);
    if (u < (long)t + 2 || u > (long)t + 1 + kReach || code[u].kind != I_EW) continue;
    const Instr& e = code[u];
    const Instr& pv = code[u - 1];
    const bool fed_by_prev = pv.dst.space == SP_TAN && pv.dst.slot >= 0 &&
                             ((e.a.space == SP_TAN && e.a.slot == pv.dst.slot) ||
                              (e.b.slot >= 0 && e.b.space == SP_TAN && e.b.slot == pv.dst.slot));
    if (fed_by_prev) continue;
    std::rotate(code.begin() + t, code.begin() + t + 1, code.begin() + u);  // the pass now at u - 1
    --t;  // look again at what moved into position t
  }
}

static void mark_fusions(std::vector<Instr>& code, int nvirt, const std::vector<DevPlanX>& plans) {
  pull_passes_to_readers(code, nvirt);
  std::vector<int> reads(nvirt, 0);
  for (Instr& in : code) {
    in.fuse = 0;
    if (in.a.slot >= 0 && in.a.space == SP_TAN) reads[in.a.slot]++;
    if (in.b.slot >= 0 && in.b.space == SP_TAN) reads[in.b.slot]++;
  }
  for (size_t t = 0; t < code.size(); ++t) {
    Instr& h = code[t];
    if ((h.kind != I_LIN && h.kind != I_EW) || h.dst.space != SP_TAN || h.dst.slot < 0) continue;
    int size = h.size;
    if (h.kind == I_LIN) {
      if (h.plan < 0 || h.plan >= (int)plans.size()) continue;
      const kad::Passes& P = plans[h.plan].P;
      int od[3] = {h.din[0], h.din[1], h.din[2]};
      for (int p = 0; p < P.npass; ++p) od[P.axis[p]] = P.nout[p];
      size = od[0] * od[1] * od[2];
    }
    int m = 0;
    while (m < kad::kMaxChain && t + m + 1 < code.size()) {
      const Instr& prev = code[t + m];
      const Instr& e = code[t + m + 1];
      if (e.kind != I_EW || e.dst.space != SP_TAN || e.size != size || e.ew < 0 || e.ew > EW_SDIV) break;
      const bool in_a = e.a.space == SP_TAN && e.a.slot == prev.dst.slot;
      const bool in_b = e.b.slot >= 0 && e.b.space == SP_TAN && e.b.slot == prev.dst.slot;
      if (in_a == in_b || reads[prev.dst.slot] != 1) break;
      ++m;
    }
    h.fuse = m;
    t += m;  // the absorbed instructions run inside this one
  }
}

// The absorbed instructions code[t+1 .. t+fuse] as steps on the value of code[t].
static kad::PostOp chain_of(const std::vector<Instr>& code, size_t t, const std::function<double*(const Opnd&)>& ptr,
                            const std::function<int64_t(const Opnd&)>& stride) {
  kad::PostOp post;
  const int m = code[t].fuse;
  for (int k = 0; k < m; ++k) {
    const Instr& prev = code[t + k];
    const Instr& e = code[t + k + 1];
    // the chain value is prev's result; the other operand is the instruction's other one
    const int side = (e.a.space == SP_TAN && e.a.slot == prev.dst.slot) ? 0 : 1;
    const Opnd& other = side == 0 ? e.b : e.a;
    kad::EwStep& st = post.op[k];
    st.kind = e.ew;
    st.side = side;
    st.o = other.slot >= 0 ? ptr(other) : nullptr;
    st.os = other.slot >= 0 ? stride(other) : 0;
    st.s = e.s;
  }
  post.n = m;
  post.dst = ptr(code[t + m].dst);
  post.ds = stride(code[t + m].dst);
  return post;
}

// Liveness-based reuse of tangent slots within one program.
static int allocate_slots(std::vector<Instr>& code, int nvirt, const std::vector<DevPlanX>& plans) {
  mark_fusions(code, nvirt, plans);
  std::vector<int> last(nvirt, -1);
  auto use = [&](const Opnd& o, int t) {
    if (o.slot >= 0 && o.space == SP_TAN) last[o.slot] = std::max(last[o.slot], t);
  };
  for (int t = 0; t < (int)code.size(); ++t) {
    const Instr& in = code[t];
    use(in.a, t);
    use(in.b, t);
    use(in.dst, t);
  }
  std::vector<int> phys(nvirt, -1), freelist;
  int nphys = 0;
  for (int t = 0; t < (int)code.size(); ++t) {
    Instr& in = code[t];
    auto map_read = [&](Opnd& o) {
      if (o.slot >= 0 && o.space == SP_TAN) o.slot = phys[o.slot];
    };
    const int dv = (in.dst.space == SP_TAN) ? in.dst.slot : -1;
    const int av = (in.a.space == SP_TAN) ? in.a.slot : -1;
    const int bv = (in.b.space == SP_TAN) ? in.b.slot : -1;
    map_read(in.a);
    map_read(in.b);
    if (dv >= 0) {
      if (phys[dv] < 0) {
        if (!freelist.empty()) {
          phys[dv] = freelist.back();
          freelist.pop_back();
        } else {
          phys[dv] = nphys++;
        }
      }
      in.dst.slot = phys[dv];
    }
    // release operands whose last use is this instruction (after the destination is taken,
    // so an instruction never writes the slot it reads)
    for (int v : {av, bv}) {
      if (v >= 0 && last[v] == t && v != dv) {
        freelist.push_back(phys[v]);
        last[v] = -2;
      }
    }
    if (dv >= 0 && last[dv] == t) {
      freelist.push_back(phys[dv]);
      last[dv] = -2;
    }
  }
  return nphys;
}

int kad_system_impl::setup() {
  Tracer& tr = sys->tr;
  // slot stride: the largest domain array (coef layout), rounded to 8 doubles
  int maxsz = 0;
  for (int d = sys->dmin; d <= sys->dmax; ++d) maxsz = std::max(maxsz, sys->doms[d].coef_size());
  stride = (maxsz + 7) / 8 * 8;
  int rc = dmalloc((void**)&val_arena, (size_t)std::max(1, tr.nval_slots) * stride * sizeof(double));
  if (rc) return rc;
  KAD_HIP_CHECK(hipMemset(val_arena, 0, (size_t)std::max(1, tr.nval_slots) * stride * sizeof(double)));
  for (auto& cs : tr.const_slots)
    KAD_HIP_CHECK(hipMemcpy(val_arena + (int64_t)cs.first * stride, cs.second.data(),
                            cs.second.size() * sizeof(double), hipMemcpyHostToDevice));
  // plans
  for (auto& hp : tr.plans) {
    DevPlanX dp;
    if ((rc = upload(hp.pool, &dp.pool))) return rc;
    dp.P.npass = (int)hp.passes.size();
    dp.P.maxsize = hp.maxsize;
    dp.P.pool_size = (int)hp.pool.size();
    if (dp.P.npass > kad::MAX_PASSES) return KAD_ERR_ARG;
    for (int p = 0; p < dp.P.npass; ++p) {
      int* lm;
      if ((rc = upload(hp.passes[p].line_mat, &lm))) return rc;
      dp.P.axis[p] = hp.passes[p].axis;
      dp.P.nout[p] = hp.passes[p].nout;
      dp.P.line_mat[p] = lm;
      dp.P.nlines[p] = (int)hp.passes[p].line_mat.size();
    }
    plans.push_back(dp);
  }
  for (auto& rt : tr.row_tabs) {
    std::vector<DevRow> h(rt.size());
    for (size_t i = 0; i < rt.size(); ++i) h[i] = {rt[i].pos, rt[i].gpos, rt[i].gw};
    DevRow* d;
    if ((rc = upload(h, &d))) return rc;
    row_tabs.push_back(d);
  }
  for (auto& st : sys->seeds) {
    int* p;
    double* v;
    int* c;
    if ((rc = upload(st.pos, &p)) || (rc = upload(st.val, &v)) || (rc = upload(st.cnt, &c))) return rc;
    seed_pos.push_back(p);
    seed_val.push_back(v);
    seed_cnt.push_back(c);
  }
  for (auto& m : sys->tau_maps) {
    int *rp, *cl;
    double* w;
    if ((rc = upload(m.rowptr, &rp)) || (rc = upload(m.col, &cl)) || (rc = upload(m.w, &w))) return rc;
    csr_rowptr.push_back(rp);
    csr_col.push_back(cl);
    csr_w.push_back(w);
  }
  // tangent programs: physical slots
  for (size_t p = 0; p < sys->tprog.size(); ++p) {
    DevProgram dp;
    dp.code = sys->tprog[p].code;
    dp.nphys = allocate_slots(dp.code, sys->tprog[p].nslots, plans);
    dp.field = sys->layout[p][0];
    dp.dom = sys->layout[p][1];
    dp.off = sys->layout[p][2];
    dp.cnt = sys->layout[p][3];
    for (auto& in : dp.code)
      if (in.kind == I_SEED) dp.seed_tab = in.seed_tab;
    nphys_max = std::max(nphys_max, dp.nphys);
    tprogs.push_back(std::move(dp));
  }
  for (size_t p = 0; p < sys->obs_tprog.size(); ++p) {
    DevProgram dp;
    dp.code = sys->obs_tprog[p].code;
    dp.nphys = allocate_slots(dp.code, sys->obs_tprog[p].nslots, plans);
    dp.field = sys->obs_tprog_lay[p][0];
    dp.dom = sys->obs_tprog_lay[p][1];
    dp.off = sys->obs_tprog_lay[p][2];
    dp.cnt = sys->obs_tprog_lay[p][3];
    dp.seed_tab = sys->obs_tprog_tab[p];
    nphys_max = std::max(nphys_max, dp.nphys);
    obs_tprogs.push_back(std::move(dp));
  }
  for (size_t p = 0; p < sys->par_prog.size(); ++p) {
    DevProgram dp;
    dp.code = sys->par_prog[p].code;
    dp.nphys = allocate_slots(dp.code, sys->par_prog[p].nslots, plans);
    dp.seed_tab = sys->par_tab[p];
    dp.cnt = 1;
    nphys_max = std::max(nphys_max, dp.nphys);
    par_progs.push_back(std::move(dp));
  }
  for (size_t p = 0; p < sys->glob_tprog.size(); ++p) {
    DevProgram dp;
    dp.code = sys->glob_tprog[p].code;
    dp.nphys = allocate_slots(dp.code, sys->glob_tprog[p].nslots, plans);
    dp.field = sys->layout[p][0];
    dp.dom = sys->layout[p][1];
    dp.off = sys->layout[p][2];
    dp.cnt = sys->layout[p][3];
    dp.seed_tab = tprogs[p].seed_tab;
    nphys_max = std::max(nphys_max, dp.nphys);
    glob_tprogs.push_back(std::move(dp));
  }
  for (size_t p = 0; p < sys->glob_par_prog.size(); ++p) {
    DevProgram dp;
    dp.code = sys->glob_par_prog[p].code;
    dp.nphys = allocate_slots(dp.code, sys->glob_par_prog[p].nslots, plans);
    dp.seed_tab = sys->par_tab[p];
    dp.cnt = 1;
    nphys_max = std::max(nphys_max, dp.nphys);
    glob_par_progs.push_back(std::move(dp));
  }
  if ((rc = dmalloc((void**)&tan_arena, (size_t)std::max(1, nphys_max) * B * stride * sizeof(double)))) return rc;
  if ((rc = dmalloc((void**)&F, (size_t)sys->n_rows * sizeof(double)))) return rc;
  if ((rc = dmalloc((void**)&scratch, 64 * sizeof(double)))) return rc;
  return KAD_OK;
}

int kad_system_impl::run(const std::vector<Instr>& code, int nb, int u0, int seed_tab, void* out, int64_t col0,
                         int64_t ldo, int dtype, const double* w, hipStream_t st, double* tan) {
  hipStream_t s = st ? st : ctx->stream;
  if (!tan) tan = tan_arena;
  auto slot_ptr = [&](const Opnd& o) { return this->slot_ptr(o, tan); };
  for (size_t t = 0; t < code.size(); ++t) {
    const Instr& in = code[t];
    const Domain& dm = sys->doms[in.dom];
    switch (in.kind) {
      case I_LIN: {
        const DevPlanX& p = plans[in.plan];
        const int batch = (in.dst.space == SP_VAL) ? 1 : nb;
        kad::PostOp post;
        if (in.fuse > 0 && fuse_ew) {
          // the following element-wise instructions, applied in this pass's store phase
          post = chain_of(code, t, slot_ptr, [&](const Opnd& o) { return slot_stride(o); });
          t += in.fuse;
        }
        int rc = kad::apply3d_launch(slot_ptr(in.a), stride, slot_ptr(in.dst), stride, batch, in.din[0], in.din[1],
                                     in.din[2], p.P, p.pool, 0, 1.0, s, &post);
        if (rc) return rc;
        break;
      }
      case I_EW: {
        const int batch = (in.dst.space == SP_VAL) ? 1 : nb;
        double* dp = slot_ptr(in.dst);
        const double* ap = slot_ptr(in.a);
        const double* bp = in.b.slot >= 0 ? slot_ptr(in.b) : nullptr;
        const int64_t ds = slot_stride(in.dst), as = slot_stride(in.a), bs = in.b.slot >= 0 ? slot_stride(in.b) : 0;
        if (in.fuse > 0 && fuse_ew) {
          // this instruction and the ones it absorbed, in one pass over the arrays
          const kad::PostOp post = chain_of(code, t, slot_ptr, [&](const Opnd& o) { return slot_stride(o); });
          t += in.fuse;
          const dim3 grid((unsigned)kad::cdiv(in.size, 256 * EW_U), (unsigned)std::min(batch, 65535));
          ewc_kernel<<<grid, 256, 0, s>>>(in.ew, ap, as, bp, bs, in.s, post, in.size, batch);
          break;
        }
        auto a16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
        if (in.size % 2 == 0 && ds % 2 == 0 && as % 2 == 0 && bs % 2 == 0 && a16(dp) && a16(ap) &&
            (!bp || a16(bp))) {
          const int size2 = in.size / 2;
          const dim3 grid((unsigned)kad::cdiv(size2, 256 * EW_U), (unsigned)std::min(batch, 65535));
          ew2_kernel<<<grid, 256, 0, s>>>(in.ew, dp, ds, ap, as, bp, bs, in.s, size2, batch);
        } else {
          const dim3 grid((unsigned)kad::cdiv(in.size, 256), (unsigned)std::min(batch, 65535));
          ew_kernel<<<grid, 256, 0, s>>>(in.ew, dp, ds, ap, as, bp, bs, in.s, in.size, batch);
        }
        break;
      }
      case I_EXPORT_IN:
      case I_EXPORT_BD: {
        const int batch = (in.a.space == SP_VAL) ? 1 : nb;
        const int64_t tot = (int64_t)in.nrows * batch;
        const unsigned g = (unsigned)kad::cdiv(tot, 256);
        if (w) {
          export_dot_kernel<<<(unsigned)batch, 256, 0, s>>>(row_tabs[in.rows], in.nrows, in.kind == I_EXPORT_BD,
                                                            slot_ptr(in.a), slot_stride(in.a), dm.nr, in.bsign, w,
                                                            in.row0, (double*)out, col0, in.sign);
        } else if (in.kind == I_EXPORT_IN) {
          if (dtype == 0)
            export_in_kernel<double><<<g, 256, 0, s>>>(row_tabs[in.rows], in.nrows, slot_ptr(in.a), slot_stride(in.a),
                                                       (double*)out, in.row0, col0, ldo, in.sign, in.accumulate,
                                                       batch);
          else
            export_in_kernel<float><<<g, 256, 0, s>>>(row_tabs[in.rows], in.nrows, slot_ptr(in.a), slot_stride(in.a),
                                                      (float*)out, in.row0, col0, ldo, in.sign, in.accumulate, batch);
        } else {
          if (dtype == 0)
            export_bd_kernel<double><<<g, 256, 0, s>>>(row_tabs[in.rows], in.nrows, slot_ptr(in.a), slot_stride(in.a),
                                                       dm.nr, in.bsign, (double*)out, in.row0, col0, ldo, in.sign,
                                                       in.accumulate, batch);
          else
            export_bd_kernel<float><<<g, 256, 0, s>>>(row_tabs[in.rows], in.nrows, slot_ptr(in.a), slot_stride(in.a),
                                                      dm.nr, in.bsign, (float*)out, in.row0, col0, ldo, in.sign,
                                                      in.accumulate, batch);
        }
        break;
      }
      case I_EXPORT_INT: {
        if (w || dtype != 0) return KAD_ERR_ARG;
        const int batch = (in.a.space == SP_VAL) ? 1 : nb;
        export_int_kernel<<<(unsigned)kad::cdiv(batch, 64), 64, 0, s>>>(slot_ptr(in.a), slot_stride(in.a), dm.nr,
                                                                         dm.nt, in.aux, in.s, (double*)out, in.row0,
                                                                         col0, ldo, batch);
        break;
      }
      case I_EXPORT_GL: {
        const bool val = in.a.slot < 0 || in.a.space == SP_VAL;
        const int batch = val ? 1 : nb;
        const double* src = in.a.slot >= 0 ? slot_ptr(in.a) : nullptr;
        const int64_t ss = in.a.slot >= 0 ? slot_stride(in.a) : 0;
        const unsigned g = (unsigned)kad::cdiv(batch, 64);
        if (w || dtype == 0)
          export_gl_kernel<double><<<g, 64, 0, s>>>(src, ss, dm.nr, dm.nt, in.aux, in.bsign, in.s, in.c0,
                                                    in.accumulate, in.pt, (double*)out, in.row0, col0, ldo, w,
                                                    batch);
        else
          export_gl_kernel<float><<<g, 64, 0, s>>>(src, ss, dm.nr, dm.nt, in.aux, in.bsign, in.s, in.c0,
                                                   in.accumulate, in.pt, (float*)out, in.row0, col0, ldo, nullptr,
                                                   batch);
        break;
      }
      case I_SEED: {
        const int64_t tot = (int64_t)in.size * nb;
        const int nu = (int)sys->seeds[seed_tab].cnt.size();
        seed_kernel<<<(unsigned)kad::cdiv(tot, 256), 256, 0, s>>>(slot_ptr(in.dst), stride, in.size,
                                                                    seed_pos[seed_tab], seed_val[seed_tab],
                                                                    seed_cnt[seed_tab], u0, nu, nb);
        break;
      }
      default:
        return KAD_ERR_ARG;
    }
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

int kad_system_impl::ensure_streams(int ns, bool arenas) {
  for (int i = 1; i < ns; ++i) {
    if (!fstream[i]) KAD_HIP_CHECK(hipStreamCreateWithFlags(&fstream[i], hipStreamNonBlocking));
    if (arenas && !ftan[i]) {
      const int rc = dmalloc((void**)&ftan[i], (size_t)std::max(1, nphys_max) * B * stride * sizeof(double));
      if (rc) return rc;
    }
  }
  for (int i = 0; i < ns && ns > 1; ++i)
    if (!fev[i]) KAD_HIP_CHECK(hipEventCreateWithFlags(&fev[i], hipEventDisableTiming));
  return KAD_OK;
}

// The value program split over ns streams: every instruction but the row exports goes to
// the stream of its domain, in program order; the exports follow on the context's
// stream once all have finished.  Valid when no instruction reads a value slot that
// an instruction of another stream writes (the domains' arrays are their own) and the
// exports feed nothing but their rows.
bool kad_system_impl::split_values(int ns) {
  if (vsplit_ns == ns) return vsplit_ok;
  vsplit_ns = ns;
  vsplit_ok = false;
  vparts.assign(ns, {});
  vexports.clear();
  std::vector<int> writer(std::max(1, sys->tr.nval_slots), -1);
  auto is_export = [](const Instr& in) {
    return in.kind == I_EXPORT_IN || in.kind == I_EXPORT_BD || in.kind == I_EXPORT_INT || in.kind == I_EXPORT_GL;
  };
  for (const Instr& in : sys->tr.val.code) {
    if (is_export(in)) {
      vexports.push_back(in);
      continue;
    }
    const int k = ((in.dom - sys->dmin) % ns + ns) % ns;
    for (const Opnd* o : {&in.a, &in.b})
      if (o->slot >= 0 && o->space == SP_VAL && o->slot < (int)writer.size() && writer[o->slot] >= 0 &&
          writer[o->slot] != k)
        return false;
    if (in.dst.slot >= 0) {
      if (in.dst.space != SP_VAL || in.dst.slot >= (int)writer.size()) return false;
      writer[in.dst.slot] = k;
    }
    vparts[k].push_back(in);
  }
  int used = 0;
  for (auto& v : vparts) used += !v.empty();
  vsplit_ok = used > 1;
  return vsplit_ok;
}

int kad_system_impl::value_pass() {
  int rc;
  const int ns = std::max(1, std::min(nfill, kFillMax));
  if (ns > 1 && split_values(ns)) {
    hipStream_t s0 = ctx->stream;
    if ((rc = ensure_streams(ns, false))) return rc;
    KAD_HIP_CHECK(hipEventRecord(fev[0], s0));
    for (int i = 1; i < ns; ++i) KAD_HIP_CHECK(hipStreamWaitEvent(fstream[i], fev[0], 0));
    for (int i = 0; i < ns && rc == KAD_OK; ++i)
      rc = run(vparts[i], 1, 0, -1, F, 0, 0, 0, nullptr, i ? fstream[i] : s0, tan_arena);
    for (int i = 1; i < ns; ++i) {
      KAD_HIP_CHECK(hipEventRecord(fev[i], fstream[i]));
      KAD_HIP_CHECK(hipStreamWaitEvent(s0, fev[i], 0));
    }
    if (rc == KAD_OK) rc = run(vexports, 1, 0, -1, F, 0, 0);
  } else {
    rc = run(sys->tr.val.code, 1, 0, -1, F, 0, 0);
  }
  vals_current = (rc == KAD_OK);
  return rc;
}

int kad_system_impl::jacobian(double* Jd, int64_t ldj) {
  int rc = value_pass();
  if (rc) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(Jd, 0, (size_t)ldj * sys->n_unknowns * sizeof(double), ctx->stream));
  if ((rc = double_columns(Jd, ldj, 0, sys->ndouble, 0))) return rc;
  return fill_columns(Jd, ldj);
}

int kad_system_impl::fill_columns(double* Jd, int64_t ldj) {
  hipStream_t s0 = ctx->stream;
  const int ns = std::max(1, std::min(nfill, (int)tprogs.size()));
  {
    const int rc = ensure_streams(ns, true);
    if (rc) return rc;
  }
  if (ns > 1) {
    KAD_HIP_CHECK(hipEventRecord(fev[0], s0));  // J cleared, value pass done
    for (int i = 1; i < ns; ++i) KAD_HIP_CHECK(hipStreamWaitEvent(fstream[i], fev[0], 0));
  }
  // longest programs first, each to the stream with the least work so far
  std::vector<int> order(tprogs.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = (int)q;
  auto cost = [&](int q) { return (double)tprogs[q].cnt * (double)tprogs[q].code.size(); };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
  double load[kFillMax] = {0, 0, 0, 0};
  for (int q : order) {
    int i = 0;
    for (int k = 1; k < ns; ++k)
      if (load[k] < load[i]) i = k;
    load[i] += cost(q);
    const DevProgram& p = tprogs[q];
    for (int u0 = 0; u0 < p.cnt; u0 += B) {
      const int nb = std::min(B, p.cnt - u0);
      // exports write columns off+u0 .. off+u0+nb-1
      const int rc = run(p.code, nb, u0, p.seed_tab, Jd, (int64_t)p.off + u0, ldj, 0, nullptr, i ? fstream[i] : s0,
                         i ? ftan[i] : tan_arena);
      if (rc) return rc;
    }
  }
  for (int i = 1; i < ns; ++i) {
    KAD_HIP_CHECK(hipEventRecord(fev[i], fstream[i]));
    KAD_HIP_CHECK(hipStreamWaitEvent(s0, fev[i], 0));
  }
  return KAD_OK;
}

int kad_system_impl::update(const double* Xd, double scale) {
  vals_current = false;
  for (int k = 0; k < sys->ndouble; ++k) {
    const std::string& name = sys->par_names[sys->dvar_par[k]];
    for (int d = sys->dmin; d <= sys->dmax; ++d) {
      const int slot = sys->param_slot(name, d);
      if (slot < 0) continue;
      const int size = sys->doms[d].conf_size();
      double_update_kernel<<<(unsigned)kad::cdiv(size, 256), 256, 0, ctx->stream>>>(val_arena + (int64_t)slot * stride,
                                                                                   size, Xd, k, scale);
    }
  }
  for (size_t q = 0; q < sys->layout.size(); ++q) {
    const int f = sys->layout[q][0], d = sys->layout[q][1], off = sys->layout[q][2];
    const int ncf = sys->doms[d].coef_size();
    double* cf = val_arena + (int64_t)sys->field_slot[f][d] * stride;
    csr_update_kernel<<<(unsigned)kad::cdiv(ncf, 256), 256, 0, ctx->stream>>>(cf, csr_rowptr[q], csr_col[q], csr_w[q],
                                                                              Xd, off, ncf, scale);
  }
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

extern "C" {

int kad_xcts_create(kad_ctx* ctx, int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax,
                    double n0, double a, double ome, int order, int batch, kad_system** out) {
  if (!ctx || !bounds || nbounds < 2 || !out || batch <= 0 || order < 0 || order > 1) return KAD_ERR_ARG;
  auto* S = new kad_system_impl();
  S->ctx = reinterpret_cast<kad::Ctx*>(ctx);
  S->B = batch;
  XctsParams p;
  p.nr = nr;
  p.nt = nt;
  p.np = np;
  p.bounds.assign(bounds, bounds + nbounds);
  p.dmin = dmin;
  p.dmax = dmax;
  p.n0 = n0;
  p.a = a;
  p.ome = ome;
  p.order = order;
  try {
    S->xsys = new XctsSystem(p);
    S->sys = S->xsys;
    S->sys->build();
  } catch (const std::exception& e) {
    fprintf(stderr, "[kadath_amd] kad_xcts_create: %s\n", e.what());
    delete S;
    return KAD_ERR_ARG;
  }
  int rc = S->setup();
  if (rc) {
    delete S;
    return rc;
  }
  *out = reinterpret_cast<kad_system*>(S);
  return KAD_OK;
}

int kad_xcts_destroy(kad_system* s) {
  if (!s) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  (void)hipStreamSynchronize(S->ctx->stream);
  delete S;
  return KAD_OK;
}

int kad_xcts_sizes(kad_system* s, int* n_unknowns, int* n_rows, int* stats) {
  if (!s) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (n_unknowns) *n_unknowns = S->sys->n_unknowns;
  if (n_rows) *n_rows = S->sys->n_rows;
  if (stats) {
    size_t tcode = 0;
    for (auto& p : S->tprogs) tcode += p.code.size();
    stats[0] = (int)S->sys->tr.val.code.size();
    stats[1] = (int)tcode;
    stats[2] = S->sys->tr.nval_slots;
    stats[3] = S->nphys_max;
  }
  return KAD_OK;
}

// layout: 4 ints per (field, domain): field (0 P, 1 N, 2-4 bet x,y,z), domain, offset, count
int kad_xcts_layout(kad_system* s, int* out, int cap) {
  if (!s || !out) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  const int n = (int)S->sys->layout.size();
  if (cap < 4 * n) return KAD_ERR_ARG;
  for (int q = 0; q < n; ++q)
    for (int c = 0; c < 4; ++c) out[4 * q + c] = S->sys->layout[q][c];
  return n;
}

// 5 ints per (equation, component), in row order: kind (0 inside, 1 matching
// d|d+1, 2 outer bc, 3 inner bc), domain, component, first row, row count.
// Returns the number of entries.
int kad_xcts_row_layout(kad_system* s, int* out, int cap) {
  if (!s || !out) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  int q = 0;
  for (const auto& e : S->sys->eqs) {
    int row = e.row0;
    for (int c = 0; c < e.ncomp; ++c) {
      if (5 * (q + 1) > cap) return KAD_ERR_ARG;
      const int v[5] = {e.kind, e.d, c, row, e.nrows[c]};
      for (int t = 0; t < 5; ++t) out[5 * q + t] = v[t];
      row += e.nrows[c];
      ++q;
    }
  }
  return q;
}

// Coefficients of configuration values on the host, with the device's tables
// (Base_spectral::coef, coef.cpp:96-104: phi, then theta, then r).
static std::vector<double> host_coef(const Domain& dm, const Base& b, const std::vector<double>& conf) {
  Mat Mp = coef_matrix(b.bp, dm.np);
  std::vector<double> t1((size_t)dm.nr * dm.nt * dm.npc, 0.0), t2(t1.size(), 0.0), t3(t1.size(), 0.0);
  for (int kk = 0; kk < dm.npc; ++kk)
    for (int k = 0; k < dm.np; ++k)
      for (int j = 0; j < dm.nt; ++j)
        for (int i = 0; i < dm.nr; ++i)
          t1[i + dm.nr * (j + (size_t)dm.nt * kk)] += Mp(kk, k) * conf[i + dm.nr * (j + (size_t)dm.nt * k)];
  for (int k = 0; k < dm.npc; ++k) {
    Mat Mt = coef_matrix(b.bt[k], dm.nt);
    for (int jj = 0; jj < dm.nt; ++jj)
      for (int j = 0; j < dm.nt; ++j)
        for (int i = 0; i < dm.nr; ++i)
          t2[i + dm.nr * (jj + (size_t)dm.nt * k)] += Mt(jj, j) * t1[i + dm.nr * (j + (size_t)dm.nt * k)];
  }
  Mat Mr = coef_matrix(CHEB, dm.nr);
  for (int k = 0; k < dm.npc; ++k)
    for (int j = 0; j < dm.nt; ++j)
      for (int ii = 0; ii < dm.nr; ++ii)
        for (int i = 0; i < dm.nr; ++i)
          t3[ii + dm.nr * (j + (size_t)dm.nt * k)] += Mr(ii, i) * t2[i + dm.nr * (j + (size_t)dm.nt * k)];
  return t3;
}

// fields += affecte_tau(X0), X0 HOST (n_unknowns)
static int add_unknowns_host(kad_system_impl* S, const double* X0) {
  hipStream_t st = S->ctx->stream;
  double* dX;
  KAD_HIP_CHECK(hipMalloc(&dX, (size_t)S->sys->n_unknowns * sizeof(double)));
  KAD_HIP_CHECK(hipMemcpyAsync(dX, X0, (size_t)S->sys->n_unknowns * sizeof(double), hipMemcpyHostToDevice, st));
  int rc = S->update(dX, 1.0);
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  (void)hipFree(dX);
  return rc;
}

// Fields := flat guess (P = 1 + a/r, N = 1, bet = 0; codes/Par_version/Kerr/kerr.cpp:36-52)
// + affecte_tau(X0).  X0: HOST, n_unknowns doubles (may be NULL for the flat guess).
int kad_xcts_set_guess(kad_system* s, const double* X0) {
  if (!s) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (!S->xsys) return KAD_ERR_ARG;
  XctsSystem& sy = *S->xsys;
  hipStream_t st = S->ctx->stream;
  for (int d = sy.dmin; d <= sy.dmax; ++d) {
    const Domain& dm = sy.doms[d];
    for (int f = 0; f < XctsSystem::NF; ++f) {
      std::vector<double> conf(dm.conf_size(), 0.0);
      for (size_t q = 0; q < conf.size(); ++q) {
        if (f == 0) {
          const double v = 1.0 + sy.prm.a / dm.R[q];
          conf[q] = std::isfinite(v) ? v : 1.0;
        } else if (f == 1) {
          conf[q] = 1.0;
        }
      }
      const std::vector<double> cf = host_coef(dm, sy.field_base(f, d), conf);
      S->vals_current = false;
      KAD_HIP_CHECK(hipMemcpyAsync(S->val_arena + (int64_t)sy.field_slot[f][d] * S->stride, cf.data(),
                                   cf.size() * sizeof(double), hipMemcpyHostToDevice, st));
    }
  }
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  if (X0) return add_unknowns_host(S, X0);
  return KAD_OK;
}

// sec_member: F (device, n_rows) for the current fields.
int kad_xcts_residual(kad_system* s, double* Fout) {
  if (!s || !Fout) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  int rc = S->value_pass();
  if (rc) return rc;
  KAD_HIP_CHECK(hipMemcpyAsync(Fout, S->F, (size_t)S->sys->n_rows * sizeof(double), hipMemcpyDeviceToDevice,
                               S->ctx->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

// Full Jacobian (device, column-major, ld ldj >= n_rows), all do_col_J columns.
int kad_xcts_jacobian(kad_system* s, double* Jd, int64_t ldj) {
  if (!s || !Jd) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (ldj < S->sys->n_rows) return KAD_ERR_ARG;
  int rc = S->jacobian(Jd, ldj);
  if (rc) return rc;
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

// fields -= affecte_tau(X)  (X device, n_unknowns)
int kad_xcts_update(kad_system* s, const double* Xd) {
  if (!s || !Xd) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  int rc = S->update(Xd, -1.0);
  if (rc) return rc;
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

// Coefficients of field f (0 P, 1 N, 2-4 bet) in domain d: HOST out, nr*nt*(np+2) doubles.
int kad_xcts_get_coef(kad_system* s, int f, int d, double* out) {
  if (!s || !out) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  if (f < 0 || f >= sy.nfields || d < sy.dmin || d > sy.dmax) return KAD_ERR_ARG;
  KAD_HIP_CHECK(hipMemcpyAsync(out, S->val_arena + (int64_t)sy.field_slot[f][d] * S->stride,
                               (size_t)sy.doms[d].coef_size() * sizeof(double), hipMemcpyDeviceToHost,
                               S->ctx->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

// Overwrite the coefficients of field f in domain d (HOST in, nr*nt*(np+2)): the
// counterpart of kad_xcts_get_coef, for fields prepared elsewhere (Scalar::set_domain).
int kad_xcts_set_coef(kad_system* s, int f, int d, const double* in) {
  if (!s || !in) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  if (f < 0 || f >= sy.nfields || d < sy.dmin || d > sy.dmax) return KAD_ERR_ARG;
  S->vals_current = false;
  KAD_HIP_CHECK(hipMemcpyAsync(S->val_arena + (int64_t)sy.field_slot[f][d] * S->stride, in,
                               (size_t)sy.doms[d].coef_size() * sizeof(double), hipMemcpyHostToDevice,
                               S->ctx->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

// Global quantities of the current fields at the outer boundary of the compactified
// domain (codes/Par_version/Kerr/reader.cpp:114-125): out[0] ADM mass, out[1] Komar
// mass, out[2] ADM angular momentum about z.  HOST out.
int kad_xcts_observables(kad_system* s, double* out) {
  if (!s || !out) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  if (sy.obs.empty() || sy.obs.size() > 32) return KAD_ERR_ARG;
  hipStream_t st = S->ctx->stream;
  int rc = S->value_pass();
  if (rc) return rc;
  if ((rc = S->run(sy.obs_prog.code, 1, 0, -1, S->F, 0, 0))) return rc;
  const Domain& dm = sy.doms[sy.dmax];
  const int nq = (int)sy.obs.size();
  for (int q = 0; q < nq; ++q) {
    const auto& r = sy.obs[q];
    if (r.slot < 0) {
      KAD_HIP_CHECK(hipMemsetAsync(S->scratch + q, 0, sizeof(double), st));
      continue;
    }
    surface_integral_kernel<<<1, 64, 0, st>>>(S->val_arena + (int64_t)r.slot * S->stride, dm.nr, dm.nt, r.bt0, 1,
                                              r.fac, S->scratch + q);
    KAD_HIP_CHECK(hipGetLastError());
  }
  KAD_HIP_CHECK(hipMemcpyAsync(out, S->scratch, nq * sizeof(double), hipMemcpyDeviceToHost, st));
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  return KAD_OK;
}

// The solvers need as many equations as unknowns (System_of_eqs::check_size_VS_unknowns,
// do_newton.cpp:113-118 and :600-609, which abort).
static int check_square(const SystemBase& sy) {
  if (sy.n_rows == sy.n_unknowns) return KAD_OK;
  fprintf(stderr, "Number of unknowns is different from the number of equations\nnbr equations = %d\n"
                  "nbr unknowns  = %d\n", sy.n_rows, sy.n_unknowns);
  return KAD_ERR_ARG;
}

// An interface condition that matches normal derivatives ("dn..." of the hand-traced
// system, or a string holding dn( or dr( ) rather than values.
static bool matches_derivative(const std::string& which) {
  return which.rfind("dn", 0) == 0 || which.find("dn(") != std::string::npos ||
         which.find("dr(") != std::string::npos;
}

// Blocks of the domain-decomposed solve: one per domain, with the domain's unknowns
// and the rows of the equations that live there.  An interface's matching condition
// lives in the inner domain when it matches values and in the outer one when it
// matches normal derivatives, which makes every block square (the domain's operator
// with a Dirichlet-type outer and a Neumann-type inner interface condition).  Integral
// equations and double unknowns couple every domain: such systems have no blocks.
static int dd_blocks(const SystemBase& sy, std::vector<int>& size, std::vector<int>& rows, std::vector<int>& cols) {
  if (sy.ndouble > 0 || sy.n_rows != sy.n_unknowns) return KAD_ERR_ARG;
  const int nb = sy.dmax - sy.dmin + 1;
  std::vector<std::vector<int>> R(nb), C(nb);
  for (auto& e : sy.eqs) {
    if (e.kind == 4) return KAD_ERR_ARG;
    const int d = (e.kind == 1 && matches_derivative(e.which)) ? e.d + 1 : e.d;
    if (d < sy.dmin || d > sy.dmax) return KAD_ERR_ARG;
    int len = 0;
    for (int c : e.nrows) len += c;
    for (int q = 0; q < len; ++q) R[d - sy.dmin].push_back(e.row0 + q);
  }
  for (auto& L : sy.layout)
    for (int q = 0; q < L[3]; ++q) C[L[1] - sy.dmin].push_back(L[2] + q);
  size.clear();
  rows.clear();
  cols.clear();
  for (int b = 0; b < nb; ++b) {
    if (R[b].size() != C[b].size() || R[b].empty()) return KAD_ERR_ARG;
    std::sort(R[b].begin(), R[b].end());
    size.push_back((int)R[b].size());
    rows.insert(rows.end(), R[b].begin(), R[b].end());
    cols.insert(cols.end(), C[b].begin(), C[b].end());
  }
  return KAD_OK;
}

// Set a parameter of the system: which indexes its parameters (XCTS: 0 = Ome, the
// horizon rotation of kerr.cpp:91-95).  The constant field is rewritten in every
// domain that holds it.
static int set_param_impl(kad_system_impl* S, int which, double value) {
  SystemBase& sy = *S->sys;
  if (which < 0 || which >= (int)sy.par_names.size()) return KAD_ERR_ARG;
  for (int d = sy.dmin; d <= sy.dmax; ++d) {
    const int slot = sy.param_slot(sy.par_names[which], d);
    if (slot < 0) continue;
    std::vector<double> h(sy.doms[d].conf_size(), value);
    S->vals_current = false;
    KAD_HIP_CHECK(hipMemcpyAsync(S->val_arena + (int64_t)slot * S->stride, h.data(), h.size() * sizeof(double),
                                 hipMemcpyHostToDevice, S->ctx->stream));
    KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  }
  if (S->xsys && sy.par_names[which] == "ome") S->xsys->prm.ome = value;
  return KAD_OK;
}

int kad_xcts_set_param(kad_system* s, int which, double value) {
  if (!s || which != 0) return KAD_ERR_ARG;
  return set_param_impl(reinterpret_cast<kad_system_impl*>(s), which, value);
}

// Sensitivities of the global quantities Q_q at a solution F(u, p) = 0 with respect
// to the parameter p = par, by the adjoint method:
//   J^T lambda_q = dQ_q/du   (the transposed solve reuses the LU factors of J),
//   dQ_q/dp = dQ_q/dp|_u - lambda_q . dF/dp   (the first term is 0: the quantities
//   read no parameter).
// dQ/du comes from tangent programs of the quantities, dF/dp from the tangent
// program of the parameter; J is refilled and factored at the current fields.
// out: HOST, one per quantity; dQdu (HOST, nq x n, may be NULL) receives dQ/du.
static int sensitivity_impl(kad_system_impl* S, int par, double* out, double* dQdu) {
  SystemBase& sy = *S->sys;
  hipStream_t st = S->ctx->stream;
  const int n = sy.n_unknowns;
  if (check_square(sy)) return KAD_ERR_ARG;
  const int nq = (int)sy.obs.size();
  if (nq == 0 || par < 0 || par >= (int)S->par_progs.size()) return KAD_ERR_ARG;
  int rc;
  if (!S->J) {
    if ((rc = S->dmalloc((void**)&S->J, (size_t)n * n * sizeof(double)))) return rc;
    if ((rc = S->dmalloc((void**)&S->ipiv, (size_t)n * sizeof(int)))) return rc;
    if ((rc = S->dmalloc((void**)&S->perm, (size_t)n * sizeof(int)))) return rc;
    if ((rc = S->dmalloc((void**)&S->X, (size_t)n * sizeof(double)))) return rc;
  }
  if (!S->dQ) {
    if ((rc = S->dmalloc((void**)&S->dQ, (size_t)nq * n * sizeof(double)))) return rc;
    if ((rc = S->dmalloc((void**)&S->dFdp, (size_t)n * sizeof(double)))) return rc;
  }
  if ((rc = S->jacobian(S->J, n))) return rc;  // includes the value pass at the current fields
  if ((rc = S->run(sy.obs_prog.code, 1, 0, -1, S->F, 0, 0))) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(S->dQ, 0, (size_t)nq * n * sizeof(double), st));
  KAD_HIP_CHECK(hipMemsetAsync(S->dFdp, 0, (size_t)n * sizeof(double), st));
  for (auto& p : S->obs_tprogs)
    for (int u0 = 0; u0 < p.cnt; u0 += S->B) {
      const int nb = std::min(S->B, p.cnt - u0);
      if ((rc = S->run(p.code, nb, u0, p.seed_tab, S->dQ, (int64_t)p.off + u0, n))) return rc;
    }
  const DevProgram& pp = S->par_progs[par];
  if ((rc = S->run(pp.code, 1, 0, pp.seed_tab, S->dFdp, 0, n))) return rc;
  if (dQdu) {
    KAD_HIP_CHECK(hipMemcpyAsync(dQdu, S->dQ, (size_t)nq * n * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  rc = kad_lu_factor(reinterpret_cast<kad_ctx*>(S->ctx), KAD_PREC_F64, n, S->J, n, S->ipiv, S->perm, 0);
  if (rc < 0) return rc;
  if (rc > 0) return KAD_ERR_SINGULAR;
  for (int q = 0; q < nq; ++q) {
    KAD_HIP_CHECK(hipMemcpyAsync(S->X, S->dQ + (size_t)q * n, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                                 st));
    if ((rc = kad_lu_solve(reinterpret_cast<kad_ctx*>(S->ctx), KAD_PREC_F64, n, S->J, n, S->perm, S->X, 1)))
      return rc;
    out[q] = -S->ctx->dot(S->X, S->dFdp, n);
  }
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  return KAD_OK;
}

// dst[q n + k] = src[q] for every quantity q (a double unknown's column of dQ/du)
__global__ void scatter_q_kernel(const double* __restrict__ src, int nq, double* __restrict__ dst, int64_t n, int k) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nq) dst[(int64_t)q * n + k] = src[q];
}

// Global quantities of a parsed system at the current fields: out[q] (HOST).
static int globals_impl(kad_system_impl* S, double* out) {
  SystemBase& sy = *S->sys;
  const int nq = (int)sy.glob_names.size();
  if (nq == 0 || nq > 32) return KAD_ERR_ARG;
  int rc = S->ensure_values();
  if (rc) return rc;
  if ((rc = S->run(sy.glob_prog.code, 1, 0, -1, S->scratch, 0, 0))) return rc;
  KAD_HIP_CHECK(hipMemcpyAsync(out, S->scratch, nq * sizeof(double), hipMemcpyDeviceToHost, S->ctx->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

// Adjoint sensitivities of the parsed system's global quantities:
// dQ_q/dp = dQ_q/dp|_u - lambda_q . dF/dp with J^T lambda_q = dQ_q/du, J refilled and
// factored at the current fields.  out[q] HOST; dQdu (HOST nq x n, may be NULL).
static int glob_sensitivity_impl(kad_system_impl* S, int par, double* out, double* dQdu) {
  SystemBase& sy = *S->sys;
  hipStream_t st = S->ctx->stream;
  const int n = sy.n_unknowns;
  if (check_square(sy)) return KAD_ERR_ARG;
  const int nq = (int)sy.glob_names.size();
  if (nq == 0 || nq > 16 || par < 0 || par >= (int)S->par_progs.size()) return KAD_ERR_ARG;
  int rc;
  if (!S->J) {
    if ((rc = S->dmalloc((void**)&S->J, (size_t)n * n * sizeof(double)))) return rc;
    if ((rc = S->dmalloc((void**)&S->ipiv, (size_t)n * sizeof(int)))) return rc;
    if ((rc = S->dmalloc((void**)&S->perm, (size_t)n * sizeof(int)))) return rc;
    if ((rc = S->dmalloc((void**)&S->X, (size_t)n * sizeof(double)))) return rc;
  }
  if (!S->dQ) {
    if ((rc = S->dmalloc((void**)&S->dQ, (size_t)std::max(3, nq) * n * sizeof(double)))) return rc;
    if ((rc = S->dmalloc((void**)&S->dFdp, (size_t)n * sizeof(double)))) return rc;
  }
  if ((rc = S->jacobian(S->J, n))) return rc;  // value pass + every column, doubles included
  double* tmp = S->scratch + 32;   // nq: a double unknown's dQ column, then dQ/dp at fixed fields
  KAD_HIP_CHECK(hipMemsetAsync(S->dQ, 0, (size_t)nq * n * sizeof(double), st));
  KAD_HIP_CHECK(hipMemsetAsync(S->dFdp, 0, (size_t)n * sizeof(double), st));
  for (int k = 0; k < sy.ndouble; ++k) {
    const DevProgram& gp = S->glob_par_progs[sy.dvar_par[k]];
    KAD_HIP_CHECK(hipMemsetAsync(tmp, 0, nq * sizeof(double), st));
    if ((rc = S->run(gp.code, 1, 0, gp.seed_tab, tmp, 0, 1))) return rc;
    scatter_q_kernel<<<1, 64, 0, st>>>(tmp, nq, S->dQ, n, k);
  }
  for (auto& p : S->glob_tprogs)
    for (int u0 = 0; u0 < p.cnt; u0 += S->B) {
      const int nb = std::min(S->B, p.cnt - u0);
      if ((rc = S->run(p.code, nb, u0, p.seed_tab, S->dQ, (int64_t)p.off + u0, 1))) return rc;
    }
  const DevProgram& pp = S->par_progs[par];
  if ((rc = S->run(pp.code, 1, 0, pp.seed_tab, S->dFdp, 0, n))) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(tmp, 0, nq * sizeof(double), st));
  if ((rc = S->run(S->glob_par_progs[par].code, 1, 0, S->glob_par_progs[par].seed_tab, tmp, 0, 1))) return rc;
  double direct[16];
  KAD_HIP_CHECK(hipMemcpyAsync(direct, tmp, nq * sizeof(double), hipMemcpyDeviceToHost, st));
  if (dQdu) KAD_HIP_CHECK(hipMemcpyAsync(dQdu, S->dQ, (size_t)nq * n * sizeof(double), hipMemcpyDeviceToHost, st));
  rc = kad_lu_factor(reinterpret_cast<kad_ctx*>(S->ctx), KAD_PREC_F64, n, S->J, n, S->ipiv, S->perm, 0);
  if (rc < 0) return rc;
  if (rc > 0) return KAD_ERR_SINGULAR;
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  for (int q = 0; q < nq; ++q) {
    KAD_HIP_CHECK(hipMemcpyAsync(S->X, S->dQ + (size_t)q * n, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                                 st));
    if ((rc = kad_lu_solve(reinterpret_cast<kad_ctx*>(S->ctx), KAD_PREC_F64, n, S->J, n, S->perm, S->X, 1)))
      return rc;
    out[q] = direct[q] - S->ctx->dot(S->X, S->dFdp, n);
  }
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  return KAD_OK;
}

// XCTS: out[3] = d(M_ADM, M_Komar, J_ADM)/dOme.
int kad_xcts_sensitivity(kad_system* s, double* out, double* dQdu) {
  if (!s || !out) return KAD_ERR_ARG;
  return sensitivity_impl(reinterpret_cast<kad_system_impl*>(s), 0, out, dQdu);
}

// ||F||_inf at the current fields (the value pass leaves F on the device).
static int newton_residual(kad_system_impl* S, double* err) {
  hipStream_t st = S->ctx->stream;
  int rc;
  if ((rc = S->value_pass())) return rc;
  absmax_kernel<<<1, 256, 0, st>>>(S->F, S->sys->n_rows, S->scratch);
  KAD_HIP_CHECK(hipMemcpyAsync(err, S->scratch, sizeof(double), hipMemcpyDeviceToHost, st));
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  return KAD_OK;
}

// The Newton direction X (S->X) with J X = F: Jacobian fill, then the solver's
// factorisation and solve.  err = ||F||_inf (the forcing term of the low-precision
// refinement).  Gauss-Newton (solver 4) solves J^T J X = J^T F instead, as
// do_newton.cpp:264-311 (pdgemv_ 'T', pdgemm_ 'T','N', pdgesv_).  times as for
// kad_xcts_do_newton from index 1; t1 = wall clock at entry.
static int newton_direction(kad_system_impl* S, int solver, double err, double* times, double t1) {
  SystemBase& sy = *S->sys;
  hipStream_t st = S->ctx->stream;
  kad_ctx* cx = reinterpret_cast<kad_ctx*>(S->ctx);
  const int n = sy.n_unknowns;
  int rc;
  if ((rc = check_square(sy))) return rc;
  if (!S->J) {
    if ((rc = S->dmalloc((void**)&S->J, (size_t)n * n * sizeof(double)))) return rc;
    if ((rc = S->dmalloc((void**)&S->ipiv, (size_t)n * sizeof(int)))) return rc;
    if ((rc = S->dmalloc((void**)&S->perm, (size_t)n * sizeof(int)))) return rc;
    if ((rc = S->dmalloc((void**)&S->X, (size_t)n * sizeof(double)))) return rc;
  }
  if ((solver == 1 || solver == 2 || solver == 3) && !S->Jlo)
    if ((rc = S->dmalloc((void**)&S->Jlo, (size_t)n * n * sizeof(float)))) return rc;
  if (solver == 5 && !S->Jfac)
    if ((rc = S->dmalloc((void**)&S->Jfac, (size_t)n * n * sizeof(double)))) return rc;
  if (solver == 4 && !S->Jt) {
    if ((rc = S->dmalloc((void**)&S->Jt, (size_t)n * n * sizeof(double)))) return rc;
    if ((rc = S->dmalloc((void**)&S->JtJ, (size_t)n * n * sizeof(double)))) return rc;
  }
  // Jacobian (the value pass of the residual already holds the intermediates)
  if (!S->vals_current)
    if ((rc = S->value_pass())) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(S->J, 0, (size_t)n * n * sizeof(double), st));
  if ((rc = S->double_columns(S->J, n, 0, sy.ndouble, 0))) return rc;
  if ((rc = S->fill_columns(S->J, n))) return rc;
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  double t2 = now_s();
  if (times) times[1] = t2 - t1;
  int info = 0;
  double tf0 = 0, tf1 = 0;
  // The refinement target of the iterative solvers.  The default runs it to its
  // attainable accuracy (the tau system's near-null modes, high radial orders in the
  // compactified domain, are barely seen by a residual test).  With a forcing factor c
  // the target follows ||F||: c min(1, ||F||) relative, which keeps Newton quadratic
  // while early iterations need only a few inner steps.
  const double eta = S->forcing > 0.0 ? std::max(1e-13, S->forcing * std::min(1.0, err)) : 1e-13;
  // A refinement result is a usable Newton direction below this relative residual.
  constexpr double kAcceptResid = 1e-6;
  // Static pivoting (solvers 6-7): the kept orders are rejected above this multiplier
  // (partial pivoting keeps every one at most 1), and their factors are replaced when
  // the refinement needs more inner steps than this (2 with the exact Woodbury solve).
  // The refinement on the fp64 J is the real check: on the 17x13x12 spinning hole the
  // orders of one Newton step give the next step's blocks multipliers up to 4e5 (in
  // columns of rounding-level entries) and the preconditioned refinement still reaches
  // 2e-15 in two inner steps, as with partial pivoting; the bound only turns away
  // factors past any use (and the infinities of an exactly zero pivot).
  constexpr double kDdMaxGrowth = 1e8;
  constexpr int kDdStaticInner = 6;
  // Orders kept over several steps drift from the Jacobian (the same hole: multipliers
  // 4e5, 6e9, 4e14 one, two, three steps on, and then 700 inner steps), so a step whose
  // static factors show multipliers above this takes them, and the next step factors
  // with partial pivoting again for fresh orders.
  constexpr double kDdRefresh = 1e3;
  if (solver == 6 || solver == 7) {
    // domain-decomposed factors (the blocks factored together, then the Woodbury
    // coupling) precondition flexible GMRES on the current J; solver 7 keeps them for
    // later Jacobians while the refinement takes at most `limit` inner steps
    S->lo_valid = false;
    S->fac_valid = false;
    if (!S->dd) {
      std::vector<int> bs, rows, cols;
      if (dd_blocks(sy, bs, rows, cols)) {
        fprintf(stderr, "domain-decomposed solve: the system's domain blocks are not square\n");
        return KAD_ERR_ARG;
      }
      if ((rc = kad_dd_create(cx, n, (int)bs.size(), bs.data(), rows.data(), cols.data(), &S->dd))) return rc;
    }
    const int limit = S->reuse_inner > 0 ? S->reuse_inner : 30;
    bool solved = false;
    if (solver == 7 && S->dd_valid && S->prev_inner <= limit) {
      rc = kad_dd_solve(S->dd, S->J, n, S->F, S->X, eta, 8, S->last_iters, &S->last_resid);
      if (rc < 0 && rc != KAD_ERR_NOCONV) return rc;
      S->prev_inner = S->last_iters[1];
      solved = rc == KAD_OK || S->last_resid <= std::max(eta, kAcceptResid);
      S->last_factored = 0;
    }
    if (!solved) {
      KAD_HIP_CHECK(hipStreamSynchronize(st));
      const double tg0 = now_s();
      // static pivoting first when the solve has pivot orders; partial pivoting when
      // they do not apply (a multiplier above kDdMaxGrowth, other coupling rows)
      int dd_info = 1;
      bool stat = false;
      if (S->dd_static && S->dd_orders) {
        if ((rc = kad_dd_factor_static(S->dd, S->J, n, kDdMaxGrowth)) < 0) return rc;
        dd_info = rc;
        stat = rc == 0;
        double growth = 0.0;
        if (stat && kad_dd_pivoting(S->dd, nullptr, &growth) == KAD_OK && growth > kDdRefresh)
          S->dd_orders = false;  // used once; the next step refreshes them
      }
      // keep: with static pivoting on, the capacitance matrix keeps its order through
      // the blocks' partial pivoting while its multipliers stay within kDdMaxGrowth
      auto factor_pivoting = [&](bool keep) -> int {
        const int frc = keep ? kad_dd_factor_keep(S->dd, S->J, n, kDdMaxGrowth) : kad_dd_factor(S->dd, S->J, n);
        S->dd_orders = frc == 0;
        return frc;
      };
      bool kept_cap = false;
      if (!stat) {
        if ((dd_info = factor_pivoting(S->dd_static)) < 0) return dd_info;
        double cg = 0.0;
        kept_cap = dd_info == 0 && S->dd_static && kad_dd_pivoting(S->dd, nullptr, &cg) == KAD_OK && cg > 0.0;
      }
      tf1 += now_s() - tg0;
      // dd_info > 0: a singular domain block or capacitance matrix
      S->dd_valid = (dd_info == 0);
      S->last_factored = 1;
      bool fallback = dd_info != 0;
      if (dd_info == 0) {
        rc = kad::dd_solve_after_factor(S->dd, S->J, n, S->F, S->X, eta, 8, S->last_iters, &S->last_resid);
        if (rc < 0 && rc != KAD_ERR_NOCONV) return rc;
        bool poor = rc == KAD_ERR_NOCONV && !(S->last_resid <= std::max(eta, kAcceptResid));
        if ((stat || kept_cap) && (poor || S->last_iters[1] > kDdStaticInner)) {
          // the kept orders made a poor preconditioner: partial pivoting throughout,
          // solved again
          KAD_HIP_CHECK(hipStreamSynchronize(st));
          const double tg2 = now_s();
          if ((dd_info = factor_pivoting(false)) < 0) return dd_info;
          tf1 += now_s() - tg2;
          S->dd_valid = (dd_info == 0);
          poor = dd_info != 0;
          if (dd_info == 0) {
            rc = kad::dd_solve_after_factor(S->dd, S->J, n, S->F, S->X, eta, 8, S->last_iters, &S->last_resid);
            if (rc < 0 && rc != KAD_ERR_NOCONV) return rc;
            poor = rc == KAD_ERR_NOCONV && !(S->last_resid <= std::max(eta, kAcceptResid));
          }
        }
        S->prev_inner = S->last_iters[1];
        fallback = poor;
      }
      if (fallback) {
        // a singular block (J itself need not be singular) or not a usable Newton
        // direction: the fp64 LU of the same, untouched J decides
        S->last_fallback = 1;
        const double tg1 = now_s();
        if ((rc = kad_lu_factor(cx, KAD_PREC_F64, n, S->J, n, S->ipiv, S->perm, 0)) < 0) return rc;
        tf1 += now_s() - tg1;
        info = rc;
        S->dd_valid = false;
        if (info == 0) {
          KAD_HIP_CHECK(hipMemcpyAsync(S->X, S->F, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
          if ((rc = kad_lu_solve(cx, KAD_PREC_F64, n, S->J, n, S->perm, S->X, 0))) return rc;
        }
      }
    }
  } else if (solver == 5) {
    // fp64 factors of an earlier J precondition flexible GMRES on the current J while
    // the previous refinement took at most `limit` inner steps; otherwise the current J
    // is factored in place and solved directly, and its factors are kept (the next fill
    // goes to the other buffer)
    S->lo_valid = false;
    const int limit = S->reuse_inner > 0 ? S->reuse_inner : 30;
    bool solved = false;
    if (S->fac_valid && S->prev_inner <= limit) {
      rc = kad_ir_solve_scaled(cx, n, S->J, n, KAD_PREC_F64, S->Jfac, n, S->perm, nullptr, nullptr, S->F, S->X,
                               eta, 8, 1, 0, S->last_iters, &S->last_resid);
      if (rc < 0 && rc != KAD_ERR_NOCONV) return rc;
      S->prev_inner = S->last_iters[1];
      solved = rc == KAD_OK || S->last_resid <= std::max(eta, kAcceptResid);
      S->last_factored = 0;
    }
    if (!solved) {
      KAD_HIP_CHECK(hipStreamSynchronize(st));
      const double tg0 = now_s();
      if ((rc = kad_lu_factor(cx, KAD_PREC_F64, n, S->J, n, S->ipiv, S->perm, 0)) < 0) return rc;
      tf1 += now_s() - tg0;
      info = rc;
      if (info == 0) {
        KAD_HIP_CHECK(hipMemcpyAsync(S->X, S->F, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
        if ((rc = kad_lu_solve(cx, KAD_PREC_F64, n, S->J, n, S->perm, S->X, 0))) return rc;
      }
      std::swap(S->J, S->Jfac);
      S->fac_valid = (info == 0);
      S->prev_inner = 0;
      S->last_factored = 1;
    }
  } else if (solver == 0 || solver == 4) {
    S->lo_valid = false;  // the fp64 factorisation takes over ipiv / perm
    S->fac_valid = false;
    S->dd_valid = false;
    S->last_factored = 1;
    double* M = S->J;
    if (solver == 4) {
      // J^T F, then J^T J = (J^T) J with an explicit transpose and the MFMA GEMM; J
      // itself stays unfactored
      if ((rc = kad_gemv64(cx, n, S->J, n, S->F, nullptr, S->X, 1))) return rc;
      dim3 g((unsigned)kad::cdiv(n, 64), (unsigned)kad::cdiv(n, 64));
      transpose_kernel<<<g, 256, 0, st>>>(S->J, n, S->Jt, n, n);
      KAD_HIP_CHECK(hipGetLastError());
      if ((rc = kad_gemm(cx, KAD_PREC_F64, 1, n, n, n, S->Jt, n, S->J, n, S->JtJ, n))) return rc;
      M = S->JtJ;
    }
    KAD_HIP_CHECK(hipStreamSynchronize(st));
    tf0 = now_s();
    if ((rc = kad_lu_factor(cx, KAD_PREC_F64, n, M, n, S->ipiv, S->perm, 0)) < 0) return rc;
    tf1 = now_s();
    info = rc;
    if (info == 0) {
      if (solver == 0)
        KAD_HIP_CHECK(hipMemcpyAsync(S->X, S->F, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
      if ((rc = kad_lu_solve(cx, KAD_PREC_F64, n, M, n, S->perm, S->X, 0))) return rc;
    }
  } else {
    S->fac_valid = false;  // the low-precision factorisation takes over ipiv / perm
    // low-precision factors of the equilibrated R J C (exact powers of two); the
    // refinement keeps the unscaled fp64 J, so its stopping test is on ||F - J X||.
    if (!S->Rs) {
      if ((rc = S->dmalloc((void**)&S->Rs, (size_t)n * sizeof(double)))) return rc;
      if ((rc = S->dmalloc((void**)&S->Cs, (size_t)n * sizeof(double)))) return rc;
    }
    const int prec = (solver == 3) ? KAD_PREC_BF16 : KAD_PREC_F32;
    // The factors of an earlier J precondition the refinement on the current one as
    // long as it keeps converging quickly; J itself (the residuals) is always current.
    bool reuse = S->reuse_inner > 0 && S->lo_valid && S->lo_prec == prec && S->prev_inner <= S->reuse_inner;
    for (int attempt = 0; attempt < 2; ++attempt) {
      if (!reuse) {
        if ((rc = kad_cast_equilibrated(cx, n, S->J, n, S->Jlo, n, S->Rs, S->Cs))) return rc;
        KAD_HIP_CHECK(hipStreamSynchronize(st));
        const double tg0 = now_s();
        if ((rc = kad_lu_factor(cx, prec, n, S->Jlo, n, S->ipiv, S->perm, 0)) < 0) return rc;
        tf1 += now_s() - tg0;
        info = rc;
        S->lo_valid = (info == 0);
        S->lo_prec = prec;
        S->last_factored = 1;
        if (info != 0) break;
      }
      rc = kad_ir_solve_scaled(cx, n, S->J, n, prec, S->Jlo, n, S->perm, S->Rs, S->Cs, S->F, S->X, eta,
                               solver == 1 ? 30 : 8, solver == 1 ? 0 : 1, 0, S->last_iters, &S->last_resid);
      if (rc < 0 && rc != KAD_ERR_NOCONV) return rc;
      S->prev_inner = S->last_iters[1];
      // reused factors that no longer precondition well: factor the current J and retry
      if (reuse && rc == KAD_ERR_NOCONV && !(S->last_resid <= std::max(eta, 1e-6))) {
        reuse = false;
        continue;
      }
      break;
    }
    if (info == 0) {
      // The refinement stops at its attainable accuracy.  A step whose relative linear
      // residual is above kAcceptResid is not a usable Newton direction (a refinement
      // that made no progress leaves X = 0): the step is then solved with an fp64
      // factorisation of the same (unscaled, untouched) J, as the reference's pdgesv_
      // would, and the fallback is reported through kad_xcts_last_solve.
      if (rc == KAD_ERR_NOCONV && !(S->last_resid <= std::max(eta, kAcceptResid))) {
        S->last_fallback = 1;
        S->lo_valid = false;  // the fp64 factorisation below takes over ipiv / perm
        const double tg0 = now_s();
        rc = kad_lu_factor(cx, KAD_PREC_F64, n, S->J, n, S->ipiv, S->perm, 0);
        if (rc < 0) return rc;
        tf1 += now_s() - tg0;  // both factorisations count as factor time
        info = rc;
        if (info == 0) {
          KAD_HIP_CHECK(hipMemcpyAsync(S->X, S->F, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
          if ((rc = kad_lu_solve(cx, KAD_PREC_F64, n, S->J, n, S->perm, S->X, 0))) return rc;
        }
      }
    }
  }
  if (times) times[4] = (solver == 0 || solver == 4) ? tf1 - tf0 : tf1;  // the others accumulate in tf1
  if (info != 0) {
    fprintf(stderr, "Jacobian is singular! Stopping.\n");
    return KAD_ERR_SINGULAR;
  }
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  if (times) times[2] = now_s() - t2;
  return KAD_OK;
}

static void reset_solve_record(kad_system_impl* S) {
  S->last_iters[0] = S->last_iters[1] = 0;
  S->last_resid = 0.0;
  S->last_fallback = 0;
  S->last_factored = 0;
}

// One Newton-Raphson iteration (System_of_eqs::do_newton, src_par/do_newton.cpp:65-430).
// solver: 0 fp64 LU; 1 fp32 LU + fp64 IR; 2 fp32 LU + fp64 GMRES-IR; 3 bf16-update LU +
// fp64 GMRES-IR; 4 Gauss-Newton (fp64 J^T J X = J^T F, the reference's GAUSS_NEWTON);
// 5 fp64 LU of an earlier J (refactored when it stops preconditioning well) + fp64
// flexible GMRES on the current J; 6 domain-decomposed fp64 factors (kad_dd_factor) +
// fp64 flexible GMRES; 7 the same, the factors of an earlier J kept while they
// precondition well.
// Returns 1 when ||F||_inf < precision at entry (converged; no step taken), 0 after a step,
// < 0 on error, KAD_ERR_SINGULAR (-5) when the factorisation meets an exact zero pivot.
// times[5] (may be NULL): residual, Jacobian fill, factor+solve, update, factor alone (seconds).
int kad_xcts_do_newton(kad_system* s, double precision, double* error, int solver, double* times) {
  if (!s || !error || solver < 0 || solver > 7) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  hipStream_t st = S->ctx->stream;
  int rc;
  double t0 = now_s();
  double err = 0;
  if ((rc = newton_residual(S, &err))) return rc;
  *error = err;
  double t1 = now_s();
  if (times) times[0] = t1 - t0;
  reset_solve_record(S);
  if (err < precision) return 1;  // false for a NaN error: a NaN residual is never converged
  if ((rc = newton_direction(S, solver, err, times, t1))) return rc;
  double t3 = now_s();
  if ((rc = S->update(S->X, -1.0))) return rc;
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  if (times) times[3] = now_s() - t3;
  return 0;
}

// Newton iteration with a backtracking line search
// (System_of_eqs::do_newton_with_linesearch, src_par/do_newton.cpp:454-517, after
// Numerical Recipes 2007 section 9.7.1).  f = sum F^2 (compute_f, :594-599), slope =
// -2 f_old; the Newton step p (J p = F, by `solver`) is scaled by stepmax / max(p) when
// its (signed) maximum exceeds stepmax; u(lambda) = u_old - lambda p.  lambda starts at
// 1 and is taken from the quadratic model, then the cubic one, kept in [0.1, 0.5] of the
// previous lambda, until f <= 5e-13, f <= f_old + 1e-4 lambda slope, lambda <= 1e-10, or
// ntrymax tries (the last computed lambda is then applied).
// Returns 1 when ||F||_inf < precision at entry, 0 after a step; KAD_ERR_NOCONV where the
// reference aborts on roundoff ("roundoff error in do_newton_with_linesearch": a
// non-negative slope or a negative cubic discriminant).  lambda_out, ntry_out (may be
// NULL): the lambda applied and the number of residual evaluations of the search.
int kad_system_do_newton_linesearch(kad_system* s, double precision, double* error, int ntrymax, double stepmax,
                                    int solver, double* lambda_out, int* ntry_out) {
  if (!s || !error || ntrymax < 1 || !(stepmax > 0.0) || solver < 0 || solver > 7) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  hipStream_t st = S->ctx->stream;
  int rc;
  double err = 0;
  if ((rc = newton_residual(S, &err))) return rc;
  *error = err;
  reset_solve_record(S);
  if (lambda_out) *lambda_out = 0.0;
  if (ntry_out) *ntry_out = 0;
  if (err < precision) return 1;
  auto sumsq_max = [&](const double* x, int n, double out[2]) -> int {
    sumsq_max_kernel<<<1, 256, 0, st>>>(x, n, S->scratch + 8);
    KAD_HIP_CHECK(hipMemcpyAsync(out, S->scratch + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    KAD_HIP_CHECK(hipStreamSynchronize(st));
    return KAD_OK;
  };
  double r[2];
  if ((rc = sumsq_max(S->F, sy.n_rows, r))) return rc;
  const double fold = r[0], slope = -2.0 * fold;
  if (!(slope < 0.0)) {
    fprintf(stderr, "roundoff error in do_newton_with_linesearch\n");
    return KAD_ERR_NOCONV;
  }
  if ((rc = newton_direction(S, solver, err, nullptr, now_s()))) return rc;
  if ((rc = sumsq_max(S->X, sy.n_unknowns, r))) return rc;
  const double pmax = r[1];
  const double pscale = (pmax > stepmax) ? stepmax / pmax : 1.0;
  // snapshot of the unknowns' storage: field coefficients and double-unknown constants
  std::vector<std::pair<double*, int>> held;
  for (auto& L : sy.layout)
    held.push_back({S->val_arena + (int64_t)sy.field_slot[L[0]][L[1]] * S->stride, sy.doms[L[1]].coef_size()});
  for (int k = 0; k < sy.ndouble; ++k)
    for (int d = sy.dmin; d <= sy.dmax; ++d) {
      const int slot = sy.param_slot(sy.par_names[sy.dvar_par[k]], d);
      if (slot >= 0) held.push_back({S->val_arena + (int64_t)slot * S->stride, sy.doms[d].conf_size()});
    }
  size_t total = 0;
  for (auto& h : held) total += (size_t)h.second;
  if (S->held_cap < total) {
    if ((rc = S->dmalloc((void**)&S->held, total * sizeof(double)))) return rc;
    S->held_cap = total;
  }
  auto copy_held = [&](bool save) -> int {
    size_t o = 0;
    for (auto& h : held) {
      double* a = S->held + o;
      KAD_HIP_CHECK(hipMemcpyAsync(save ? a : h.first, save ? h.first : a, (size_t)h.second * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
      o += (size_t)h.second;
    }
    return KAD_OK;
  };
  auto apply = [&](double lam) -> int {
    int q = copy_held(false);
    if (q) return q;
    return S->update(S->X, -lam * pscale);
  };
  if ((rc = copy_held(true))) return rc;
  const double alpha = 1.0e-4, lambdamin = 1.0e-10;
  double lambda = 1.0, lambda2 = 0.0, f2 = 0.0;
  int tries = 0;
  for (int itry = 1; itry <= ntrymax; ++itry) {
    if ((rc = apply(lambda))) return rc;
    if ((rc = S->value_pass())) return rc;
    if ((rc = sumsq_max(S->F, sy.n_rows, r))) return rc;
    const double f = r[0];
    tries = itry;
    if (f <= 5.0e-13) break;
    if (f <= fold + alpha * lambda * slope || lambda <= lambdamin) break;
    double lt;
    if (lambda == 1.0) {
      lt = -0.5 * slope / (f - fold - slope);
    } else {
      const double rhs1 = f - fold - lambda * slope, rhs2 = f2 - fold - lambda2 * slope;
      const double a = (rhs1 / (lambda * lambda) - rhs2 / (lambda2 * lambda2)) / (lambda - lambda2);
      const double b = (-lambda2 * rhs1 / (lambda * lambda) + lambda * rhs2 / (lambda2 * lambda2)) / (lambda - lambda2);
      if (a == 0.0) {
        lt = -0.5 * slope / b;
      } else {
        const double delta = b * b - 3.0 * a * slope;
        if (delta < 0.0) {
          fprintf(stderr, "roundoff error in do_newton_with_linesearch\n");
          return KAD_ERR_NOCONV;
        }
        lt = (-b + std::sqrt(delta)) / (3.0 * a);
      }
      if (lt > 0.5 * lambda) lt = 0.5 * lambda;
    }
    lambda2 = lambda;
    f2 = f;
    lambda = std::max(lt, 0.1 * lambda);
    if (itry == ntrymax)
      if ((rc = apply(lambda))) return rc;
  }
  S->vals_current = false;
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  if (lambda_out) *lambda_out = lambda;
  if (ntry_out) *ntry_out = tries;
  return 0;
}

// Seeded initial perturbation amp * 2^-(i+j+k) * U(-1,1) on every admissible
// coefficient, in unknown order (HOST out, n_unknowns).  U from splitmix64.
int kad_xcts_perturbation(kad_system* s, uint64_t seed, double amp, double* out) {
  if (!s || !out) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  uint64_t x = seed;
  auto next = [&]() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  // draws in the reference unknown order (variables, then domains, then components),
  // stored at this system's layout offsets, so both orderings start from the same guess;
  // the double unknowns are not perturbed
  for (int k = 0; k < sy.ndouble; ++k) out[k] = 0.0;
  for (auto& var : sy.var_fields)
    for (int d = sy.dmin; d <= sy.dmax; ++d)
      for (int f : var) {
        int u = -1;
        for (auto& L : sy.layout)
          if (L[0] == f && L[1] == d) u = L[2];
        if (u < 0) return KAD_ERR_ARG;
        const Domain& dm = sy.doms[d];
        for (auto& kj : dm.unknown_slots(sy.field_base(f, d).bt))
          for (int i = 0; i < dm.nr; ++i)
            out[u++] = amp * std::ldexp(1.0, -(i + kj.second + kj.first)) * (2 * next() - 1);
      }
  return KAD_OK;
}

int kad_xcts_last_solve(kad_system* s, int* iters, double* resid) {
  if (!s) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (iters) {
    iters[0] = S->last_iters[0];
    iters[1] = S->last_iters[1];
    iters[2] = S->last_fallback;
    iters[3] = S->last_factored;
  }
  if (resid) *resid = S->last_resid;
  return KAD_OK;
}

}  // extern "C"

extern "C" {
// Host-only tracing (no device work): sizes of the system and of its programs.
// info[0] n_unknowns, [1] n_rows, [2] value instructions, [3] tangent instructions,
// [4] value slots, [5] tangent programs.  layout: 4 ints per (field, domain) entry.
// kad_system_dd_blocks of the hand-traced system, traced on the host only.
int kad_xcts_host_dd_blocks(int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax, int order,
                            int* sizes, int cap, int* rows, int* cols) {
  if (!bounds || !sizes || order < 0 || order > 1) return KAD_ERR_ARG;
  XctsParams p;
  p.nr = nr;
  p.nt = nt;
  p.np = np;
  p.bounds.assign(bounds, bounds + nbounds);
  p.dmin = dmin;
  p.dmax = dmax;
  p.order = order;
  try {
    XctsSystem sy(p);
    sy.build();
    std::vector<int> bs, rr, cc;
    const int rc = dd_blocks(sy, bs, rr, cc);
    if (rc) return rc;
    for (int b = 0; b < (int)bs.size() && b < cap; ++b) sizes[b] = bs[b];
    if (rows) std::copy(rr.begin(), rr.end(), rows);
    if (cols) std::copy(cc.begin(), cc.end(), cols);
    return (int)bs.size();
  } catch (const std::exception& e) {
    fprintf(stderr, "kad_xcts_host_dd_blocks: %s\n", e.what());
    return KAD_ERR_ARG;
  }
}

int kad_xcts_host_info(int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax, double n0,
                       double a, double ome, int order, int* info, int* layout, int cap) {
  if (!bounds || !info || order < 0 || order > 1) return KAD_ERR_ARG;
  XctsParams p;
  p.nr = nr;
  p.nt = nt;
  p.np = np;
  p.bounds.assign(bounds, bounds + nbounds);
  p.dmin = dmin;
  p.dmax = dmax;
  p.n0 = n0;
  p.a = a;
  p.ome = ome;
  p.order = order;
  try {
    XctsSystem sy(p);
    sy.build();
    size_t tcode = 0;
    for (auto& pr : sy.tprog) tcode += pr.code.size();
    info[0] = sy.n_unknowns;
    info[1] = sy.n_rows;
    info[2] = (int)sy.tr.val.code.size();
    info[3] = (int)tcode;
    info[4] = sy.tr.nval_slots;
    info[5] = (int)sy.tprog.size();
    if (layout) {
      if (cap < 4 * (int)sy.layout.size()) return KAD_ERR_ARG;
      for (size_t q = 0; q < sy.layout.size(); ++q)
        for (int c = 0; c < 4; ++c) layout[4 * q + c] = sy.layout[q][c];
    }
    return (int)sy.layout.size();
  } catch (const std::exception& e) {
    fprintf(stderr, "[kadath_amd] kad_xcts_host_info: %s\n", e.what());
    return KAD_ERR_ARG;
  }
}
}

// ---------------------------------------------------------------------------
// Systems given as equation strings, and the entry points common to every system
// ---------------------------------------------------------------------------
struct kad_sysdef_impl {
  ParsedSystem* ps = nullptr;
  std::string err;
  ~kad_sysdef_impl() { delete ps; }
};

template <typename F>
static int guarded(kad_sysdef* def, F f) {
  if (!def) return KAD_ERR_ARG;
  auto* D = reinterpret_cast<kad_sysdef_impl*>(def);
  if (!D->ps) return KAD_ERR_ARG;
  try {
    f(*D->ps);
  } catch (const std::exception& e) {
    D->err = e.what();
    fprintf(stderr, "[kadath_amd] %s\n", e.what());
    return KAD_ERR_ARG;
  }
  return KAD_OK;
}

extern "C" {

int kad_sysdef_create(int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax,
                      kad_sysdef** out) {
  if (!bounds || nbounds < 1 || !out) return KAD_ERR_ARG;
  auto* D = new kad_sysdef_impl();
  try {
    D->ps = new ParsedSystem(nr, nt, np, std::vector<double>(bounds, bounds + nbounds), dmin, dmax, 0);
  } catch (const std::exception& e) {
    fprintf(stderr, "[kadath_amd] kad_sysdef_create: %s\n", e.what());
    delete D;
    return KAD_ERR_ARG;
  }
  *out = reinterpret_cast<kad_sysdef*>(D);
  return KAD_OK;
}

int kad_sysdef_destroy(kad_sysdef* def) {
  if (!def) return KAD_ERR_ARG;
  delete reinterpret_cast<kad_sysdef_impl*>(def);
  return KAD_OK;
}

int kad_sysdef_add_var(kad_sysdef* def, const char* name, int valence) {
  if (!name) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_var(name, valence); });
}
int kad_sysdef_add_cst(kad_sysdef* def, const char* name, double value) {
  if (!name) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_cst(name, value); });
}
int kad_sysdef_add_var_double(kad_sysdef* def, const char* name, double value) {
  if (!name) return KAD_ERR_ARG;
  int idx = -1;
  int rc = guarded(def, [&](ParsedSystem& p) { idx = p.add_var_double(name, value); });
  return rc ? rc : idx;
}
int kad_sysdef_add_global(kad_sysdef* def, const char* name, int dom, int bound, const char* expr) {
  if (!name || !expr) return KAD_ERR_ARG;
  int idx = -1;
  int rc = guarded(def, [&](ParsedSystem& p) { idx = p.add_global(name, dom, bound, expr); });
  return rc ? rc : idx;
}
int kad_sysdef_add_eq_int(kad_sysdef* def, int dom, int bound, const char* eq) {
  if (!eq) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_eq_int(dom, bound, eq); });
}
int kad_sysdef_add_param(kad_sysdef* def, const char* name, double value) {
  if (!name) return KAD_ERR_ARG;
  int idx = -1;
  int rc = guarded(def, [&](ParsedSystem& p) { idx = p.add_param(name, value); });
  return rc ? rc : idx;
}
int kad_sysdef_add_cst_field(kad_sysdef* def, const char* name, int valence, const double* conf, int64_t count) {
  if (!name || !conf || count <= 0) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_cst_field(name, valence, std::vector<double>(conf, conf + count)); });
}
int kad_sysdef_set_flat_metric(kad_sysdef* def, const char* name) {
  if (!name) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.set_flat_metric(name); });
}
int kad_sysdef_add_def(kad_sysdef* def, const char* text) {
  if (!text) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_def(text); });
}
int kad_sysdef_add_eq_inside(kad_sysdef* def, int dom, const char* eq, int order) {
  if (!eq) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_eq_inside(dom, eq, order); });
}
int kad_sysdef_add_eq_bc(kad_sysdef* def, int dom, int bound, const char* eq) {
  if (!eq) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_eq_bc(dom, bound, eq); });
}
int kad_sysdef_add_eq_matching(kad_sysdef* def, int dom, int bound, const char* expr) {
  if (!expr) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.add_eq_matching(dom, bound, expr); });
}
int kad_sysdef_space_add_eq(kad_sysdef* def, const char* eq, const char* m1, const char* m2) {
  if (!eq || !m1 || !m2) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) { p.space_add_eq(eq, m1, m2); });
}
int kad_sysdef_space_add_bc(kad_sysdef* def, int bound, const char* eq) {
  if (!eq) return KAD_ERR_ARG;
  return guarded(def, [&](ParsedSystem& p) {
    if (bound == INNER_BC) p.space_add_inner_bc(eq);
    else p.space_add_outer_bc(eq);
  });
}

// Collocation points of the conf grid of domain d (r fastest, then theta, then phi):
// r (compactified domain: infinity at x = 1 is reported as +inf), theta, phi.  HOST out.
// Mark domain d as a shell of an adapted space with round boundaries (the homothetic
// shells of Space_adapted_bh): its surface integrals weight in configuration space
// (Domain_shell_inner_adapted::integ, domain_shell_inner_adapted.cpp:1134-1208).
int kad_sysdef_set_adapted(kad_sysdef* def, int d) {
  return guarded(def, [&](ParsedSystem& p) {
    if (d < p.dmin || d > p.dmax) throw std::runtime_error("kad_sysdef_set_adapted: domain out of range");
    if (p.doms[d].compact) throw std::runtime_error("kad_sysdef_set_adapted: the compactified domain is not adapted");
    p.doms[d].adapted = true;
  });
}

int kad_sysdef_coloc(kad_sysdef* def, int d, double* r, double* theta, double* phi) {
  return guarded(def, [&](ParsedSystem& p) {
    if (d < p.dmin || d > p.dmax) throw std::runtime_error("kad_sysdef_coloc: domain out of range");
    const Domain& dm = p.doms[d];
    for (int k = 0; k < dm.np; ++k)
      for (int j = 0; j < dm.nt; ++j)
        for (int i = 0; i < dm.nr; ++i) {
          const size_t q = (size_t)i + dm.nr * (j + (size_t)dm.nt * k);
          if (r) r[q] = std::isfinite(dm.R[q]) ? dm.R[q] : HUGE_VAL;
          if (theta) theta[q] = dm.th[j];
          if (phi) phi[q] = dm.ph[k];
        }
  });
}

// Host-only trace (no device): sizes and row layout of the declared system.
int kad_sysdef_host_trace(kad_sysdef* def, int order, int* info, int* row_layout, int cap) {
  if (!info || (order != 0 && order != 1)) return KAD_ERR_ARG;
  int q = 0;
  int rc = guarded(def, [&](ParsedSystem& p) {
    p.order_mode = order;
    p.build();
    size_t tcode = 0;
    for (auto& pr : p.tprog) tcode += pr.code.size();
    info[0] = p.n_unknowns;
    info[1] = p.n_rows;
    info[2] = (int)p.tr.val.code.size();
    info[3] = (int)tcode;
    info[4] = p.tr.nval_slots;
    info[5] = (int)p.tprog.size();
    for (const auto& e : p.eqs) {
      int row = e.row0;
      for (int c = 0; c < e.ncomp; ++c) {
        if (row_layout) {
          if (5 * (q + 1) > cap) throw std::runtime_error("kad_sysdef_host_trace: row layout capacity");
          const int v[5] = {e.kind, e.d, c, row, e.nrows[c]};
          for (int t = 0; t < 5; ++t) row_layout[5 * q + t] = v[t];
        }
        row += e.nrows[c];
        ++q;
      }
    }
  });
  if (rc) return rc;
  // a traced declaration is not built again
  auto* D = reinterpret_cast<kad_sysdef_impl*>(def);
  delete D->ps;
  D->ps = nullptr;
  return q;
}

// Trace the declared system into device programs and upload them.  The declaration
// is consumed (def is freed, also on failure).
int kad_system_build(kad_ctx* ctx, kad_sysdef* def, int order, int batch, kad_system** out) {
  if (!def) return KAD_ERR_ARG;
  auto* D = reinterpret_cast<kad_sysdef_impl*>(def);
  if (!ctx || !out || batch <= 0 || (order != 0 && order != 1) || !D->ps) {
    delete D;
    return KAD_ERR_ARG;
  }
  auto* S = new kad_system_impl();
  S->ctx = reinterpret_cast<kad::Ctx*>(ctx);
  S->B = batch;
  S->sys = D->ps;
  D->ps = nullptr;
  delete D;
  try {
    static_cast<ParsedSystem*>(S->sys)->order_mode = order;
    S->sys->build();
  } catch (const std::exception& e) {
    fprintf(stderr, "[kadath_amd] kad_system_build: %s\n", e.what());
    delete S;
    return KAD_ERR_ARG;
  }
  int rc = S->setup();
  if (rc) {
    delete S;
    return rc;
  }
  *out = reinterpret_cast<kad_system*>(S);
  return KAD_OK;
}

int kad_system_destroy(kad_system* s) { return kad_xcts_destroy(s); }

// n_unknowns, n_rows, scalar field components
int kad_system_sizes(kad_system* s, int* n_unknowns, int* n_rows, int* nfields) {
  if (!s) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (n_unknowns) *n_unknowns = S->sys->n_unknowns;
  if (n_rows) *n_rows = S->sys->n_rows;
  if (nfields) *nfields = S->sys->nfields;
  return KAD_OK;
}

int kad_system_layout(kad_system* s, int* out, int cap) { return kad_xcts_layout(s, out, cap); }
int kad_system_row_layout(kad_system* s, int* out, int cap) { return kad_xcts_row_layout(s, out, cap); }

// Values of field component f in domain d at the collocation points (HOST in, conf
// order): the field's coefficients become their transform (Scalar::set_domain then
// std_base / coef, tensor.cpp:609-650).
int kad_system_set_field(kad_system* s, int f, int d, const double* conf) {
  if (!s || !conf) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  if (f < 0 || f >= sy.nfields || d < sy.dmin || d > sy.dmax) return KAD_ERR_ARG;
  const Domain& dm = sy.doms[d];
  std::vector<double> c(conf, conf + dm.conf_size());
  for (double& x : c)
    if (!std::isfinite(x)) return KAD_ERR_ARG;
  const std::vector<double> cf = host_coef(dm, sy.field_base(f, d), c);
  S->vals_current = false;
  KAD_HIP_CHECK(hipMemcpyAsync(S->val_arena + (int64_t)sy.field_slot[f][d] * S->stride, cf.data(),
                               cf.size() * sizeof(double), hipMemcpyHostToDevice, S->ctx->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

// Values of field component f in domain d at the collocation points (HOST out):
// the synthesis of its coefficients (Val_domain::coef_i).
int kad_system_get_field(kad_system* s, int f, int d, double* conf) {
  if (!s || !conf) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  if (f < 0 || f >= sy.nfields || d < sy.dmin || d > sy.dmax) return KAD_ERR_ARG;
  const Domain& dm = sy.doms[d];
  std::vector<double> cf(dm.coef_size());
  int rc = kad_xcts_get_coef(s, f, d, cf.data());
  if (rc) return rc;
  const Base b = sy.field_base(f, d);
  // coef_i: r, then theta, then phi (coef_i.cpp:96-104)
  Mat Mr = coef_i_matrix(CHEB, dm.nr);
  std::vector<double> t1(cf.size(), 0.0), t2(cf.size(), 0.0);
  for (int k = 0; k < dm.npc; ++k)
    for (int j = 0; j < dm.nt; ++j)
      for (int i = 0; i < dm.nr; ++i)
        for (int ii = 0; ii < dm.nr; ++ii)
          t1[i + dm.nr * (j + (size_t)dm.nt * k)] += Mr(i, ii) * cf[ii + dm.nr * (j + (size_t)dm.nt * k)];
  for (int k = 0; k < dm.npc; ++k) {
    Mat Mt = coef_i_matrix(b.bt[k], dm.nt);
    for (int j = 0; j < dm.nt; ++j)
      for (int jj = 0; jj < dm.nt; ++jj)
        for (int i = 0; i < dm.nr; ++i)
          t2[i + dm.nr * (j + (size_t)dm.nt * k)] += Mt(j, jj) * t1[i + dm.nr * (jj + (size_t)dm.nt * k)];
  }
  Mat Mp = coef_i_matrix(b.bp, dm.np);
  for (int k = 0; k < dm.np; ++k)
    for (int j = 0; j < dm.nt; ++j)
      for (int i = 0; i < dm.nr; ++i) {
        double v = 0.0;
        for (int kk = 0; kk < dm.npc; ++kk) v += Mp(k, kk) * t2[i + dm.nr * (j + (size_t)dm.nt * kk)];
        conf[i + dm.nr * (j + (size_t)dm.nt * k)] = v;
      }
  return KAD_OK;
}

int kad_system_get_coef(kad_system* s, int f, int d, double* out) { return kad_xcts_get_coef(s, f, d, out); }
int kad_system_set_coef(kad_system* s, int f, int d, const double* in) { return kad_xcts_set_coef(s, f, d, in); }
int kad_system_update(kad_system* s, const double* X) { return kad_xcts_update(s, X); }
int kad_system_perturbation(kad_system* s, uint64_t seed, double amp, double* out) {
  return kad_xcts_perturbation(s, seed, amp, out);
}
int kad_system_add_unknowns(kad_system* s, const double* X0) {
  if (!s || !X0) return KAD_ERR_ARG;
  return add_unknowns_host(reinterpret_cast<kad_system_impl*>(s), X0);
}
// Number of double unknowns (unknowns 0 .. count-1).
int kad_system_ndouble(kad_system* s) {
  if (!s) return KAD_ERR_ARG;
  return reinterpret_cast<kad_system_impl*>(s)->sys->ndouble;
}
// Current value of double unknown k (HOST out).
int kad_system_get_double(kad_system* s, int k, double* value) {
  if (!s || !value) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  SystemBase& sy = *S->sys;
  if (k < 0 || k >= sy.ndouble) return KAD_ERR_ARG;
  const std::string& name = sy.par_names[sy.dvar_par[k]];
  for (int d = sy.dmin; d <= sy.dmax; ++d) {
    const int slot = sy.param_slot(name, d);
    if (slot < 0) continue;
    KAD_HIP_CHECK(hipMemcpyAsync(value, S->val_arena + (int64_t)slot * S->stride, sizeof(double),
                                 hipMemcpyDeviceToHost, S->ctx->stream));
    KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
    return KAD_OK;
  }
  return KAD_ERR_ARG;  // the unknown appears in no equation
}
// Global quantities (kad_sysdef_add_global) at the current fields, HOST out.
int kad_system_globals(kad_system* s, double* out) {
  if (!s || !out) return KAD_ERR_ARG;
  return globals_impl(reinterpret_cast<kad_system_impl*>(s), out);
}
// dQ_q/dp of the global quantities with respect to parameter `which`, by the adjoint
// method at the current (converged) fields; HOST out[nq], dQdu (HOST nq x n) may be NULL.
int kad_system_sensitivity(kad_system* s, int which, double* out, double* dQdu) {
  if (!s || !out) return KAD_ERR_ARG;
  return glob_sensitivity_impl(reinterpret_cast<kad_system_impl*>(s), which, out, dQdu);
}
int kad_system_set_param(kad_system* s, int which, double value) {
  if (!s) return KAD_ERR_ARG;
  return set_param_impl(reinterpret_cast<kad_system_impl*>(s), which, value);
}
int kad_system_do_newton(kad_system* s, double precision, double* error, int solver, double* times) {
  return kad_xcts_do_newton(s, precision, error, solver, times);
}
int kad_system_last_solve(kad_system* s, int* iters, double* resid) { return kad_xcts_last_solve(s, iters, resid); }

// The domain blocks of solvers 6-7 (HOST out): sizes (cap entries), rows and cols (n
// each, may be NULL).  Returns the number of blocks, KAD_ERR_ARG when they are not
// square (integral equations, double unknowns, unpaired interface conditions).
int kad_system_dd_blocks(kad_system* s, int* sizes, int cap, int* rows, int* cols) {
  if (!s || !sizes) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  std::vector<int> bs, rr, cc;
  int rc = dd_blocks(*S->sys, bs, rr, cc);
  if (rc) return rc;
  for (int b = 0; b < (int)bs.size() && b < cap; ++b) sizes[b] = bs[b];
  if (rows) std::copy(rr.begin(), rr.end(), rows);
  if (cols) std::copy(cc.begin(), cc.end(), cols);
  return (int)bs.size();
}

// Sizes and set-up times of the last domain-decomposed factorisation (kad_dd_info).
int kad_system_dd_info(kad_system* s, int* sizes, double* times) {
  if (!s) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (!S->dd) return KAD_ERR_ARG;
  return kad_dd_info(S->dd, sizes, times);
}

// Pivoting of the last domain-decomposed factorisation (kad_dd_pivoting).
int kad_system_dd_pivoting(kad_system* s, int* static_orders, double* growth) {
  if (!s) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (!S->dd) return KAD_ERR_ARG;
  return kad_dd_pivoting(S->dd, static_orders, growth);
}

// F(u) (device, n_rows): System_of_eqs::sec_member.
int kad_residual(kad_system* s, double* F) { return kad_xcts_residual(s, F); }

// Per (equation, component) max |F| over its rows (HOST out, cap entries), in row-layout
// order: System_of_eqs::check_equations (solver.cpp:24-47).  Returns the entry count.
int kad_check_equations(kad_system* s, double* errors, int cap) {
  if (!s || !errors) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  int rc = S->value_pass();
  if (rc) return rc;
  std::vector<double> h(S->sys->n_rows);
  KAD_HIP_CHECK(hipMemcpyAsync(h.data(), S->F, h.size() * sizeof(double), hipMemcpyDeviceToHost, S->ctx->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  int q = 0;
  for (const auto& e : S->sys->eqs) {
    int row = e.row0;
    for (int c = 0; c < e.ncomp; ++c) {
      if (q >= cap) return KAD_ERR_ARG;
      double m = 0.0;
      for (int r = row; r < row + e.nrows[c]; ++r) {
        const double a = std::fabs(h[r]);
        if (a > m || a != a) m = a;
      }
      errors[q++] = m;
      row += e.nrows[c];
    }
  }
  return q;
}

// Columns [col_begin, col_begin + col_count) of J (every do_col_J of that range,
// solver.cpp:120-214) into the device tile J_tile (column-major, ld >= n_rows), fp64
// (dtype 0) or fp32 (dtype 1).  The tile's columns are cleared first; the values are
// those of the current fields (the value pass runs when they changed).
int kad_jvp_fill(kad_system* s, int col_begin, int col_count, void* J_tile, int64_t ld, int dtype) {
  if (!s || !J_tile || col_begin < 0 || col_count <= 0 || (dtype != 0 && dtype != 1)) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  const int n = S->sys->n_unknowns, m = S->sys->n_rows;
  if (ld < m || col_begin + (int64_t)col_count > n) return KAD_ERR_ARG;
  hipStream_t st = S->ctx->stream;
  int rc = S->ensure_values();
  if (rc) return rc;
  const size_t es = dtype == 0 ? sizeof(double) : sizeof(float);
  KAD_HIP_CHECK(hipMemset2DAsync(J_tile, (size_t)ld * es, 0, (size_t)m * es, (size_t)col_count, st));
  const int c1 = col_begin + col_count;
  if ((rc = S->double_columns(J_tile, ld, col_begin, c1, col_begin, dtype))) return rc;
  for (auto& p : S->tprogs) {
    const int lo = std::max(col_begin, p.off), hi = std::min(c1, p.off + p.cnt);
    for (int u = lo; u < hi; u += S->B) {
      const int nb = std::min(S->B, hi - u);
      if ((rc = S->run(p.code, nb, u - p.off, p.seed_tab, J_tile, (int64_t)u - col_begin, ld, dtype))) return rc;
    }
  }
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  return KAD_OK;
}

// out = J^T w (device, n_unknowns; w device, n_rows), matrix-free: every tangent
// program runs as for the fill, and its row exports are replaced by their transposes
// (per column, the dot product of w with the exported rows), so J is never formed.
int kad_vjp(kad_system* s, const double* w, double* out) {
  if (!s || !w || !out) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  hipStream_t st = S->ctx->stream;
  int rc = S->ensure_values();
  if (rc) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(out, 0, (size_t)S->sys->n_unknowns * sizeof(double), st));
  for (int k = 0; k < S->sys->ndouble; ++k) {
    const DevProgram& pp = S->par_progs[S->sys->dvar_par[k]];
    if ((rc = S->run(pp.code, 1, 0, pp.seed_tab, out, k, 0, 0, w))) return rc;
  }
  for (auto& p : S->tprogs)
    for (int u0 = 0; u0 < p.cnt; u0 += S->B) {
      const int nb = std::min(S->B, p.cnt - u0);
      if ((rc = S->run(p.code, nb, u0, p.seed_tab, out, (int64_t)p.off + u0, 0, 0, w))) return rc;
    }
  KAD_HIP_CHECK(hipStreamSynchronize(st));
  return KAD_OK;
}

// dF/dp of parameter `which` at the current fields (device out, n_rows).
int kad_param_tangent(kad_system* s, int which, double* out) {
  if (!s || !out) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (which < 0 || which >= (int)S->par_progs.size()) return KAD_ERR_ARG;
  int rc = S->ensure_values();
  if (rc) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(out, 0, (size_t)S->sys->n_rows * sizeof(double), S->ctx->stream));
  const DevProgram& pp = S->par_progs[which];
  if ((rc = S->run(pp.code, 1, 0, pp.seed_tab, out, 0, S->sys->n_rows))) return rc;
  KAD_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
  return KAD_OK;
}

}  // extern "C"

extern "C" {
// Options of the low-precision solvers: which 0 = forcing factor c (refinement target
// max(1e-13, c min(1, ||F||_inf)); 0, the default, runs it to its attainable
// accuracy), which 1 = reuse the low-precision factors while the previous refinement
// took at most `value` inner steps (0, the default: factor every step).
int kad_system_set_option(kad_system* s, int which, double value) {
  if (!s || !(value >= 0.0)) return KAD_ERR_ARG;
  auto* S = reinterpret_cast<kad_system_impl*>(s);
  if (which == 0) {
    S->forcing = value;
  } else if (which == 1) {
    S->reuse_inner = (int)value;
  } else if (which == 2) {
    // forget every factorisation kept for reuse, and the kept pivot orders: the next
    // step factors its own J with partial pivoting
    S->lo_valid = S->fac_valid = S->dd_valid = S->dd_orders = false;
    S->prev_inner = 0;
  } else if (which == 5) {
    // element-wise instructions applied in the spectral pass that feeds them (1, the
    // default) or launched on their own (0)
    if (value != 0.0 && value != 1.0) return KAD_ERR_ARG;
    S->fuse_ew = value != 0.0;
  } else if (which == 4) {
    // static pivoting in solvers 6-7 (1, the default) or partial pivoting at every
    // factorisation (0)
    if (value != 0.0 && value != 1.0) return KAD_ERR_ARG;
    S->dd_static = value != 0.0;
  } else if (which == 3) {
    // streams of the Jacobian fill
    if (value < 1.0 || value > (double)kad_system_impl::kFillMax || value != (double)(int)value) return KAD_ERR_ARG;
    S->nfill = (int)value;
  } else {
    return KAD_ERR_ARG;
  }
  return KAD_OK;
}
}  // extern "C"
