// Device context + C-ABI for the dense part of the Newton step: LU factorisation,
// triangular solves, and fp64 iterative refinement (classical IR and GMRES-IR)
// on low-precision (fp32 / bf16-updated) factors.
//
// Reference behaviour replaced: pdgesv_ in System_of_eqs::do_newton
// (src_par/do_newton.cpp:317-330; info != 0 -> "Jacobian is singular!") and
// Matrice::set_lu / Matrice::solve (src/Matrice/matrice.cpp:192-250).
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include "../../include/kadath_amd.h"
#include <vector>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <cstdlib>
#include <functional>

namespace kad {
int gemm_launch(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                int64_t ldc, hipStream_t s);
template <typename T>
int getrf(T* A, int64_t lda, int n, int* ipiv, int nb, int gemm_prec, LuWork& w, const LuBatch& bt = LuBatch());
int lu_solve_permuted_batched(const double* LU, int64_t lda, int n, int nbat, int64_t sA, double* B, int64_t ldb,
                              int64_t sB, int nrhs, LuWork& w, const int* zcols, const int* zstart, int zng);
size_t trsv_table_doubles(int n);
size_t trsv_flag_ints(int n);
size_t ffp_bytes();
template <typename T>
int trsv_prepare(const T* LU, int64_t lda, int n, int trans, double* W, int* flag, TrsvExtents& ext, hipStream_t s);
template <typename T>
int getrs_fast(const T* LU, int64_t lda, int n, const int* perm, const double* W, const TrsvExtents& ext, double* b,
               int trans, double* work, hipStream_t s);
bool ipiv_to_perm(const int* ipiv, int n, int* perm);
int noalloc_init(int device);
int gemv64(const double* A, int64_t lda, int n, const double* x, const double* b, double* y, int trans, double* part,
           hipStream_t s);
__global__ void axpy_kernel(int n, double a, const double* x, double* y);
__global__ void absmax_partial_kernel(const double* x, int n, double* part);
__global__ void dot_partial_kernel(const double* x, const double* y, int n, double* part);
__global__ void multidot_kernel(const double* __restrict__ V, int n, int k, const double* __restrict__ w,
                                double* __restrict__ h);
__global__ void multiaxpy_kernel(const double* __restrict__ V, int n, int k, const double* __restrict__ h,
                                 double* __restrict__ w, int sub);
__global__ void cast_mat_d2f_kernel(const double* __restrict__ a, int64_t lda, float* __restrict__ b, int64_t ldb,
                                    int n);
__global__ void row_absmax_kernel(const double* __restrict__ A, int64_t lda, int n, double* __restrict__ R);
__global__ void col_scale_kernel(double* __restrict__ A, int64_t lda, int n, const double* __restrict__ R,
                                 double* __restrict__ C);
__global__ void vec_scale_kernel(int n, const double* __restrict__ s, double* __restrict__ v);
__global__ void col_scale_only_kernel(const double* __restrict__ A, int64_t lda, int n, const double* __restrict__ R,
                                      double* __restrict__ C);
__global__ void cast_scaled_kernel(const double* __restrict__ a, int64_t lda, float* __restrict__ b, int64_t ldb,
                                   int n, const double* __restrict__ R, const double* __restrict__ C);

static constexpr int RED_BLOCKS = 256;
GemmTimer* g_gemm_timer = nullptr;
GemmTimer* g_apply_timer = nullptr;

int GemmTimer::begin(hipStream_t s, double work, bool scaled_by_device) {
  if (used >= kCap) {
    overflow = true;
    return 1;
  }
  if (2 * used + 2 > ev.size()) {
    const size_t old = ev.size();
    ev.resize(old + 2048);
    for (size_t q = old; q < ev.size(); ++q)
      if (hipEventCreate(&ev[q]) != hipSuccess) return KAD_ERR_HIP;
  }
  flops.push_back(work);
  scaled.push_back(scaled_by_device ? 1 : 0);
  return hipEventRecord(ev[2 * used], s) == hipSuccess ? KAD_OK : KAD_ERR_HIP;
}

int GemmTimer::end(hipStream_t s) {
  if (hipEventRecord(ev[2 * used + 1], s) != hipSuccess) return KAD_ERR_HIP;
  ++used;
  return KAD_OK;
}

int Ctx::ensure(void** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return KAD_OK;
  if (*p) KAD_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  KAD_HIP_CHECK(hipMalloc(p, bytes));
  *cap = bytes;
  return KAD_OK;
}

int Ctx::lu_work(int n, int nb) {
  const int red = (int)cdiv(n, 64) + 1;
  int rc;
  if ((rc = ensure((void**)&lu.red_val, &cap_redv, 2 * sizeof(double) * red))) return rc;
  if ((rc = ensure((void**)&lu.red_idx, &cap_redi, 2 * sizeof(int) * red))) return rc;
  if ((rc = ensure((void**)&lu.map, &cap_map, 2 * sizeof(int) * (size_t)n))) return rc;
  const size_t nt = (size_t)cdiv(n, 128) + 1;
  if ((rc = ensure((void**)&lu.mask_r, &cap_mask, sizeof(int) * (7 * nt + 2)))) return rc;
  lu.mask_c = lu.mask_r + nt;
  lu.tile_list = lu.mask_r + 2 * nt;  // 2 nt + 2
  lu.mask_c2 = lu.mask_r + 4 * nt + 2;
  lu.mask_r2 = lu.mask_r + 5 * nt + 2;
  lu.mask_r3 = lu.mask_r + 6 * nt + 2;  // [6 nt + 2, 7 nt + 2)
  if ((rc = ensure((void**)&lu.ext, &lu.cap_ext, sizeof(int) * (size_t)cdiv(n, 128) * cdiv(n, 2048)))) return rc;
  if (!lu.stream2) {
    // same (default) priority as the main stream
    KAD_HIP_CHECK(hipStreamCreateWithFlags(&lu.stream2, hipStreamNonBlocking));
    for (int q = 0; q < 2; ++q) {
      KAD_HIP_CHECK(hipEventCreateWithFlags(&lu.ev_panel[q], hipEventDisableTiming));
      KAD_HIP_CHECK(hipEventCreateWithFlags(&lu.ev_next[q], hipEventDisableTiming));
    }
    KAD_HIP_CHECK(hipEventCreateWithFlags(&lu.ev_start, hipEventDisableTiming));
  }
  lu.red_cap = red;
  if (!lu.counter) {
    // counter, info (+16 B), the persistent panel's words (+32 B), batch info (+64 B),
    // the static-pivot panel's arrival counters, then its growth words (8-byte aligned)
    constexpr size_t words = 64 + 2 * LU_MAX_BATCH * sizeof(int) + LU_MAX_BATCH * sizeof(unsigned long long);
    KAD_HIP_CHECK(hipMalloc((void**)&lu.counter, words));
    KAD_HIP_CHECK(hipMemsetAsync(lu.counter, 0, words, stream));
    lu.info = (int*)((char*)lu.counter + 16);
    lu.binfo = (int*)((char*)lu.counter + 64);
    lu.bcnt = (unsigned*)((char*)lu.counter + 64 + LU_MAX_BATCH * sizeof(int));
    lu.bgrowth = (unsigned long long*)((char*)lu.counter + 64 + 2 * LU_MAX_BATCH * sizeof(int));
  }
  lu.ffp_sync = (int*)((char*)lu.counter + 32);  // [2]: abort flag (zeroed above)
  if (!lu.ffp) {
    KAD_HIP_CHECK(hipMalloc(&lu.ffp, ffp_bytes()));
    KAD_HIP_CHECK(hipMemsetAsync(lu.ffp, 0, ffp_bytes(), stream));  // flag 0: no column
    lu.ffp_seq = 0;
    KAD_HIP_CHECK(hipDeviceGetAttribute(&lu.ncu, hipDeviceAttributeMultiprocessorCount, device));
    const char* e = std::getenv("KAD_LU_PANEL");
    lu.ffp_off = e && std::strcmp(e, "percolumn") == 0;
    // a small limit makes the persistent panel abandon its waits (and the
    // factorisation report an error) instead of waiting: the failure path's test
    const char* sp = std::getenv("KAD_LU_SPIN");
    if (sp) lu.ffp_spin = std::max(0, std::atoi(sp));
  }
  if ((rc = ensure(&lu.tinv, &cap_tinv, (size_t)nb * nb * 8))) return rc;
  // GEMM scratch for L11^{-1} A12: nb x chunk (bounded at 64 MiB).
  const size_t want = std::min<size_t>((size_t)nb * (size_t)n * 8, (size_t)64 << 20);
  if ((rc = ensure(&lu.tmp, &lu.tmp_bytes, std::max<size_t>(want, (size_t)nb * nb * 8)))) return rc;
  // the main stream's own copies (sized after tmp: the look-ahead runs both at once)
  if ((rc = ensure(&lu.tinv_m, &cap_tinv_m, (size_t)nb * nb * 8))) return rc;
  if ((rc = ensure(&lu.tmp_m, &lu.tmp_m_bytes, lu.tmp_bytes))) return rc;
  if ((rc = ensure(&lu.tmp2, &lu.tmp2_bytes, std::max<size_t>((size_t)nb * nb * 8, (size_t)n * 32 * 8)))) return rc;
  lu.stream = stream;
  return KAD_OK;
}

int Ctx::vec_work(int n) {
  int rc;
  const size_t part = (size_t)cdiv(n, 1024) * n * sizeof(double);
  if ((rc = ensure((void**)&gemv_part, &cap_part, std::max<size_t>(part, RED_BLOCKS * sizeof(double))))) return rc;
  if ((rc = ensure((void**)&vwork, &cap_vwork, (size_t)8 * n * sizeof(double)))) return rc;
  if ((rc = ensure((void**)&trs_tab, &cap_trs_tab, trsv_table_doubles(n) * sizeof(double)))) return rc;
  if ((rc = ensure((void**)&trs_flag, &cap_trs_flag, trsv_flag_ints(n) * sizeof(int)))) return rc;
  return KAD_OK;
}

// Triangular-solve tables of one set of factors (for op(A), trans as in solve),
// then b <- op(A)^{-1} b with them.
int Ctx::solve_prepare(int prec, int n, const void* LU, int64_t lda, int trans) {
  if (prec == F64) return trsv_prepare<double>((const double*)LU, lda, n, trans, trs_tab, trs_flag, trs_ext, stream);
  return trsv_prepare<float>((const float*)LU, lda, n, trans, trs_tab, trs_flag, trs_ext, stream);
}

int Ctx::solve(int prec, int n, const void* LU, int64_t lda, const int* perm, double* b, int trans) {
  double* w = vwork + (size_t)4 * n;  // 3 n doubles; vwork[0:4n] belongs to the refinement loop
  if (prec == F64) return getrs_fast<double>((const double*)LU, lda, n, perm, trs_tab, trs_ext, b, trans, w, stream);
  return getrs_fast<float>((const float*)LU, lda, n, perm, trs_tab, trs_ext, b, trans, w, stream);
}

double Ctx::absmax(const double* x, int n) {
  absmax_partial_kernel<<<RED_BLOCKS, 256, 0, stream>>>(x, n, gemv_part);
  double h[RED_BLOCKS];
  (void)hipMemcpyAsync(h, gemv_part, sizeof(h), hipMemcpyDeviceToHost, stream);
  (void)hipStreamSynchronize(stream);
  double m = 0.0;
  for (int i = 0; i < RED_BLOCKS; ++i)
    if (h[i] > m || h[i] != h[i]) m = h[i];
  return m;
}

double Ctx::dot(const double* x, const double* y, int n) {
  dot_partial_kernel<<<RED_BLOCKS, 256, 0, stream>>>(x, y, n, gemv_part);
  double h[RED_BLOCKS];
  (void)hipMemcpyAsync(h, gemv_part, sizeof(h), hipMemcpyDeviceToHost, stream);
  (void)hipStreamSynchronize(stream);
  double s = 0.0;
  for (int i = 0; i < RED_BLOCKS; ++i) s += h[i];
  return s;
}

int Ctx::upload(const void* host, size_t bytes, void** dev) {
  void* p = nullptr;
  KAD_HIP_CHECK(hipMalloc(&p, bytes < 16 ? 16 : bytes));
  owned.push_back(p);
  KAD_HIP_CHECK(hipMemcpy(p, host, bytes, hipMemcpyHostToDevice));
  *dev = p;
  return KAD_OK;
}

Ctx::~Ctx() {
  for (auto& kv : plans) delete kv.second;
  for (void* p : owned) (void)hipFree(p);
  void* bufs[] = {lu.red_val, lu.red_idx, lu.counter, lu.tinv,    lu.tmp,    lu.tmp2,  lu.map,  lu.mask_r,
                  lu.swap_map, lu.ext,     lu.ffp,     gemv_part, vwork,     krylov,   trs_tab, trs_flag, bsolve};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (int q = 0; q < 2; ++q) {
    if (lu.ev_panel[q]) (void)hipEventDestroy(lu.ev_panel[q]);
    if (lu.ev_next[q]) (void)hipEventDestroy(lu.ev_next[q]);
  }
  if (lu.ev_start) (void)hipEventDestroy(lu.ev_start);
  if (lu.tinv_m) (void)hipFree(lu.tinv_m);
  if (lu.tmp_m) (void)hipFree(lu.tmp_m);
  if (lu.stream2) (void)hipStreamDestroy(lu.stream2);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

int ir_loop(Ctx* c, int n, const double* A64, int64_t lda, const std::function<int(double*)>& precond,
            const double* b, double* x, double tol, int max_outer, int mode, int trans, int* iters, double* resid,
            bool map_ready = false);

}  // namespace kad

using kad::Ctx;

extern "C" {

int kad_ctx_create(kad_ctx** out, int device) {
  if (!out) return KAD_ERR_ARG;
  KAD_HIP_CHECK(hipSetDevice(device));
  // the no-allocation entry points' static scratch: its address on this device
  int rc0 = kad::noalloc_init(device);
  if (rc0) return rc0;
  Ctx* c = new Ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return KAD_ERR_HIP;
  }
  c->own_stream = true;
  *out = reinterpret_cast<kad_ctx*>(c);
  return KAD_OK;
}

int kad_ctx_destroy(kad_ctx* ctx) {
  if (!ctx) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  (void)hipStreamSynchronize(c->stream);
  delete c;
  return KAD_OK;
}

// Live timing of one kernel class: kind 0 = MFMA GEMM launches (work in flops),
// kind 1 = fused 3-D spectral passes (work in HBM bytes).  on: start a new
// measurement; off: stop.
int kad_kernel_timing(kad_ctx* ctx, int kind, int on) {
  if (!ctx || kind < 0 || kind > 1) return KAD_ERR_ARG;
  static kad::GemmTimer timers[2];
  kad::GemmTimer*& slot = kind == 0 ? kad::g_gemm_timer : kad::g_apply_timer;
  if (on) {
    kad::GemmTimer& t = timers[kind];
    KAD_HIP_CHECK(hipSetDevice(reinterpret_cast<Ctx*>(ctx)->device));
    KAD_HIP_CHECK(hipDeviceSynchronize());  // earlier launches may still write dscale
    if (!t.dscale) KAD_HIP_CHECK(hipMalloc((void**)&t.dscale, sizeof(double) * kad::GemmTimer::kCap));
    t.used = 0;
    t.overflow = false;
    t.flops.clear();
    t.scaled.clear();
    slot = &t;
  } else {
    slot = nullptr;
  }
  return KAD_OK;
}

// Total work and kernel milliseconds of the launches of one class since
// kad_kernel_timing(ctx, kind, 1).
int kad_kernel_timing_read(kad_ctx* ctx, int kind, double* work, double* ms, int* launches) {
  if (!ctx || kind < 0 || kind > 1) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  KAD_HIP_CHECK(hipSetDevice(c->device));
  // launches of the window may sit on any of the context's streams
  KAD_HIP_CHECK(hipDeviceSynchronize());
  kad::GemmTimer* tm = kind == 0 ? kad::g_gemm_timer : kad::g_apply_timer;
  double f = 0, tot = 0;
  int n = 0;
  if (tm) {
    if (tm->overflow) return KAD_ERR_ALLOC;
    std::vector<double> sc(tm->used, 1.0);
    if (tm->used)
      KAD_HIP_CHECK(hipMemcpy(sc.data(), tm->dscale, sizeof(double) * tm->used, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < tm->used; ++q) {
      float e = 0;
      KAD_HIP_CHECK(hipEventElapsedTime(&e, tm->ev[2 * q], tm->ev[2 * q + 1]));
      tot += e;
      f += tm->flops[q] * (tm->scaled[q] ? sc[q] : 1.0);
      ++n;
    }
  }
  if (work) *work = f;
  if (ms) *ms = tot;
  if (launches) *launches = n;
  return KAD_OK;
}

void* kad_ctx_stream(kad_ctx* ctx) { return ctx ? (void*)reinterpret_cast<Ctx*>(ctx)->stream : nullptr; }

int kad_sync(kad_ctx* ctx) {
  if (!ctx) return KAD_ERR_ARG;
  KAD_HIP_CHECK(hipStreamSynchronize(reinterpret_cast<Ctx*>(ctx)->stream));
  return KAD_OK;
}

int kad_gemm(kad_ctx* ctx, int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B,
             int64_t ldb, void* C, int64_t ldc) {
  if (!ctx || prec < 0 || prec > 2 || (op != 0 && op != 1)) return KAD_ERR_ARG;
  if (m < 0 || n < 0 || k < 0 || lda < std::max(1, m) || ldb < std::max(1, k) || ldc < std::max(1, m))
    return KAD_ERR_ARG;
  return kad::gemm_launch(prec, op, m, n, k, A, lda, B, ldb, C, ldc, reinterpret_cast<Ctx*>(ctx)->stream);
}

// In-place LU.  prec: 0 = fp64 storage/fp64 MFMA, 1 = fp32/fp32 MFMA, 2 = fp32 storage,
// bf16 MFMA trailing update with fp32 accumulation.  ipiv (device, n ints) receives
// 0-based LAPACK-style interchanges; perm (device, n ints) the composed row order.
// Returns 0, a negative error, or i > 0 when U(i-1,i-1) is exactly zero (LAPACK info).
int kad_lu_factor(kad_ctx* ctx, int prec, int n, void* A, int64_t lda, int* ipiv, int* perm, int nb) {
  if (!ctx || !A || !ipiv || n <= 0 || lda < n || prec < 0 || prec > 2) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (nb <= 0) {
    // block size: KAD_LU_NB (a multiple of the 128-row tiles), else 256 (512 for BF16)
    nb = (prec == kad::BF16) ? 512 : 256;
    const char* e = std::getenv("KAD_LU_NB");
    if (e && std::atoi(e) > 0 && std::atoi(e) % 128 == 0) nb = std::atoi(e);
  }
  if (nb % 16 != 0) return KAD_ERR_ARG;
  int rc = c->lu_work(n, nb);
  if (rc) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(c->lu.info, 0, sizeof(int), c->stream));
  if (prec == kad::F64)
    rc = kad::getrf<double>((double*)A, lda, n, ipiv, nb, kad::F64, c->lu);
  else
    rc = kad::getrf<float>((float*)A, lda, n, ipiv, nb, prec, c->lu);
  // both streams of the factorisation have finished before its words are read (also
  // when getrf returned early)
  KAD_HIP_CHECK(hipStreamSynchronize(c->lu.stream2));
  // info (counter + 16 B) and the persistent panel's barrier words (counter + 32 B)
  int words[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  KAD_HIP_CHECK(hipMemcpyAsync(words, (char*)c->lu.counter + 16, sizeof(words), hipMemcpyDeviceToHost, c->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
  const int info = words[0];
  if (words[6] != 0) {
    // a wait of the persistent panel reached its spin limit: the factors are not
    // valid; clear the abort word for the next call
    std::fprintf(stderr, "kad_lu_factor: persistent panel wait timed out\n");
    KAD_HIP_CHECK(hipMemsetAsync(c->lu.ffp_sync, 0, 4 * sizeof(int), c->stream));
    KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return KAD_ERR_HIP;
  }
  if (rc) return rc;
  if (perm) {
    std::vector<int> hp(n), hperm(n);
    KAD_HIP_CHECK(hipMemcpyAsync(hp.data(), ipiv, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (!kad::ipiv_to_perm(hp.data(), n, hperm.data())) return KAD_ERR_HIP;
    KAD_HIP_CHECK(hipMemcpyAsync(perm, hperm.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
  return info;
}

// Solve op(A) x = b in place (b: device fp64, n).  trans = 1 solves A^T x = b
// (the adjoint solve J^T lambda = dQ/du reuses the same factors).
int kad_lu_solve(kad_ctx* ctx, int prec, int n, const void* LU, int64_t lda, const int* perm, double* b, int trans) {
  if (!ctx || !LU || !perm || !b || n <= 0 || lda < n) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = c->vec_work(n);
  if (rc) return rc;
  if ((rc = c->solve_prepare(prec, n, LU, lda, trans))) return rc;
  return c->solve(prec, n, LU, lda, perm, b, trans);
}

}  // extern "C"

namespace {
// Block size of the batched factorisations: KAD_LU_NB when it is a multiple of 256 that
// divides n, else 256.
int batched_nb(int n) {
  const char* e = std::getenv("KAD_LU_NB");
  const int v = e ? std::atoi(e) : 0;
  return (v > 0 && v % 256 == 0 && n % v == 0) ? v : 256;
}
}  // namespace

extern "C" {

// nbat matrices of order n (a multiple of 256) at A + z strideA, factored together.
int kad_lu_factor_batched(kad_ctx* ctx, int prec, int n, int nbat, void* A, int64_t lda, int64_t strideA, int* ipiv,
                          int* perm, int* info) {
  if (!ctx || !A || !ipiv || !info || n <= 0 || n % 256 != 0 || nbat < 1 || nbat > kad::LU_MAX_BATCH || lda < n ||
      prec < 0 || prec > 2 || (nbat > 1 && strideA < lda * (int64_t)n)) {
    std::fprintf(stderr, "kad_lu_factor_batched: bad arguments (n %d, nbat %d, lda %lld, strideA %lld)\n", n, nbat,
                 (long long)lda, (long long)strideA);
    return KAD_ERR_ARG;
  }
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const int nb = batched_nb(n);
  int rc = c->lu_work(n, nb);
  if (rc) return rc;
  // every matrix's workgroups of the persistent panel are resident together
  if ((int)kad::cdiv(n, 256) * nbat > c->lu.ncu) {
    std::fprintf(stderr, "kad_lu_factor_batched: %d matrices of order %d exceed the %d CUs\n", nbat, n, c->lu.ncu);
    return KAD_ERR_ARG;
  }
  KAD_HIP_CHECK(hipMemsetAsync(c->lu.binfo, 0, nbat * sizeof(int), c->stream));
  kad::LuBatch bt;
  bt.nbat = nbat;
  bt.sA = strideA;
  bt.sP = n;
  bt.info = c->lu.binfo;
  if (prec == kad::F64)
    rc = kad::getrf<double>((double*)A, lda, n, ipiv, nb, kad::F64, c->lu, bt);
  else
    rc = kad::getrf<float>((float*)A, lda, n, ipiv, nb, prec, c->lu, bt);
  KAD_HIP_CHECK(hipStreamSynchronize(c->lu.stream2));
  int words[4] = {0, 0, 0, 0};
  KAD_HIP_CHECK(hipMemcpyAsync(words, c->lu.ffp_sync, sizeof(words), hipMemcpyDeviceToHost, c->stream));
  KAD_HIP_CHECK(hipMemcpyAsync(info, c->lu.binfo, nbat * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (words[2] != 0) {
    std::fprintf(stderr, "kad_lu_factor_batched: persistent panel wait timed out\n");
    KAD_HIP_CHECK(hipMemsetAsync(c->lu.ffp_sync, 0, 4 * sizeof(int), c->stream));
    KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return KAD_ERR_HIP;
  }
  if (rc) return rc;
  if (perm) {
    std::vector<int> hp((size_t)n * nbat), hperm((size_t)n * nbat);
    KAD_HIP_CHECK(hipMemcpyAsync(hp.data(), ipiv, hp.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int z = 0; z < nbat; ++z)
      if (!kad::ipiv_to_perm(hp.data() + (size_t)z * n, n, hperm.data() + (size_t)z * n)) return KAD_ERR_HIP;
    KAD_HIP_CHECK(hipMemcpyAsync(perm, hperm.data(), hperm.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return KAD_OK;
}

// The same without row interchanges (static pivoting): the caller has put the rows of
// every matrix in the order of an earlier factorisation's pivots (one matrix may have
// any order; a batch, orders that are multiples of 256).  ipiv (nbat x n) gets
// the identity; info[z] the first exactly-zero pivot; growth[z] the largest |multiplier|
// of matrix z (partial pivoting keeps it at most 1; a large one means the order no longer
// suits the matrix).  HOST info and growth.
int kad_lu_factor_batched_static(kad_ctx* ctx, int prec, int n, int nbat, void* A, int64_t lda, int64_t strideA,
                                 int* ipiv, int* info, double* growth) {
  if (!ctx || !A || !ipiv || !info || !growth || n <= 0 || (nbat > 1 && n % 256 != 0) || nbat < 1 ||
      nbat > kad::LU_MAX_BATCH || lda < n || prec < 0 || prec > 2 || (nbat > 1 && strideA < lda * (int64_t)n)) {
    std::fprintf(stderr, "kad_lu_factor_batched_static: bad arguments (n %d, nbat %d, lda %lld, strideA %lld)\n", n,
                 nbat, (long long)lda, (long long)strideA);
    return KAD_ERR_ARG;
  }
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const int nb = batched_nb(n);
  int rc = c->lu_work(n, nb);
  if (rc) return rc;
  KAD_HIP_CHECK(hipMemsetAsync(c->lu.binfo, 0, nbat * sizeof(int), c->stream));
  kad::LuBatch bt;
  bt.nbat = nbat;
  bt.sA = strideA;
  bt.sP = n;
  bt.info = c->lu.binfo;
  bt.nopiv = true;
  bt.growth = c->lu.bgrowth;
  if (prec == kad::F64)
    rc = kad::getrf<double>((double*)A, lda, n, ipiv, nb, kad::F64, c->lu, bt);
  else
    rc = kad::getrf<float>((float*)A, lda, n, ipiv, nb, prec, c->lu, bt);
  KAD_HIP_CHECK(hipStreamSynchronize(c->lu.stream2));
  unsigned long long g[kad::LU_MAX_BATCH];
  KAD_HIP_CHECK(hipMemcpyAsync(info, c->lu.binfo, nbat * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  KAD_HIP_CHECK(hipMemcpyAsync(g, c->lu.bgrowth, nbat * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (rc) return rc;
  for (int z = 0; z < nbat; ++z) {
    double v;
    std::memcpy(&v, &g[z], sizeof v);
    growth[z] = v;
  }
  return KAD_OK;
}

namespace {
__global__ void gather_rows_batched_kernel(const double* __restrict__ B, int64_t ldb, int64_t sB,
                                           const int* __restrict__ perm, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y, z = blockIdx.z;
  if (i < n) out[i + (int64_t)n * (col + (int64_t)gridDim.y * z)] = B[perm[(int64_t)z * n + i] + col * ldb + z * sB];
}
__global__ void put_back_batched_kernel(const double* __restrict__ in, int n, double* __restrict__ B, int64_t ldb,
                                        int64_t sB) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y, z = blockIdx.z;
  if (i < n) B[i + col * ldb + z * sB] = in[i + (int64_t)n * (col + (int64_t)gridDim.y * z)];
}
}  // namespace

// B_z <- A_z^{-1} B_z (nrhs columns each) with the factors of kad_lu_factor_batched (FP64).
int kad_lu_solve_batched(kad_ctx* ctx, int n, int nbat, const double* LU, int64_t lda, int64_t strideA, const int* perm,
                         double* B, int64_t ldb, int64_t strideB, int nrhs) {
  if (!ctx || !LU || !perm || !B || n <= 0 || n % 32 != 0 || nbat < 1 || nbat > 65535 || lda < n || ldb < n ||
      nrhs < 1 || nrhs > 65535 || (nbat > 1 && strideB < ldb * (int64_t)nrhs))
    return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = c->lu_work(n, 256);
  if (rc) return rc;
  const size_t need = (size_t)n * nrhs * nbat * sizeof(double);
  if ((rc = c->ensure((void**)&c->bsolve, &c->cap_bsolve, need))) return rc;
  dim3 g((unsigned)kad::cdiv(n, 256), (unsigned)nrhs, (unsigned)nbat);
  gather_rows_batched_kernel<<<g, 256, 0, c->stream>>>(B, ldb, strideB, perm, n, c->bsolve);
  KAD_HIP_CHECK(hipGetLastError());
  if ((rc = kad::lu_solve_permuted_batched(LU, lda, n, nbat, strideA, c->bsolve, n, (int64_t)n * nrhs, nrhs, c->lu, nullptr, nullptr, 0)))
    return rc;
  put_back_batched_kernel<<<g, 256, 0, c->stream>>>(c->bsolve, n, B, ldb, strideB);
  KAD_HIP_CHECK(hipGetLastError());
  KAD_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KAD_OK;
}

int kad_gemv64(kad_ctx* ctx, int n, const double* A, int64_t lda, const double* x, const double* b, double* y,
               int trans) {
  if (!ctx || n <= 0 || lda < n) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = c->vec_work(n);
  if (rc) return rc;
  return kad::gemv64(A, lda, n, x, b, y, trans, c->gemv_part, c->stream);
}

// A <- diag(R) A diag(C) in place with power-of-two R, C (device vectors, n each).
int kad_equilibrate(kad_ctx* ctx, int n, double* A, int64_t lda, double* R, double* C) {
  if (!ctx || n <= 0 || lda < n || !A || !R || !C) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  kad::row_absmax_kernel<<<(unsigned)kad::cdiv(n, 256), 256, 0, c->stream>>>(A, lda, n, R);
  kad::col_scale_kernel<<<n, 256, 0, c->stream>>>(A, lda, n, R, C);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// Equilibrated low-precision copy: computes R, C (powers of two, as
// kad_equilibrate) and B = (float)(diag(R) A diag(C)) without modifying A.
int kad_cast_equilibrated(kad_ctx* ctx, int n, const double* A, int64_t lda, float* B, int64_t ldb, double* R,
                          double* C) {
  if (!ctx || n <= 0 || lda < n || ldb < n || !A || !B || !R || !C) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  kad::row_absmax_kernel<<<(unsigned)kad::cdiv(n, 256), 256, 0, c->stream>>>(A, lda, n, R);
  kad::col_scale_only_kernel<<<n, 256, 0, c->stream>>>(A, lda, n, R, C);
  dim3 g((unsigned)kad::cdiv(n, 256), n);
  kad::cast_scaled_kernel<<<g, 256, 0, c->stream>>>(A, lda, B, ldb, n, R, C);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// v <- s * v (elementwise, device)
int kad_vec_scale(kad_ctx* ctx, int n, const double* s, double* v) {
  if (!ctx || n <= 0 || !s || !v) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  kad::vec_scale_kernel<<<(unsigned)kad::cdiv(n, 256), 256, 0, c->stream>>>(n, s, v);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// fp64 n x n column-major -> fp32 copy (device).
int kad_cast_f64_to_f32(kad_ctx* ctx, int n, const double* A, int64_t lda, float* B, int64_t ldb) {
  if (!ctx || n <= 0 || lda < n || ldb < n) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  dim3 g((unsigned)kad::cdiv(n, 256), n);
  kad::cast_mat_d2f_kernel<<<g, 256, 0, c->stream>>>(A, lda, B, ldb, n);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

// Mixed-precision refinement: x = A^{-1} b with A given in fp64 (A64, for the
// residual) and factored in prec_lo (LUlo, perm).  mode 0: classical IR;
// mode 1: GMRES-IR (right preconditioned by the low-precision factors, CGS2,
// restart m = kad_ir_set_restart, default 100).  Stops when
// ||b - A x||_inf <= tol * ||b||_inf, or at the attainable accuracy: an outer step
// that gains less than a factor 2 ends the loop, and one that increases the residual
// is undone.  trans = 1 solves A^T x = b.
// iters[0] = outer steps, iters[1] = total inner GMRES steps; resid = final relative residual.
int kad_ir_solve_scaled(kad_ctx* ctx, int n, const double* A64, int64_t lda, int prec_lo, const void* LUlo,
                        int64_t ldlu, const int* perm, const double* Rs, const double* Cs, const double* b, double* x,
                        double tol, int max_outer, int mode, int trans, int* iters, double* resid);

int kad_ir_solve(kad_ctx* ctx, int n, const double* A64, int64_t lda, int prec_lo, const void* LUlo, int64_t ldlu,
                 const int* perm, const double* b, double* x, double tol, int max_outer, int mode, int trans,
                 int* iters, double* resid) {
  return kad_ir_solve_scaled(ctx, n, A64, lda, prec_lo, LUlo, ldlu, perm, nullptr, nullptr, b, x, tol, max_outer,
                             mode, trans, iters, resid);
}

// As kad_ir_solve, with the low-precision factors being those of diag(Rs) A diag(Cs):
// the preconditioner applies v -> Cs .* LU^{-1} (Rs .* v) (Rs <-> Cs for the transpose),
// while residuals and the stopping test use the unscaled A.
//
// GMRES-IR is the flexible variant: the preconditioned vectors z_j = M^{-1} v_j are
// kept, and the correction is Z y.  The Arnoldi relation A Z = V H then holds to
// fp64 GEMV accuracy whatever the rounding of the low-precision triangular solves
// (which is not a linear map), so the attainable residual is set by the fp64
// products and not by the preconditioner.  Products with A skip its exactly-zero
// tiles (scanned once per call).
int kad_ir_solve_scaled(kad_ctx* ctx, int n, const double* A64, int64_t lda, int prec_lo, const void* LUlo,
                        int64_t ldlu, const int* perm, const double* Rs, const double* Cs, const double* b, double* x,
                        double tol, int max_outer, int mode, int trans, int* iters, double* resid) {
  if (!ctx || n <= 0 || lda < n || ldlu < n || !A64 || !LUlo || !perm || !b || !x || max_outer < 1 ||
      (mode != 0 && mode != 1) || (trans != 0 && trans != 1) || prec_lo < 0 || prec_lo > 2)
    return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = c->vec_work(n);
  if (rc) return rc;
  hipStream_t s = c->stream;
  const unsigned g = (unsigned)kad::cdiv(n, 256);
  if ((rc = c->solve_prepare(prec_lo, n, LUlo, ldlu, trans))) return rc;
  const double* pre = trans ? Cs : Rs;   // applied before the factor solve
  const double* post = trans ? Rs : Cs;  // applied after
  auto solve_lo = [&](double* v) {
    if (pre) kad::vec_scale_kernel<<<g, 256, 0, s>>>(n, pre, v);
    const int r = c->solve(prec_lo, n, LUlo, ldlu, perm, v, trans);
    if (!r && post) kad::vec_scale_kernel<<<g, 256, 0, s>>>(n, post, v);
    return r;
  };
  return kad::ir_loop(c, n, A64, lda, solve_lo, b, x, tol, max_outer, mode, trans, iters, resid);
}

}  // extern "C"

namespace kad {

// The refinement loop of kad_ir_solve_scaled for any preconditioner: precond(v)
// replaces v by M^{-1} v (op(M) for trans = 1).  Residuals and products use the
// fp64 A; mode 0 is classical refinement, mode 1 flexible GMRES.  map_ready: c->jmap
// already holds A's nonzero tiles (adopted from a scan of this same A), no new scan.
int ir_loop(Ctx* c, int n, const double* A64, int64_t lda, const std::function<int(double*)>& precond,
            const double* b, double* x, double tol, int max_outer, int mode, int trans, int* iters, double* resid,
            bool map_ready) {
  int rc = c->vec_work(n);
  if (rc) return rc;
  hipStream_t s = c->stream;
  const unsigned g = (unsigned)kad::cdiv(n, 256);
  double* r = c->vwork + (size_t)2 * n;
  double* d = c->vwork + (size_t)3 * n;
  if (!map_ready || c->jmap.A != A64 || c->jmap.lda != lda || c->jmap.n != n)
    if ((rc = c->jmap.build(A64, lda, n, s))) return rc;
  const kad::GemvMap& Am = c->jmap;
  auto& solve_lo = precond;
  const double bnorm = c->absmax(b, n);
  c->hist_inner.clear();
  c->hist_outer.assign(1, bnorm);
  KAD_HIP_CHECK(hipMemsetAsync(x, 0, n * sizeof(double), s));
  KAD_HIP_CHECK(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  int outer = 0, inner_total = 0, slow = 0;
  double rel = (bnorm == 0.0) ? 0.0 : 1.0;
  if (bnorm != bnorm) rel = bnorm;  // NaN right-hand side: never reported as solved
  const int m = std::max(2, c->gmres_m);  // Krylov dimension per outer step
  const double inner_tol = 1e-14;
  std::vector<double> H, cs, sn, gvec, y;
  double* V = nullptr;
  double* Z = nullptr;
  double* hbuf = nullptr;
  if (mode == 1) {
    // V: m + 1 basis vectors, Z: m preconditioned vectors, then the small h / y buffer
    rc = c->ensure((void**)&c->krylov, &c->cap_krylov,
                   (size_t)(2 * m + 1) * n * sizeof(double) + (m + 8) * sizeof(double));
    if (rc) return rc;
    V = c->krylov;
    Z = c->krylov + (size_t)(m + 1) * n;
    hbuf = c->krylov + (size_t)(2 * m + 1) * n;
  }
  double* corr = d;  // the correction added by the current outer step
  while (rel > tol && outer < max_outer) {
    if (mode == 0) {
      KAD_HIP_CHECK(hipMemcpyAsync(d, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      if ((rc = solve_lo(d))) return rc;
      kad::axpy_kernel<<<g, 256, 0, s>>>(n, 1.0, d, x);
      corr = d;
    } else {
      // Flexible right-preconditioned GMRES on A M^{-1} u = r (M = low-precision LU),
      // CGS2 with batched projections; the Givens residual is the inner residual.
      const double beta = std::sqrt(c->dot(r, r, n));
      if (!(beta > 0.0)) break;
      KAD_HIP_CHECK(hipMemcpyAsync(V, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      kad::axpy_kernel<<<g, 256, 0, s>>>(n, 1.0 / beta - 1.0, V, V);
      H.assign((size_t)(m + 1) * m, 0.0);
      cs.assign(m, 0.0);
      sn.assign(m, 0.0);
      gvec.assign(m + 1, 0.0);
      gvec[0] = beta;
      int kdim = 0;
      std::vector<double> hh(m + 1);
      for (int j = 0; j < m; ++j) {
        double* vj1 = V + (size_t)(j + 1) * n;
        double* zj = Z + (size_t)j * n;
        KAD_HIP_CHECK(hipMemcpyAsync(zj, V + (size_t)j * n, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        if ((rc = solve_lo(zj))) return rc;  // z_j = M^{-1} v_j (kept)
        if ((rc = Am.apply(zj, nullptr, vj1, trans, c->gemv_part, s))) return rc;
        for (int pass = 0; pass < 2; ++pass) {
          kad::multidot_kernel<<<j + 1, 256, 0, s>>>(V, n, j + 1, vj1, hbuf);
          kad::multiaxpy_kernel<<<g, 256, 0, s>>>(V, n, j + 1, hbuf, vj1, 1);
          KAD_HIP_CHECK(hipMemcpyAsync(hh.data(), hbuf, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
          KAD_HIP_CHECK(hipStreamSynchronize(s));
          for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] += hh[i];
        }
        const double hn = std::sqrt(c->dot(vj1, vj1, n));
        H[(size_t)(j + 1) * m + j] = hn;
        if (hn > 0.0) kad::axpy_kernel<<<g, 256, 0, s>>>(n, 1.0 / hn - 1.0, vj1, vj1);
        for (int i = 0; i < j; ++i) {
          const double a = H[(size_t)i * m + j], bb = H[(size_t)(i + 1) * m + j];
          H[(size_t)i * m + j] = cs[i] * a + sn[i] * bb;
          H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
        }
        const double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
        const double rr = std::hypot(a, bb);
        cs[j] = (rr == 0.0) ? 1.0 : a / rr;
        sn[j] = (rr == 0.0) ? 0.0 : bb / rr;
        H[(size_t)j * m + j] = rr;
        H[(size_t)(j + 1) * m + j] = 0.0;
        gvec[j + 1] = -sn[j] * gvec[j];
        gvec[j] = cs[j] * gvec[j];
        kdim = j + 1;
        ++inner_total;
        c->hist_inner.push_back(std::fabs(gvec[j + 1]));
        // |g_{j+1}| is ||r - A Z y||_2 >= the inf-norm the outer test uses
        if (std::fabs(gvec[j + 1]) <= std::max(inner_tol * beta, 0.5 * tol * bnorm) || hn == 0.0) break;
      }
      y.assign(kdim, 0.0);
      for (int i = kdim - 1; i >= 0; --i) {
        double sacc = gvec[i];
        for (int l = i + 1; l < kdim; ++l) sacc -= H[(size_t)i * m + l] * y[l];
        y[i] = (H[(size_t)i * m + i] != 0.0) ? sacc / H[(size_t)i * m + i] : 0.0;
      }
      KAD_HIP_CHECK(hipMemcpyAsync(hbuf, y.data(), kdim * sizeof(double), hipMemcpyHostToDevice, s));
      kad::multiaxpy_kernel<<<g, 256, 0, s>>>(Z, n, kdim, hbuf, d, 0);  // correction Z y
      kad::axpy_kernel<<<g, 256, 0, s>>>(n, 1.0, d, x);
      corr = d;
    }
    // r = b - op(A) x in fp64
    if ((rc = Am.apply(x, b, r, trans, c->gemv_part, s))) return rc;
    const double rn = c->absmax(r, n);
    const double rel_new = (bnorm == 0.0) ? 0.0 : rn / bnorm;
    c->hist_outer.push_back(rn);
    ++outer;
    if (!(rel_new <= rel)) {
      // the step made the residual worse (attainable accuracy reached) or produced a
      // NaN: undo it
      kad::axpy_kernel<<<g, 256, 0, s>>>(n, -1.0, corr, x);
      if ((rc = Am.apply(x, b, r, trans, c->gemv_part, s))) return rc;
      break;
    }
    // two outer steps in a row that gain less than a halving: attainable accuracy
    slow = (rel_new > 0.5 * rel) ? slow + 1 : 0;
    rel = rel_new;
    if (slow >= 2) break;
  }
  if (iters) {
    iters[0] = outer;
    iters[1] = inner_total;
  }
  if (resid) *resid = rel;
  KAD_HIP_CHECK(hipStreamSynchronize(s));
  return (rel <= tol) ? KAD_OK : KAD_ERR_NOCONV;
}

}  // namespace kad

extern "C" {

// GMRES-IR Krylov dimension per restart of the context (default 100).
int kad_ir_set_restart(kad_ctx* ctx, int m) {
  if (!ctx || m < 2 || m > 4096) return KAD_ERR_ARG;
  reinterpret_cast<Ctx*>(ctx)->gmres_m = m;
  return KAD_OK;
}

// Residual history of the last refinement: kind 0 = inner GMRES residual 2-norms
// (one per inner step), kind 1 = ||b||_inf then the true ||b - op(A) x||_inf after
// each outer step.  Returns the number of entries (copies at most cap).
int kad_ir_history(kad_ctx* ctx, int kind, double* out, int cap) {
  if (!ctx || kind < 0 || kind > 1) return KAD_ERR_ARG;
  const auto& h = kind == 0 ? reinterpret_cast<Ctx*>(ctx)->hist_inner : reinterpret_cast<Ctx*>(ctx)->hist_outer;
  if (out)
    for (int i = 0; i < (int)h.size() && i < cap; ++i) out[i] = h[i];
  return (int)h.size();
}

}  // extern "C"
