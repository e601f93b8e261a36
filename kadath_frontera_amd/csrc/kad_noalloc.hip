// Triangular solves with LU factors that allocate nothing: the variant of
// kad_lu_solve for callers that cannot allocate device memory (code being captured
// into a HIP graph, tight serving loops, harnesses that forbid hipMalloc).
//
// kad_lu_solve keeps its scratch in the context (the per-block tables W_k of the
// sweeps, the tile flags of the band extents, three work vectors), allocated on
// first use.  This variant takes the tables and work vectors from one file-scope
// device array of fixed size instead, so it serves n <= KAD_NOALLOC_MAX_N; it skips
// the band-extent scan (its host readback and flag buffer), and its sweeps cover
// every row below (or above) the diagonal block -- the same kernels, the same
// arithmetic order, hence the same results bit for bit.
//
// The array is per GPU and lives until the process ends.  Every call on a GPU uses
// that GPU's copy and overwrites it completely before reading it (the tables are
// rebuilt from the factors, the work vectors from b), so nothing needs clearing; two
// calls that overlap on one GPU (different streams or host threads) race.
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include "../../include/kadath_amd.h"
#include <cstdio>

namespace kad {

size_t trsv_table_doubles(int n);
template <typename T>
int trsv_prepare(const T* LU, int64_t lda, int n, int trans, double* W, int* flag, TrsvExtents& ext, hipStream_t s);
template <typename T>
int getrs_fast(const T* LU, int64_t lda, int n, const int* perm, const double* W, const TrsvExtents& ext, double* b,
               int trans, double* work, hipStream_t s);

constexpr int kNoallocMaxN = 8192;
// tables of both sweeps (2 x 64 blocks x 128 x 256) + 3 work vectors
constexpr size_t kNoallocTable = (size_t)2 * (kNoallocMaxN / 128) * 128 * 256;
constexpr size_t kNoallocDoubles = kNoallocTable + (size_t)3 * kNoallocMaxN;

static __device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU

// Set-up, once per device with that device current (called by kad_ctx_create): the
// host-side name `scratch` is not a device address, so the address is looked up.
int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return KAD_ERR_ARG;
  if (g_scratch[device]) return KAD_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    std::fprintf(stderr, "[kadath_amd] hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
                 hipGetErrorString(e));
    return KAD_ERR_HIP;
  }
  g_scratch[device] = static_cast<double*>(p);
  return KAD_OK;
}

}  // namespace kad

extern "C" {

// Solve op(A) x = b in place with the factors, as kad_lu_solve, for n <= 8192,
// without allocating: scratch is the library's static per-GPU array.  Work is queued
// on the context's stream.
int kad_lu_solve_noalloc(kad_ctx* ctx, int prec, int n, const void* LU, int64_t lda, const int* perm, double* b,
                         int trans) {
  if (!ctx || !LU || !perm || !b || n <= 0 || lda < n || prec < 0 || prec > 2 || (trans != 0 && trans != 1))
    return KAD_ERR_ARG;
  if (n > kad::kNoallocMaxN) {
    std::fprintf(stderr, "[kadath_amd] kad_lu_solve_noalloc: n = %d exceeds its static scratch (n <= %d)\n", n,
                 kad::kNoallocMaxN);
    return KAD_ERR_ARG;
  }
  kad::Ctx* c = reinterpret_cast<kad::Ctx*>(ctx);
  if (c->device < 0 || c->device >= kad::kMaxDevices || !kad::g_scratch[c->device]) return KAD_ERR_HIP;
  double* W = kad::g_scratch[c->device];
  double* work = W + kad::kNoallocTable;
  kad::TrsvExtents full;  // empty: every row (no allocation)
  int rc;
  if (prec == kad::F64) {
    if ((rc = kad::trsv_prepare<double>((const double*)LU, lda, n, trans, W, nullptr, full, c->stream))) return rc;
    return kad::getrs_fast<double>((const double*)LU, lda, n, perm, W, full, b, trans, work, c->stream);
  }
  if ((rc = kad::trsv_prepare<float>((const float*)LU, lda, n, trans, W, nullptr, full, c->stream))) return rc;
  return kad::getrs_fast<float>((const float*)LU, lda, n, perm, W, full, b, trans, work, c->stream);
}

}  // extern "C"
