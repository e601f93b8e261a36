// Device context: one per (process, GPU).  Owns the stream and every scratch
// buffer of the Newton path, so the C-ABI entry points carry no hidden globals
// (the reference keeps FFT plans and transform scratch in statics:
// src/Coef/coef_1d.cpp:34-44, include/memory.hpp:55-69).
#pragma once
#include "kad_common.hpp"
#include "kad_spectral.hpp"
#include <map>
#include <string>
#include <vector>

namespace kad {

// Selection of the C tiles a GEMM computes (all NULL: every tile).
//  mask mode: rmask[tm] == 0 / cmask[tn] == 0 marks an exactly-zero 128-row tile of A
//    / 128-column tile of B; those C tiles are skipped (C -= AB) or zeroed (C = AB).
//  list mode (cnt != NULL): only the pairs rlist[0..cnt[0]) x clist[0..cnt[1]) are
//    computed, compacted over the grid so the active tiles spread over every XCD
//    (C -= AB only).
//  batch: nbat products of one shape, launched together (grid z); product z reads
//    A + z sa and B + z sb and writes C + z sc (the masks apply to every product;
//    lists: nbat = 1 only).
struct TileSel {
  const int* rmask = nullptr;
  const int* cmask = nullptr;
  const int* rlist = nullptr;
  const int* clist = nullptr;
  const int* cnt = nullptr;
  // batch only (device, nbat entries): matrix z computes only the tiles holding its
  // first zcols[z] columns of C; the others are left as they are (OP_SUB only)
  const int* zcols = nullptr;
  // batch only (device, nbat x zng): the k rows of B_z are rows zthr - k .. zthr - 1 of
  // a taller right-hand side whose columns 64g .. 64g+63 are zero above its row
  // zstart[z zng + g]; a tile whose column groups all start at or past zthr would only
  // subtract zeros and is skipped (OP_SUB only)
  const int* zstart = nullptr;
  int zng = 0, zthr = 0;
  int nbat = 1;
  int64_t sa = 0, sb = 0, sc = 0;
};

// A batch of equally sized matrices factored together: matrix z at A + z sA (same
// leading dimension), its interchanges at ipiv + z sP, indices local to the matrix.
struct LuBatch {
  int nbat = 1;
  int64_t sA = 0;
  int sP = 0;
  int* info = nullptr;  // nbat words (first zero pivot of each matrix); NULL: LuWork::info
  // static pivoting: no row interchanges (the rows were put in the order of an earlier
  // factorisation's pivots); growth[z] receives the largest |multiplier| of matrix z
  // (bits of a non-negative double, zeroed by getrf)
  bool nopiv = false;
  unsigned long long* growth = nullptr;
};
constexpr int LU_MAX_BATCH = 16;

// Scratch of the LU factorisation.
struct LuWork {
  hipStream_t stream = nullptr;
  double* red_val = nullptr;    // per-block partial maxima of the pivot search
  int* red_idx = nullptr;
  unsigned* counter = nullptr;  // last-block counter (reset by the last block)
  int* info = nullptr;          // first exactly-zero pivot (1-based), 0 if none
  int* binfo = nullptr;         // the same per matrix of a batched factorisation (LU_MAX_BATCH)
  unsigned* bcnt = nullptr;     // arrival counters of the static-pivot base panel (LU_MAX_BATCH, kept at 0)
  unsigned long long* bgrowth = nullptr;  // largest |multiplier| per matrix of a static-pivot factorisation
  int red_cap = 0;
  void* tinv = nullptr;         // unit-lower inverse, nb x nb
  void* tmp = nullptr;          // GEMM scratch for L11^{-1} A12
  size_t tmp_bytes = 0;
  void* tmp2 = nullptr;         // scratch of the recursive triangular inverse / panel row gather
  size_t tmp2_bytes = 0;
  int* map = nullptr;           // 2 x n logical -> physical row maps of the base panel
  int* mask_r = nullptr;        // per 128-row tile of the L panel: any nonzero
  int* mask_c = nullptr;        // per 128-column tile of the U row panel: any nonzero
  int* tile_list = nullptr;     // compacted active row tiles, column tiles, and the 2 counts
  int* swap_map = nullptr;      // row gathers of the deferred left interchanges
  size_t swap_cap = 0;          // ints
  int* ext = nullptr;           // last nonzero row per (128-column tile, 2048-row chunk)
  size_t cap_ext = 0;
  int* mask_c2 = nullptr;       // column tiles of U12 beyond the next panel
  int* mask_r2 = nullptr;       // second and third row-tile masks (rotating by step)
  int* mask_r3 = nullptr;
  // look-ahead: the next panel is updated and factored on a second stream
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_panel[2] = {nullptr, nullptr}, ev_next[2] = {nullptr, nullptr}, ev_start = nullptr;
  // triangular-inverse and GEMM scratch of the main stream (the panel stream uses tinv / tmp)
  void* tinv_m = nullptr;
  void* tmp_m = nullptr;
  size_t tmp_m_bytes = 0;
  // persistent base panel: the flagged exchange slots, the abort word (ffp_sync[2])
  // and the CU count that bounds its grid
  void* ffp = nullptr;
  int* ffp_sync = nullptr;
  int ncu = 0;
  bool ffp_off = false;         // KAD_LU_PANEL=percolumn selects the per-column kernels
  int ffp_spin = 1 << 22;       // spin limit of the exchange's waits (KAD_LU_SPIN overrides)
  unsigned ffp_seq = 0;         // persistent-panel launches since the slots were cleared
};

// Optional live timing of one kernel class (HIP events on the launch stream) and
// its algorithmic work per launch: flops of the MFMA GEMMs, bytes of the fused
// 3-D spectral passes.
struct GemmTimer {
  static constexpr size_t kCap = size_t(1) << 18;  // timed launches per window
  std::vector<hipEvent_t> ev;   // pairs
  std::vector<double> flops;    // work per pair (flops or bytes)
  std::vector<char> scaled;     // pair's work is multiplied by dscale[pair] at read time
  double* dscale = nullptr;     // device, kCap entries: executed fraction written by the device
  size_t used = 0;
  bool overflow = false;        // more than kCap launches in the window (read reports an error)
  // Device slot for the executed fraction of the next pair (nullptr when the window is full).
  double* scale_slot() { return (dscale && used < kCap) ? dscale + used : nullptr; }
  // Opens a pair; returns 1 (nothing recorded) when the window is full.
  int begin(hipStream_t s, double work, bool scaled_by_device = false);
  int end(hipStream_t s);
};
extern GemmTimer* g_gemm_timer;   // GEMM launches (work = flops)
extern GemmTimer* g_apply_timer;  // apply3d launches (work = HBM bytes: arrays in + out)

// Nonzero 256 x 1024 tiles of an n x n fp64 matrix, for products that skip the
// structural zeros (kad_struct.hip).  Valid for the contents A had at build().
constexpr int GM_ROWS = 256;   // rows per GEMV tile
constexpr int GM_COLS = 1024;  // columns per GEMV tile (= partial-sum chunk)

struct GemvMap {
  const double* A = nullptr;
  int64_t lda = 0;
  int n = 0, nrb = 0, nchunk = 0, nact = 0;
  int* map = nullptr;   // nchunk x nrb flags
  int* list = nullptr;  // (row tile, chunk) pairs of the nonzero tiles
  size_t cap = 0;
  int build(const double* A, int64_t lda, int n, hipStream_t s);
  // as build, with the tile flags already made by another scan of the same A (device,
  // nchunk x nrb, 256-row by 1024-column tiles)
  int adopt(const double* A, int64_t lda, int n, const int* dmap, hipStream_t s);
  int apply(const double* x, const double* b, double* y, int trans, double* part, hipStream_t s) const;
  ~GemvMap();
};

// Row extents of the panel updates of the two triangular sweeps, per 128-column
// block (kad_trsv.hip): the forward update ends at row fhi[b], the backward one
// starts at row blo[b]; the factor is exactly zero outside.
struct TrsvExtents {
  std::vector<int> fhi, blo;
};

// Device copy of a chain of 1-D passes (matrix pool + per-line tables).
struct DevPlan {
  Passes P;
  double* pool = nullptr;
  int* lines = nullptr;
  int d_in[3] = {0, 0, 0};
  int d_out[3] = {0, 0, 0};
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  LuWork lu;
  size_t cap_redv = 0, cap_redi = 0, cap_tinv = 0, cap_map = 0, cap_mask = 0, cap_tinv_m = 0;
  double* gemv_part = nullptr;
  size_t cap_part = 0;
  double* vwork = nullptr;
  size_t cap_vwork = 0;
  double* krylov = nullptr;
  size_t cap_krylov = 0;
  double* bsolve = nullptr;     // permuted right-hand sides of the batched solve
  size_t cap_bsolve = 0;
  double* trs_tab = nullptr;    // triangular-solve tables of the current factors
  size_t cap_trs_tab = 0;
  int* trs_flag = nullptr;      // nonzero flags of the factor tiles (set-up scratch)
  size_t cap_trs_flag = 0;
  TrsvExtents trs_ext;          // row extents of the current factors' sweeps
  GemvMap jmap;                 // nonzero tiles of the refinement's fp64 matrix
  int gmres_m = 100;                // GMRES-IR Krylov dimension per restart
  std::vector<double> hist_inner;   // last refinement: inner GMRES residual 2-norms
  std::vector<double> hist_outer;   // last refinement: true residual inf-norms

  std::map<std::string, DevPlan*> plans;  // transform plans by key
  std::vector<void*> owned;                // device buffers freed with the context

  int ensure(void** p, size_t* cap, size_t bytes);
  int upload(const void* host, size_t bytes, void** dev);
  int lu_work(int n, int nb);
  int vec_work(int n);
  int solve_prepare(int prec, int n, const void* LU, int64_t lda, int trans);
  int solve(int prec, int n, const void* LU, int64_t lda, const int* perm, double* b, int trans);
  double absmax(const double* x, int n);
  double dot(const double* x, const double* y, int n);
  ~Ctx();
};

}  // namespace kad
