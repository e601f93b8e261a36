// MFMA GEMM for the trailing update of the blocked LU (the ScaLAPACK pdgesv
// replacement, reference call site src_par/do_newton.cpp:317-320).
//
//   OP_SUB : C -= A * B      (trailing update A22 -= L21 * U12)
//   OP_SET : C  = A * B      (triangular solve through an explicit inverse)
//
// All operands column-major with leading dimensions, as the dense Jacobian is
// stored (the reference packs J column-major for pdgesv, do_newton.cpp:178-185).
//
// Tile: 128x128 per workgroup, 4 waves in a 2x2 grid, each wave 64x64 = 4x4
// MFMA 16x16 tiles.  K is staged through LDS with register double buffering.
// The MFMA is issued with the roles of A and B swapped (D' = B^T A^T = C^T), so
// lanes 0..15 of an accumulator register hold 16 consecutive ROWS of one column
// of C: the column-major epilogue is coalesced.
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include <type_traits>

namespace kad {

constexpr int GBM = 128, GBN = 128, GNT = 256;
constexpr int OP_SUB = 0, OP_SET = 1;

template <int P> struct GemmCfg;
template <> struct GemmCfg<F64> { using T = double; static constexpr int BK = 16; };
template <> struct GemmCfg<F32> { using T = float;  static constexpr int BK = 32; };
template <> struct GemmCfg<BF16> { using T = float; static constexpr int BK = 64; };

// Bijective block-id remap: consecutive tiles of the grouped order land on the
// same XCD (hardware dispatches workgroups round-robin over the 8 XCDs), so the
// A row-panels and B column-panels they share stay in that XCD's L2.
__device__ inline void tile_coords(int bid, int nbm, int nbn, int& tm, int& tn) {
  const int nblocks = nbm * nbn;
  const int full = (nblocks / 8) * 8;
  int id = bid;
  if (bid < full) id = (bid % 8) * (full / 8) + bid / 8;
  const int GROUP_M = 8;
  const int in_group = GROUP_M * nbn;
  const int group = id / in_group;
  const int first_m = group * GROUP_M;
  const int gsize = min(nbm - first_m, GROUP_M);
  const int r = id % in_group;
  tm = first_m + r % gsize;
  tn = r / gsize;
}

// Tile of this workgroup under a TileSel; false: nothing to compute (the caller
// zeroes the tile for C = AB in mask mode, returns otherwise).
// Masks and lists are per 128-row / 128-column tile; a kernel with BM-wide tiles
// (BM = 64 or 128) maps its tile to the 128-tile containing it (lists: BM = 128 only).
template <int BM = GBM>
__device__ inline bool select_tile(const TileSel& sel, int nbm, int nbn, int& tm, int& tn, bool& zero_fill) {
  zero_fill = false;
  if (sel.cnt) {
    const int nr = sel.cnt[0], nc = sel.cnt[1];
    if ((int)blockIdx.x >= nr * nc) return false;
    int i, j;
    tile_coords(blockIdx.x, nr, nc, i, j);
    tm = sel.rlist[i];
    tn = sel.clist[j];
    return true;
  }
  tile_coords(blockIdx.x, nbm, nbn, tm, tn);
  if (sel.zcols && tn * BM >= sel.zcols[blockIdx.z]) return false;
  if (sel.zstart) {  // the tile's 64-column groups (tn * BM < n <= 64 zng)
    const int* st = sel.zstart + (int64_t)blockIdx.z * sel.zng;
    const int g0 = tn * BM / 64;
    int m0 = st[g0];
    if (BM > 64 && g0 + 1 < sel.zng) m0 = min(m0, st[g0 + 1]);
    if (m0 >= sel.zthr) return false;
  }
  if ((sel.rmask && sel.rmask[tm * BM / 128] == 0) || (sel.cmask && sel.cmask[tn * BM / 128] == 0)) {
    zero_fill = true;
    return false;
  }
  return true;
}
template <typename T, int BM = GBM>
__device__ inline void zero_tile(T* C, int64_t ldc, int m, int n, int row0, int col0, int tid) {
  for (int e = tid; e < BM * BM; e += GNT) {
    const int r = row0 + e % BM, c = col0 + e / BM;
    if (r < m && c < n) C[r + (int64_t)c * ldc] = T(0);
  }
}

// ---------------------------------------------------------------------------
// f64 / f32 path: 16x16x4 MFMA, LDS tiles stored k-major: As[k][m], Bs[k][n].
// ---------------------------------------------------------------------------
// BM x BM tile per workgroup (BM = 128: the bulk updates; BM = 64: small and narrow
// products -- 4x the workgroups, and at most ~100 VGPRs and 35 KB of LDS, so they run
// beside the bulk update's 128-tiles on the same CUs instead of queueing behind them).
template <int P, int OP, int BM = GBM>
__global__ __launch_bounds__(GNT) void gemm_k4_kernel(int m, int n, int k,
                                                      const typename GemmCfg<P>::T* __restrict__ A, int64_t lda,
                                                      const typename GemmCfg<P>::T* __restrict__ B, int64_t ldb,
                                                      typename GemmCfg<P>::T* __restrict__ C, int64_t ldc,
                                                      TileSel sel) {
  using T = typename GemmCfg<P>::T;
  // the small products (panel recursion, triangular-solve updates) lie on the
  // factorisation's critical path: issue priority over the bulk update's waves
  if constexpr (BM == 64) __builtin_amdgcn_s_setprio(2);
  constexpr int BK = GemmCfg<P>::BK;
  constexpr int PADA = (sizeof(T) == 8) ? 4 : 8;
  constexpr int LDA_S = BM + PADA, LDB_S = BM + PADA;
  constexpr int NA = BM * BK / GNT;  // elements of A per thread per stage
  constexpr int NB = BM * BK / GNT;
  constexpr int WT = BM / 2, TI = WT / 16;  // wave tile and its 16x16 MFMA tiles per side
  using acc_t = typename std::conditional<sizeof(T) == 8, d4v, f4v>::type;
  A += (int64_t)blockIdx.z * sel.sa;
  B += (int64_t)blockIdx.z * sel.sb;
  C += (int64_t)blockIdx.z * sel.sc;

  __shared__ T As[2][BK * LDA_S];
  __shared__ T Bs[2][BK * LDB_S];

  const int nbm = (int)cdiv(m, BM), nbn = (int)cdiv(n, BM);
  int tm = 0, tn = 0;
  bool zero_fill;
  const bool active = select_tile<BM>(sel, nbm, nbn, tm, tn, zero_fill);
  const int row0 = tm * BM, col0 = tn * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * WT, wn = (wave >> 1) * WT;
  if (!active) {
    if (OP == OP_SET && zero_fill) zero_tile<T, BM>(C, ldc, m, n, row0, col0, tid);
    return;
  }

  acc_t acc[TI][TI];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TI; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // The k loop, with bounds checks on the loads only for tiles at the matrix's edge
  // (EDGE) or a k that is not a whole number of steps.  Inside the matrix the loop is
  // one basic block, so the accumulators stay in the matrix-core registers across it
  // (guarded loads split the loop, and the compiler then copied every accumulator out
  // and back at each step, draining the MFMA pipeline).
  auto k_loop = [&](auto edge_tag) {
    constexpr bool EDGE = decltype(edge_tag)::value;
    T ra[NA], rb[NB];
    auto gload = [&](int k0) {
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        const int idx = tid + e * GNT;
        const int r = idx % BM, kk = idx / BM;
        const int gr = row0 + r, gk = k0 + kk;
        if constexpr (EDGE) {
          ra[e] = (gr < m && gk < k) ? A[gr + (int64_t)gk * lda] : T(0);
        } else {
          // a wave-uniform column (idx / BM is the same for the wave's lanes) on a
          // scalar base, the row as the lane's 32-bit offset: one address register
          const int kc = min(k0 + __builtin_amdgcn_readfirstlane(kk), k - 1);  // clamped: the last step's prefetch
          const T* col = A + row0 + (int64_t)kc * lda;
          ra[e] = col[r];
        }
      }
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        const int idx = tid + e * GNT;
        const int kk = idx % BK, c = idx / BK;
        const int gk = k0 + kk, gc = col0 + c;
        if constexpr (EDGE) {
          rb[e] = (gk < k && gc < n) ? B[gk + (int64_t)gc * ldb] : T(0);
        } else {
          // scalar base for the element's column group, lane offset kk + (c - c_e) ldb
          const int k0c = min(k0, k - BK);  // the last step's prefetch re-reads the last step
          const int ce = col0 + e * (GNT / BK);
          const T* base = B + k0c + (int64_t)ce * ldb;
          rb[e] = base[kk + (c - e * (GNT / BK)) * (int)ldb];
        }
      }
    };
    auto sstore = [&](int buf) {
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        const int idx = tid + e * GNT;
        As[buf][(idx / BM) * LDA_S + (idx % BM)] = ra[e];
      }
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        const int idx = tid + e * GNT;
        Bs[buf][(idx % BK) * LDB_S + (idx / BK)] = rb[e];
      }
    };
    const int nk = (int)cdiv(k, BK);
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      // inside the matrix the last step prefetches (clamped) and stores a tile that is
      // never read: no branch in the loop
      if (!EDGE || kt + 1 < nk) gload((kt + 1) * BK);
      const T* as = As[cur];
      const T* bs = Bs[cur];
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        T af[TI], bf[TI];
#pragma unroll
        for (int t = 0; t < TI; ++t) {
          af[t] = as[kr * LDA_S + wm + t * 16 + (lane & 15)];
          bf[t] = bs[kr * LDB_S + wn + t * 16 + (lane & 15)];
        }
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TI; ++j) {
            if constexpr (sizeof(T) == 8)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
          }
      }
      if (!EDGE || kt + 1 < nk) sstore(cur ^ 1);
      __syncthreads();
    }
  };
  if (row0 + BM <= m && col0 + BM <= n && k % BK == 0)
    k_loop(std::false_type{});
  else
    k_loop(std::true_type{});

  // Epilogue: lane holds row (lane&15) of tile i and 4 columns of tile j:
  // f32 16x16x4 accumulators are D'[4*(lane>>4)+r][lane&15], f64 16x16x4 ones
  // D'[4*r+(lane>>4)][lane&15] (measured lane layouts of the two instructions).
  // Inside the matrix there are no guards: C -= AB then reads half of the lane's C
  // entries in one batch (a register array) before it writes them back.  With a guard
  // around each entry every read of C was its own basic block, waited for alone: 64
  // memory round trips per tile, as long as the k = 256 loop itself.
  if (row0 + BM <= m && col0 + BM <= n) {
#pragma unroll
    for (int h = 0; h < TI; h += TI / 2) {
      T cv[TI / 2][TI][4];
      if constexpr (OP == OP_SUB) {
#pragma unroll
        for (int i = 0; i < TI / 2; ++i)
#pragma unroll
          for (int j = 0; j < TI; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int gr = row0 + wm + (h + i) * 16 + (lane & 15);
              const int gc = col0 + wn + j * 16 + (sizeof(T) == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r);
              cv[i][j][r] = C[gr + (int64_t)gc * ldc];
            }
      }
#pragma unroll
      for (int i = 0; i < TI / 2; ++i)
#pragma unroll
        for (int j = 0; j < TI; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int gr = row0 + wm + (h + i) * 16 + (lane & 15);
            const int gc = col0 + wn + j * 16 + (sizeof(T) == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r);
            C[gr + (int64_t)gc * ldc] = OP == OP_SUB ? cv[i][j][r] - acc[h + i][j][r] : acc[h + i][j][r];
          }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int gr = row0 + wm + i * 16 + (lane & 15);
    if (gr >= m) continue;
#pragma unroll
    for (int j = 0; j < TI; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = col0 + wn + j * 16 + (sizeof(T) == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r);
        if (gc < n) {
          T* p = C + gr + (int64_t)gc * ldc;
          if (OP == OP_SUB) *p -= acc[i][j][r];
          else *p = acc[i][j][r];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// bf16 path: fp32 storage, operands rounded to bf16 in the LDS store,
// 16x16x32 MFMA with fp32 accumulation.  LDS tiles are k-contiguous:
// As[m][k], Bs[n][k], so each lane's 8 consecutive k values are one 16-byte read.
// ---------------------------------------------------------------------------
template <int OP>
__global__ __launch_bounds__(GNT) void gemm_bf16_kernel(int m, int n, int k, const float* __restrict__ A, int64_t lda,
                                                        const float* __restrict__ B, int64_t ldb, float* __restrict__ C,
                                                        int64_t ldc, TileSel sel) {
  constexpr int BK = GemmCfg<BF16>::BK;
  constexpr int LDK = BK + 8;  // bf16 elements per LDS row (pad 16 bytes)
  constexpr int NA = GBM * BK / GNT, NB = GBN * BK / GNT;
  A += (int64_t)blockIdx.z * sel.sa;
  B += (int64_t)blockIdx.z * sel.sb;
  C += (int64_t)blockIdx.z * sel.sc;
  __shared__ __bf16 As[2][GBM * LDK];
  __shared__ __bf16 Bs[2][GBN * LDK];

  const int nbm = (int)cdiv(m, GBM), nbn = (int)cdiv(n, GBN);
  int tm = 0, tn = 0;
  bool zero_fill;
  const bool active = select_tile(sel, nbm, nbn, tm, tn, zero_fill);
  const int row0 = tm * GBM, col0 = tn * GBN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  if (!active) {
    if (OP == OP_SET && zero_fill) zero_tile(C, ldc, m, n, row0, col0, tid);
    return;
  }

  f4v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0, 0, 0, 0};

  float ra[NA], rb[NB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + e * GNT;
      const int r = idx % GBM, kk = idx / GBM;
      const int gr = row0 + r, gk = k0 + kk;
      ra[e] = (gr < m && gk < k) ? A[gr + (int64_t)gk * lda] : 0.f;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = tid + e * GNT;
      const int kk = idx % BK, c = idx / BK;
      const int gk = k0 + kk, gc = col0 + c;
      rb[e] = (gk < k && gc < n) ? B[gk + (int64_t)gc * ldb] : 0.f;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + e * GNT;
      As[buf][(idx % GBM) * LDK + idx / GBM] = (__bf16)ra[e];
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = tid + e * GNT;
      Bs[buf][(idx / BK) * LDK + idx % BK] = (__bf16)rb[e];
    }
  };

  const int nk = (int)cdiv(k, BK);
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    // the last step loads (clamped, in bounds) and stores a tile nobody reads
    gload((kt + 1) * BK);
    const __bf16* as = As[cur];
    const __bf16* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      const int kof = kk + 8 * (lane >> 4);
      bf8v af[4], bf[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        af[t] = *reinterpret_cast<const bf8v*>(as + (wm + t * 16 + (lane & 15)) * LDK + kof);
        bf[t] = *reinterpret_cast<const bf8v*>(bs + (wn + t * 16 + (lane & 15)) * LDK + kof);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gr = row0 + wm + i * 16 + (lane & 15);
    if (gr >= m) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gc = col0 + wn + j * 16 + 4 * (lane >> 4) + r;
        if (gc < n) {
          float* p = C + gr + (int64_t)gc * ldc;
          if (OP == OP_SUB) *p -= acc[i][j][r];
          else *p = acc[i][j][r];
        }
      }
  }
}

// Fraction of the (nbm x nbn) tile pairs of C that a launch with `sel` executes: the
// list counts, or the product of the active fractions of the row and column masks.
// One block; timing runs only.
__global__ void tile_fraction_kernel(TileSel sel, int nbm, int nbn, double* out) {
  __shared__ int cr, cc;
  if (sel.cnt) {
    if (threadIdx.x == 0) *out = ((double)sel.cnt[0] / nbm) * ((double)sel.cnt[1] / nbn);
    return;
  }
  if (sel.zcols || sel.zstart) {  // no masks with a batch: the column tiles of each matrix
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int z = 0; z < sel.nbat; ++z)
        for (int q = 0; q < nbn; ++q) {
          bool on = !sel.zcols || q * GBN < sel.zcols[z];
          if (on && sel.zstart) {
            const int* st = sel.zstart + (int64_t)z * sel.zng;
            int m0 = st[2 * q];
            if (2 * q + 1 < sel.zng) m0 = min(m0, st[2 * q + 1]);
            on = m0 < sel.zthr;
          }
          t += on;
        }
      *out = t / ((double)nbn * sel.nbat);
    }
    return;
  }
  if (threadIdx.x == 0) {
    cr = 0;
    cc = 0;
  }
  __syncthreads();
  int r = 0, c = 0;
  for (int q = threadIdx.x; q < nbm; q += blockDim.x) r += sel.rmask ? (sel.rmask[q] != 0) : 1;
  for (int q = threadIdx.x; q < nbn; q += blockDim.x) c += sel.cmask ? (sel.cmask[q] != 0) : 1;
  atomicAdd(&cr, r);
  atomicAdd(&cc, c);
  __syncthreads();
  if (threadIdx.x == 0) *out = ((double)cr / nbm) * ((double)cc / nbn);
}

// Host launcher: C (-)= A*B for the given precision.  Returns KAD_OK or an error.
int gemm_launch_sel(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                    void* C, int64_t ldc, const TileSel& sel, hipStream_t s);

int gemm_launch(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                int64_t ldc, hipStream_t s) {
  return gemm_launch_sel(prec, op, m, n, k, A, lda, B, ldb, C, ldc, TileSel(), s);
}

int gemm_launch_batched(int prec, int op, int m, int n, int k, const void* A, int64_t lda, int64_t sa, const void* B,
                        int64_t ldb, int64_t sb, void* C, int64_t ldc, int64_t sc, int nbat, hipStream_t s) {
  TileSel sel;
  sel.nbat = nbat;
  sel.sa = sa;
  sel.sb = sb;
  sel.sc = sc;
  return gemm_launch_sel(prec, op, m, n, k, A, lda, B, ldb, C, ldc, sel, s);
}

int gemm_launch_masked(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                       void* C, int64_t ldc, const int* rmask, const int* cmask, hipStream_t s) {
  TileSel sel;
  sel.rmask = rmask;
  sel.cmask = cmask;
  return gemm_launch_sel(prec, op, m, n, k, A, lda, B, ldb, C, ldc, sel, s);
}

// As gemm_launch, computing only the C tiles that sel selects (list mode: OP_SUB only;
// masks and lists are device arrays over 128-row / 128-column tiles of C).
int gemm_launch_sel(int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                    void* C, int64_t ldc, const TileSel& sel, hipStream_t s) {
  if ((sel.cnt || sel.zcols || sel.zstart) && op != OP_SUB) return KAD_ERR_ARG;
  if (sel.zstart && (int64_t)sel.zng * 64 < n) return KAD_ERR_ARG;
  if (sel.nbat < 1 || sel.nbat > 65535 || (sel.nbat > 1 && sel.cnt)) {
    fprintf(stderr, "[kadath_amd] gemm: batch of %d with tile lists\n", sel.nbat);
    return KAD_ERR_ARG;
  }
  if (m <= 0 || n <= 0) return KAD_OK;
  if (k <= 0) {
    if (op == OP_SET) {
      const size_t es = (prec == F64) ? 8 : 4;
      for (int z = 0; z < sel.nbat; ++z)
        KAD_HIP_CHECK(hipMemset2DAsync((char*)C + (size_t)z * sel.sc * es, ldc * es, 0, (size_t)m * es, n, s));
    }
    return KAD_OK;
  }
  const int nblocks = (int)(cdiv(m, GBM) * cdiv(n, GBN));
  // fewer 128-tiles than two per CU (and not a compacted tile list): 64-tiles
  const bool small = !sel.cnt && prec != BF16 && (int64_t)nblocks * sel.nbat < 512;
  dim3 grid(small ? (unsigned)(cdiv(m, 64) * cdiv(n, 64)) : (unsigned)nblocks, 1, (unsigned)sel.nbat), block(GNT);
  GemmTimer* tm = g_gemm_timer;
  bool timed = false;
  if (tm) {
    // executed work: the device counts the tile pairs that are not skipped, just before
    // the launch on the same stream (outside the timed bracket, no host synchronisation)
    const int nbm = (int)cdiv(m, GBM), nbn = (int)cdiv(n, GBN);
    double* slot = tm->scale_slot();
    const bool masked = sel.cnt || sel.rmask || sel.cmask || sel.zcols || sel.zstart;
    if (slot && masked) tile_fraction_kernel<<<1, 256, 0, s>>>(sel, nbm, nbn, slot);
    const int rc = tm->begin(s, 2.0 * m * n * (double)k * sel.nbat, masked);
    if (rc < 0) return rc;
    timed = rc == 0;
  }
  if (prec == F64) {
    const double *a = (const double*)A, *b = (const double*)B;
    double* c = (double*)C;
    if (op == OP_SUB) {
      if (small)
        gemm_k4_kernel<F64, OP_SUB, 64><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
      else
        gemm_k4_kernel<F64, OP_SUB><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
    } else {
      if (small)
        gemm_k4_kernel<F64, OP_SET, 64><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
      else
        gemm_k4_kernel<F64, OP_SET><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
    }
  } else if (prec == F32) {
    const float *a = (const float*)A, *b = (const float*)B;
    float* c = (float*)C;
    if (op == OP_SUB) {
      if (small)
        gemm_k4_kernel<F32, OP_SUB, 64><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
      else
        gemm_k4_kernel<F32, OP_SUB><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
    } else {
      if (small)
        gemm_k4_kernel<F32, OP_SET, 64><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
      else
        gemm_k4_kernel<F32, OP_SET><<<grid, block, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, sel);
    }
  } else if (prec == BF16) {
    if (op == OP_SUB)
      gemm_bf16_kernel<OP_SUB><<<grid, block, 0, s>>>(m, n, k, (const float*)A, lda, (const float*)B, ldb, (float*)C,
                                                      ldc, sel);
    else
      gemm_bf16_kernel<OP_SET><<<grid, block, 0, s>>>(m, n, k, (const float*)A, lda, (const float*)B, ldb, (float*)C,
                                                      ldc, sel);
  } else {
    return KAD_ERR_ARG;
  }
  KAD_HIP_CHECK(hipGetLastError());
  if (timed && tm->end(s)) return KAD_ERR_HIP;
  return KAD_OK;
}

}  // namespace kad
