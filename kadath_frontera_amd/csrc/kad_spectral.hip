// Batched spectral transforms and 1-D coefficient-space operators.
//
// Replaces Base_spectral::coef / coef_i (src/Coef/coef.cpp:28-104,
// src/Coef/coef_i.cpp:28-104) and Base_spectral::ope_1d
// (include/implementation/ope_1d.cpp:25-82), which gather each 1-D line of a
// Val_domain, call FFTW (or a coefficient recurrence) on it and scatter back.
//
// Here every 1-D operation on a line of length <= 64 is a small dense matrix
// (chosen per line by the line's basis), and a whole 3-D domain array
// (<= ~33x33x34 doubles) is staged in LDS: one coalesced HBM read, any number of
// 1-D passes along r, theta, phi in LDS, one coalesced HBM write.
//
// Array layout: index 0 (r) fastest, then theta, then phi -- the reference's
// Array/Index order (include/implementation/index.cpp:84-96).
#include "kad_common.hpp"
#include "kad_spectral.hpp"
#include "kad_ctx.hpp"

namespace kad {

// One pass along any axis with NIN (12..17) input points per line, one line per
// thread: the thread reads its line into registers once and produces all of the
// line's outputs; the line's matrix is read with wave-uniform (scalar) loads from
// the read-only matrix pool in global memory, so each FMA takes one operand from
// SGPRs and one from VGPRs and LDS carries only the line in and the outputs out.
// Waves whose lines use different matrices (e.g. theta lines of even and odd phi
// index) are served one matrix at a time (a waterfall over the distinct offsets).
// x and y may be the same buffer when the pass keeps the line length: the thread
// reads its whole line before writing the line's outputs, and lines are disjoint.
template <int NIN>
__device__ __forceinline__ void pass_lines(const double* x, double* y,
                                           const double* __restrict__ pool, const int* __restrict__ lm, int ax,
                                           const int dims[3], const int od[3], int nout_ax) {
  const int nl0 = ax == 0 ? od[1] : od[0];
  const int nlines = ax == 0 ? od[1] * od[2] : (ax == 1 ? od[0] * od[2] : od[0] * od[1]);
  const int sin_ = ax == 0 ? 1 : (ax == 1 ? dims[0] : dims[0] * dims[1]);
  const int sout = ax == 0 ? 1 : (ax == 1 ? od[0] : od[0] * od[1]);
  for (int li = threadIdx.x; li - (int)(threadIdx.x & 63) < nlines; li += blockDim.x) {
    // every lane of a wave runs the same iterations (the vote below needs them all)
    const bool act = li < nlines;
    const int a = act ? li % nl0 : 0, bq = act ? li / nl0 : 0;
    int xb, yb, lmi;
    if (ax == 0) {         // line (j, k) = (a, bq)
      xb = dims[0] * (a + dims[1] * bq);
      yb = od[0] * (a + od[1] * bq);
      lmi = a + od[1] * bq;
    } else if (ax == 1) {  // line (i, k) = (a, bq)
      xb = a + dims[0] * dims[1] * bq;
      yb = a + od[0] * od[1] * bq;
      lmi = bq;
    } else {               // line (i, j) = (a, bq)
      xb = a + dims[0] * bq;
      yb = a + od[0] * bq;
      lmi = 0;
    }
    const int off = act ? lm[lmi] : -1;
    double xv[NIN];
#pragma unroll
    for (int q = 0; q < NIN; ++q) xv[q] = act ? x[xb + q * sin_] : 0.0;
    // waterfall over the distinct matrices of the wave, in uniform control flow (so
    // the matrix address is uniform and its loads are scalar): every lane computes
    // with the chosen matrix, the lanes whose line uses it store
    bool todo = act;
    while (true) {
      const unsigned long long pend = __ballot(todo);
      if (pend == 0) break;
      const int off0 = __builtin_amdgcn_readlane(off, __ffsll((long long)pend) - 1);
      const bool mine = todo && off == off0;
      todo = todo && !mine;
      if (off0 < 0) {
        for (int o = 0; o < nout_ax; ++o)
          if (mine) y[yb + o * sout] = 0.0;
        continue;
      }
      // the pool is read-only for the whole kernel: read it through the constant
      // address space, so the uniform address gives scalar loads
      typedef const double __attribute__((address_space(4)))* cdptr;
      const cdptr M = (cdptr)(pool + off0);
#pragma unroll 2
      for (int o = 0; o < nout_ax; ++o) {
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int q = 0; q + 1 < NIN; q += 2) {
          s0 += M[o * NIN + q] * xv[q];
          s1 += M[o * NIN + q + 1] * xv[q + 1];
        }
        if (NIN % 2) s0 += M[o * NIN + NIN - 1] * xv[NIN - 1];
        if (mine) y[yb + o * sout] = s0 + s1;
      }
    }
  }
}

// One workgroup per batch element.  Dynamic LDS: two ping-pong buffers, then the
// plan's matrix pool and per-line matrix tables (a few KB, staged once), so the
// contraction of every output element reads only LDS.
__global__ __launch_bounds__(256) void apply3d_kernel(const double* __restrict__ in, int64_t in_stride,
                                                      double* __restrict__ out, int64_t out_stride, int d0, int d1,
                                                      int d2, Passes P, const double* __restrict__ pool,
                                                      int accumulate, double scale, int stage_pool, int inplace,
                                                      PostOp post) {
  extern __shared__ double lds[];
  // the two ping-pong buffers are selected by value (not through an array of
  // pointers), so every access stays an LDS access and the matrix-pool loads can be
  // proven unclobbered (scalar loads); inplace: every pass keeps its line length and
  // runs line by line, so one buffer serves as input and output (half the LDS)
  double* const buf0 = lds;
  double* const buf1 = inplace ? lds : lds + P.maxsize;
  double* mats = lds + (inplace ? 1 : 2) * P.maxsize;
  int* ltab = reinterpret_cast<int*>(mats + (stage_pool ? P.pool_size : 0));
  const int64_t b = blockIdx.x;
  const int nin = d0 * d1 * d2;
  const double* src = in + b * in_stride;
  // 8 loads in flight per thread (one load per iteration would wait out the HBM
  // latency ~10 times per array)
  constexpr int UL = 8;
  for (int e0 = 0; e0 < nin; e0 += UL * 256) {
    double v[UL];
#pragma unroll
    for (int u = 0; u < UL; ++u) {
      const int e = e0 + u * 256 + (int)threadIdx.x;
      v[u] = e < nin ? src[e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UL; ++u) {
      const int e = e0 + u * 256 + (int)threadIdx.x;
      if (e < nin) buf0[e] = v[u];
    }
  }
  if (stage_pool)  // only passes outside 12..17 points per line read the matrices from LDS
    for (int e = threadIdx.x; e < P.pool_size; e += blockDim.x) mats[e] = pool[e];
  {
    int o = 0;
    for (int p = 0; p < P.npass; ++p) {
      for (int e = threadIdx.x; e < P.nlines[p]; e += blockDim.x) ltab[o + e] = P.line_mat[p][e];
      o += P.nlines[p];
    }
  }
  __syncthreads();
  int cur = 0;
  int dims[3] = {d0, d1, d2};
  int loff = 0;
  for (int p = 0; p < P.npass; ++p) {
    const int ax = P.axis[p];
    const int nout_ax = P.nout[p];
    const int nin_ax = dims[ax];
    int od[3] = {dims[0], dims[1], dims[2]};
    od[ax] = nout_ax;
    const int nout = od[0] * od[1] * od[2];
    const double* x = cur ? buf1 : buf0;
    double* y = cur ? buf0 : buf1;
    const int* lm = ltab + loff;
    const int stride = ax == 0 ? 1 : (ax == 1 ? dims[0] : dims[0] * dims[1]);
    if (nin_ax >= 12 && nin_ax <= 17) {
      switch (nin_ax) {
        case 12: pass_lines<12>(x, y, pool, lm, ax, dims, od, nout_ax); break;
        case 13: pass_lines<13>(x, y, pool, lm, ax, dims, od, nout_ax); break;
        case 14: pass_lines<14>(x, y, pool, lm, ax, dims, od, nout_ax); break;
        case 15: pass_lines<15>(x, y, pool, lm, ax, dims, od, nout_ax); break;
        case 16: pass_lines<16>(x, y, pool, lm, ax, dims, od, nout_ax); break;
        default: pass_lines<17>(x, y, pool, lm, ax, dims, od, nout_ax); break;
      }
      loff += P.nlines[p];
      dims[0] = od[0];
      dims[1] = od[1];
      dims[2] = od[2];
      cur ^= 1;
      __syncthreads();
      continue;
    }
    for (int e = threadIdx.x; e < nout; e += blockDim.x) {
      const int i = e % od[0];
      const int t = e / od[0];
      const int j = t % od[1];
      const int k = t / od[1];
      int line, o, base_idx;
      if (ax == 0) {
        line = j + od[1] * k;
        o = i;
        base_idx = dims[0] * (j + dims[1] * k);
      } else if (ax == 1) {
        line = k;
        o = j;
        base_idx = i + dims[0] * dims[1] * k;
      } else {
        line = 0;
        o = k;
        base_idx = i + dims[0] * j;
      }
      const int off = lm[line];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      if (off >= 0) {
        const double* M = mats + off + o * nin_ax;
        const double* xv = x + base_idx;
        int q = 0;
        for (; q + 4 <= nin_ax; q += 4) {
          s0 += M[q] * xv[q * stride];
          s1 += M[q + 1] * xv[(q + 1) * stride];
          s2 += M[q + 2] * xv[(q + 2) * stride];
          s3 += M[q + 3] * xv[(q + 3) * stride];
        }
        for (; q < nin_ax; ++q) s0 += M[q] * xv[q * stride];
      }
      y[e] = (s0 + s1) + (s2 + s3);
    }
    loff += P.nlines[p];
    dims[0] = od[0];
    dims[1] = od[1];
    dims[2] = od[2];
    cur ^= 1;
    __syncthreads();
  }
  const int nfin = dims[0] * dims[1] * dims[2];
  double* dst = out + b * out_stride;
  const double* fin = cur ? buf1 : buf0;
  if (accumulate) {
    for (int e0 = 0; e0 < nfin; e0 += UL * 256) {
      double v[UL];
#pragma unroll
      for (int u = 0; u < UL; ++u) {
        const int e = e0 + u * 256 + (int)threadIdx.x;
        v[u] = e < nfin ? dst[e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < UL; ++u) {
        const int e = e0 + u * 256 + (int)threadIdx.x;
        if (e < nfin) dst[e] = v[u] + scale * fin[e];
      }
    }
  } else if (post.n > 0) {
    // the element-wise instructions fed by this pass (scale is 1 here); the other
    // operands of UL elements per thread are loaded together, as the inputs were
    double* drow = post.dst + b * post.ds;
    for (int e0 = 0; e0 < nfin; e0 += UL * 256) {
      double o[UL][kMaxChain];
#pragma unroll
      for (int u = 0; u < UL; ++u) {
        const int e = e0 + u * 256 + (int)threadIdx.x;
        post_load(post, b, e < nfin ? e : 0, o[u]);
      }
#pragma unroll
      for (int u = 0; u < UL; ++u) {
        const int e = e0 + u * 256 + (int)threadIdx.x;
        if (e < nfin) drow[e] = post_ops(post, fin[e], o[u]);
      }
    }
  } else {
    for (int e = threadIdx.x; e < nfin; e += blockDim.x) dst[e] = scale * fin[e];
  }
}

// Contiguous 1-D lines: in [L][nin] -> out [L][nout], one matrix for all lines.
// A workgroup stages LINES_PER_WG lines in LDS (coalesced) together with M.
constexpr int LINES_PER_WG = 64;
__global__ __launch_bounds__(256) void apply_lines_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                          int64_t L, int nin, int nout, const double* __restrict__ M) {
  extern __shared__ double lds[];
  double* Ms = lds;                    // nout*nin
  double* xs = lds + nout * nin;       // LINES_PER_WG*nin
  const int64_t l0 = (int64_t)blockIdx.x * LINES_PER_WG;
  const int nl = (int)min<int64_t>(LINES_PER_WG, L - l0);
  for (int e = threadIdx.x; e < nout * nin; e += blockDim.x) Ms[e] = M[e];
  const double* src = in + l0 * nin;
  for (int e = threadIdx.x; e < nl * nin; e += blockDim.x) xs[e] = src[e];
  __syncthreads();
  double* dst = out + l0 * nout;
  for (int e = threadIdx.x; e < nl * nout; e += blockDim.x) {
    const int l = e / nout, o = e % nout;
    const double* m = Ms + o * nin;
    const double* x = xs + l * nin;
    double s = 0.0;
    for (int q = 0; q < nin; ++q) s += m[q] * x[q];
    dst[e] = s;
  }
}

int apply3d_launch(const double* in, int64_t in_stride, double* out, int64_t out_stride, int nbatch, int d0, int d1,
                   int d2, const Passes& P, const double* pool, int accumulate, double scale, hipStream_t s,
                   const PostOp* post) {
  if (nbatch <= 0) return KAD_OK;
  if (post && post->n > 0 && (post->n > kMaxChain || accumulate || scale != 1.0 || !post->dst)) return KAD_ERR_ARG;
  const PostOp pop = post ? *post : PostOp();
  size_t nl = 0;
  for (int p = 0; p < P.npass; ++p) nl += (size_t)P.nlines[p];
  // the matrix pool is staged in LDS only if some pass has a line length outside
  // 12..17 (the other passes read it with scalar loads)
  int stage_pool = 0, square = 1;
  {
    int dm[3] = {d0, d1, d2};
    for (int p = 0; p < P.npass; ++p) {
      const int nin_ax = dm[P.axis[p]];
      if (nin_ax < 12 || nin_ax > 17) stage_pool = 1;
      if (P.nout[p] != nin_ax) square = 0;
      dm[P.axis[p]] = P.nout[p];
    }
  }
  // one LDS buffer when every pass is a line-by-line pass that keeps its line length
  const int inplace = (!stage_pool && square) ? 1 : 0;
  const size_t lds = ((inplace ? 1 : 2) * (size_t)P.maxsize + (stage_pool ? (size_t)P.pool_size : 0)) *
                         sizeof(double) + nl * sizeof(int);
  if (lds > 160 * 1024) return KAD_ERR_ARG;
  GemmTimer* tm = g_apply_timer;
  bool timed = false;
  if (tm) {
    // algorithmic HBM bytes: the input arrays read once, the output arrays written
    // once (accumulate also reads them); the per-line matrices are cache-resident
    int od[3] = {d0, d1, d2};
    for (int p = 0; p < P.npass; ++p) od[P.axis[p]] = P.nout[p];
    const double nin = (double)d0 * d1 * d2, nout = (double)od[0] * od[1] * od[2];
    // (each fused instruction's other operand adds its array per batch element)
    double extra = 0.0;
    for (int k = 0; k < pop.n; ++k)
      if (pop.op[k].o && pop.op[k].os != 0) extra += nout;
    const int rc = tm->begin(s, 8.0 * nbatch * (nin + nout * (accumulate ? 2 : 1) + extra));
    if (rc < 0) return rc;
    timed = rc == 0;
  }
  apply3d_kernel<<<nbatch, 256, lds, s>>>(in, in_stride, out, out_stride, d0, d1, d2, P, pool, accumulate, scale,
                                          stage_pool, inplace, pop);
  KAD_HIP_CHECK(hipGetLastError());
  if (timed && tm->end(s)) return KAD_ERR_HIP;
  return KAD_OK;
}

int apply_lines_launch(const double* in, double* out, int64_t L, int nin, int nout, const double* M, hipStream_t s) {
  if (L <= 0) return KAD_OK;
  const size_t lds = ((size_t)nout * nin + (size_t)LINES_PER_WG * nin) * sizeof(double);
  if (lds > 160 * 1024) return KAD_ERR_ARG;
  apply_lines_kernel<<<(unsigned)cdiv(L, LINES_PER_WG), 256, lds, s>>>(in, out, L, nin, nout, M);
  KAD_HIP_CHECK(hipGetLastError());
  return KAD_OK;
}

}  // namespace kad
