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

namespace kad {

// One workgroup per batch element.  Dynamic LDS: two ping-pong buffers.
__global__ __launch_bounds__(256) void apply3d_kernel(const double* __restrict__ in, int64_t in_stride,
                                                      double* __restrict__ out, int64_t out_stride, int d0, int d1,
                                                      int d2, Passes P, const double* __restrict__ pool,
                                                      int accumulate, double scale) {
  extern __shared__ double lds[];
  double* buf[2] = {lds, lds + P.maxsize};
  const int64_t b = blockIdx.x;
  const int nin = d0 * d1 * d2;
  const double* src = in + b * in_stride;
  for (int e = threadIdx.x; e < nin; e += blockDim.x) buf[0][e] = src[e];
  __syncthreads();
  int cur = 0;
  int dims[3] = {d0, d1, d2};
  for (int p = 0; p < P.npass; ++p) {
    const int ax = P.axis[p];
    const int nout_ax = P.nout[p];
    const int nin_ax = dims[ax];
    int od[3] = {dims[0], dims[1], dims[2]};
    od[ax] = nout_ax;
    const int nout = od[0] * od[1] * od[2];
    const double* x = buf[cur];
    double* y = buf[cur ^ 1];
    const int* lm = P.line_mat[p];
    for (int e = threadIdx.x; e < nout; e += blockDim.x) {
      const int i = e % od[0];
      const int t = e / od[0];
      const int j = t % od[1];
      const int k = t / od[1];
      int line, o, base_idx, stride;
      if (ax == 0) {
        line = j + od[1] * k;
        o = i;
        base_idx = dims[0] * (j + dims[1] * k);
        stride = 1;
      } else if (ax == 1) {
        line = k;
        o = j;
        base_idx = i + dims[0] * dims[1] * k;
        stride = dims[0];
      } else {
        line = 0;
        o = k;
        base_idx = i + dims[0] * j;
        stride = dims[0] * dims[1];
      }
      const int off = lm[line];
      double s = 0.0;
      if (off >= 0) {
        const double* M = pool + off + (int64_t)o * nin_ax;
        for (int q = 0; q < nin_ax; ++q) s += M[q] * x[base_idx + q * stride];
      }
      y[e] = s;
    }
    dims[0] = od[0];
    dims[1] = od[1];
    dims[2] = od[2];
    cur ^= 1;
    __syncthreads();
  }
  const int nfin = dims[0] * dims[1] * dims[2];
  double* dst = out + b * out_stride;
  if (accumulate)
    for (int e = threadIdx.x; e < nfin; e += blockDim.x) dst[e] += scale * buf[cur][e];
  else
    for (int e = threadIdx.x; e < nfin; e += blockDim.x) dst[e] = scale * buf[cur][e];
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
                   int d2, const Passes& P, const double* pool, int accumulate, double scale, hipStream_t s) {
  if (nbatch <= 0) return KAD_OK;
  const size_t lds = 2 * (size_t)P.maxsize * sizeof(double);
  if (lds > 160 * 1024) return KAD_ERR_ARG;
  apply3d_kernel<<<nbatch, 256, lds, s>>>(in, in_stride, out, out_stride, d0, d1, d2, P, pool, accumulate, scale);
  KAD_HIP_CHECK(hipGetLastError());
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
