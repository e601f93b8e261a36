// 1-D spectral operators of the spherical domains (host side, builds the
// per-line matrices the GPU kernels apply).  See spectral.hpp.
#include "spectral.hpp"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

namespace kadhost {

Mat matmul(const Mat& A, const Mat& B) {
  if (A.nin != B.nout) throw std::runtime_error("matmul: shape mismatch");
  Mat C(A.nout, B.nin);
  for (int i = 0; i < A.nout; ++i)
    for (int k = 0; k < A.nin; ++k) {
      const double a = A(i, k);
      if (a == 0.0) continue;
      for (int j = 0; j < B.nin; ++j) C(i, j) += a * B(k, j);
    }
  return C;
}

Mat identity(int n) {
  Mat I(n, n);
  for (int i = 0; i < n; ++i) I(i, i) = 1.0;
  return I;
}

[[noreturn]] static void not_impl(const char* what, int base) {
  char buf[128];
  snprintf(buf, sizeof buf, "%s not implemented for basis %d", what, base);
  throw std::runtime_error(buf);
}

// ---------------------------------------------------------------------------
// der_1d  (src/Ope_1d/der_1d.cpp:31-280)
// ---------------------------------------------------------------------------
int der_1d(int base, std::vector<double>& tab) {
  const int nr = (int)tab.size();
  std::vector<double> res(nr, 0.0);
  switch (base) {
    case CHEB:  // der_1d.cpp:31-48
      for (int i = 0; i < nr; ++i)
        for (int p = i + 1; p < nr; p += 2) res[i] += p * tab[p];
      for (auto& v : res) v *= 2;
      res[0] /= 2;
      tab = res;
      return CHEB;
    case COSSIN:  // der_1d.cpp:132-148
      for (int i = 0; i < nr - 1; i += 2) res[i + 1] = -(i / 2) * tab[i];
      for (int i = 1; i < nr; i += 2) res[i - 1] = ((i - 1) / 2) * tab[i];
      res[nr - 1] = 0;
      tab = res;
      return COSSIN;
    case COS_EVEN:  // der_1d.cpp:215-228
      for (int i = 0; i < nr; ++i) res[i] = -2 * i * tab[i];
      res[nr - 1] = 0;
      tab = res;
      return SIN_EVEN;
    case COS_ODD:  // der_1d.cpp:232-244
      for (int i = 0; i < nr; ++i) res[i] = -(2 * i + 1) * tab[i];
      tab = res;
      return SIN_ODD;
    case SIN_EVEN:  // der_1d.cpp:250-262
      for (int i = 0; i < nr; ++i) res[i] = 2 * i * tab[i];
      tab = res;
      return COS_EVEN;
    case SIN_ODD:  // der_1d.cpp:266-279
      for (int i = 0; i < nr; ++i) res[i] = (2 * i + 1) * tab[i];
      res[nr - 1] = 0;
      tab = res;
      return COS_ODD;
    default:
      not_impl("der_1d", base);
  }
}

// ---------------------------------------------------------------------------
// mult_sin_1d  (src/Ope_1d/mult_sin_1d.cpp:29-145)
// ---------------------------------------------------------------------------
int mult_sin_1d(int base, std::vector<double>& tab) {
  const int nr = (int)tab.size();
  std::vector<double> res(nr, 0.0);
  switch (base) {
    case COSSIN:  // :29-51
      res[0] = 0.5 * tab[3];
      res[1] = 0;
      res[2] = 0.5 * tab[5];
      res[3] = tab[0] - 0.5 * tab[4];
      for (int k = 4; k < nr - 2; k += 2) {
        res[k] = 0.5 * (tab[3 + k] - tab[k - 1]);
        res[k + 1] = 0.5 * (tab[k - 2] - tab[k + 2]);
      }
      res[nr - 2] = -0.5 * tab[nr - 3];
      res[nr - 1] = 0.;
      tab = res;
      return COSSIN;
    case COS_EVEN:  // :84-99
      res[0] = tab[0] - 0.5 * tab[1];
      for (int k = 1; k < nr - 1; k++) res[k] = 0.5 * (tab[k] - tab[k + 1]);
      res[nr - 1] = 0;
      tab = res;
      return SIN_ODD;
    case COS_ODD:  // :101-114
      res[0] = 0.;
      for (int k = 1; k < nr; k++) res[k] = 0.5 * (tab[k - 1] - tab[k]);
      tab = res;
      return SIN_EVEN;
    case SIN_EVEN:  // :116-131
      res[0] = 0.5 * tab[1];
      for (int k = 1; k < nr - 1; k++) res[k] = 0.5 * (-tab[k] + tab[k + 1]);
      res[nr - 1] = 0;
      tab = res;
      return COS_ODD;
    case SIN_ODD:  // :133-146
      res[0] = 0.5 * tab[0];
      for (int k = 1; k < nr; k++) res[k] = 0.5 * (-tab[k - 1] + tab[k]);
      tab = res;
      return COS_EVEN;
    default:
      not_impl("mult_sin_1d", base);
  }
}

// ---------------------------------------------------------------------------
// mult_cos_1d  (src/Ope_1d/mult_cos_1d.cpp:29-145)
// ---------------------------------------------------------------------------
int mult_cos_1d(int base, std::vector<double>& tab) {
  const int nr = (int)tab.size();
  std::vector<double> res(nr, 0.0);
  switch (base) {
    case COSSIN:  // :29-48
      res[0] = 0.5 * tab[2];
      res[1] = 0;
      res[2] = tab[0] + 0.5 * tab[4];
      res[3] = 0.5 * tab[5];
      for (int k = 4; k < nr - 2; k++) res[k] = 0.5 * (tab[2 + k] + tab[k - 2]);
      res[nr - 2] = 0.5 * tab[nr - 4];
      res[nr - 1] = 0.;
      tab = res;
      return COSSIN;
    case COS_EVEN:  // :84-98
      res[0] = tab[0] + 0.5 * tab[1];
      for (int k = 1; k < nr - 1; k++) res[k] = 0.5 * (tab[k] + tab[k + 1]);
      res[nr - 1] = 0.;
      tab = res;
      return COS_ODD;
    case COS_ODD:  // :100-112
      res[0] = 0.5 * tab[0];
      for (int k = 1; k < nr; k++) res[k] = 0.5 * (tab[k - 1] + tab[k]);
      tab = res;
      return COS_EVEN;
    case SIN_EVEN:  // :114-127
      res[0] = 0.5 * tab[1];
      for (int k = 1; k < nr - 1; k++) res[k] = 0.5 * (tab[k] + tab[k + 1]);
      res[nr - 1] = 0.;
      tab = res;
      return SIN_ODD;
    case SIN_ODD:  // :129-141
      res[0] = 0.;
      for (int k = 1; k < nr; k++) res[k] = 0.5 * (tab[k - 1] + tab[k]);
      tab = res;
      return SIN_EVEN;
    default:
      not_impl("mult_cos_1d", base);
  }
}

// ---------------------------------------------------------------------------
// div_sin_1d  (src/Ope_1d/div_sin_1d.cpp:75-143)
// ---------------------------------------------------------------------------
int div_sin_1d(int base, std::vector<double>& tab) {
  const int nr = (int)tab.size();
  std::vector<double> res(nr, 0.0);
  double somme = 0;
  switch (base) {
    case COS_EVEN:  // :75-88
      res[nr - 1] = 0.;
      for (int k = nr - 2; k >= 0; k--) {
        somme -= 2 * tab[k + 1];
        res[k] = somme;
      }
      tab = res;
      return SIN_ODD;
    case COS_ODD:  // :90-103
      res[nr - 1] = 0.;
      for (int k = nr - 2; k >= 0; k--) {
        somme -= 2 * tab[k];
        res[k] = somme;
      }
      tab = res;
      return SIN_EVEN;
    case SIN_EVEN:  // :105-118
      res[nr - 1] = 0.;
      for (int k = nr - 2; k >= 0; k--) {
        somme += 2 * tab[k + 1];
        res[k] = somme;
      }
      tab = res;
      return COS_ODD;
    case SIN_ODD:  // :120-134
      res[nr - 1] = 0.;
      for (int k = nr - 2; k >= 0; k--) {
        somme += 2 * tab[k];
        res[k] = somme;
      }
      res[0] *= 0.5;
      tab = res;
      return COS_EVEN;
    default:
      not_impl("div_sin_1d", base);
  }
}

// mult_x_1d (src/Ope_1d/mult_x_1d.cpp:29-41), CHEB only here.
int mult_x_1d(int base, std::vector<double>& so) {
  if (base != CHEB) not_impl("mult_x_1d", base);
  const int nr = (int)so.size();
  std::vector<double> res(nr, 0.0);
  res[0] = 0.5 * so[1];
  res[1] = so[0] + 0.5 * so[2];
  for (int i = 2; i < nr - 1; i++) res[i] = 0.5 * (so[i + 1] + so[i - 1]);
  res[nr - 1] = 0.5 * so[nr - 2];
  so = res;
  return CHEB;
}

// mult_xm1_1d (src/Ope_1d/mult_xm1_1d.cpp:29-41), CHEB.
int mult_xm1_1d(int base, std::vector<double>& so) {
  if (base != CHEB) not_impl("mult_xm1_1d", base);
  const int nr = (int)so.size();
  std::vector<double> res(nr, 0.0);
  res[0] = -so[0] + 0.5 * so[1];
  res[1] = so[0] - so[1] + 0.5 * so[2];
  for (int i = 2; i < nr - 1; i++) res[i] = 0.5 * so[i - 1] - so[i] + 0.5 * so[i + 1];
  res[nr - 1] = 0.5 * so[nr - 2] - so[nr - 1];
  so = res;
  return CHEB;
}

// div_xm1_1d (src/Ope_1d/div_xm1_1d.cpp:29-40), CHEB.
int div_xm1_1d(int base, std::vector<double>& so) {
  if (base != CHEB) not_impl("div_xm1_1d", base);
  const int nr = (int)so.size();
  std::vector<double> res(nr, 0.0);
  for (int i = 0; i < nr - 1; i++)
    for (int j = i + 1; j < nr; j++) res[i] += 2 * (j - i) * so[j];
  res[0] *= 0.5;
  so = res;
  return CHEB;
}

Mat op_matrix(Op1d op, int base, int n, int* out_base) {
  Mat M(n, n);
  int ob = -1;
  for (int c = 0; c < n; ++c) {
    std::vector<double> e(n, 0.0);
    e[c] = 1.0;
    ob = op(base, e);
    for (int r = 0; r < n; ++r) M(r, c) = e[r];
  }
  if (out_base) *out_base = ob;
  return M;
}

// ---------------------------------------------------------------------------
// Grids and transforms.
// ---------------------------------------------------------------------------
std::vector<double> coloc_r(int n) {
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) x[i] = -std::cos(M_PI * i / (n - 1));
  return x;
}
std::vector<double> coloc_theta(int n) {
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) x[i] = M_PI / 2. * i / (n - 1);
  return x;
}
std::vector<double> coloc_phi(int n) {
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) x[i] = M_PI * 2. * i / n;
  return x;
}

int ncoef_of(int base, int nconf) { return base == COSSIN ? nconf + 2 : nconf; }

// Coefficient slots that the reference's coef_i never reads (their basis
// function vanishes at every collocation point or is outside the expansion):
// COSSIN b_0 and b_{n/2} (coef_i_1d.cpp:217-239), last slot of COS_ODD and
// SIN_ODD (coef_i_1d.cpp:351-396, 429-475), first and last of SIN_EVEN (:398-427).
static bool coef_slot_used(int base, int nconf, int j) {
  const int nc = ncoef_of(base, nconf);
  switch (base) {
    case COSSIN: return !(j == 1 || j == nc - 1);
    case COS_ODD:
    case SIN_ODD: return j != nc - 1;
    case SIN_EVEN: return !(j == 0 || j == nc - 1);
    default: return true;
  }
}

// Points at which every used basis function vanishes; coef_i writes exact zeros
// there (coef_i_1d.cpp: cos_odd tab(nbr-1)=0, sin_odd/sin_even tab(0)=0).
static bool point_is_zero(int base, int nconf, int i) {
  switch (base) {
    case COS_ODD: return i == nconf - 1;
    case SIN_ODD: return i == 0;
    case SIN_EVEN: return i == 0 || i == nconf - 1;
    default: return false;
  }
}

double basis_value(int base, int nconf, int j, int i) {
  switch (base) {
    case CHEB: {
      const double t = M_PI * i / (nconf - 1);  // x_i = -cos t  ->  T_j(x_i) = (-1)^j cos(j t)
      return ((j & 1) ? -1.0 : 1.0) * std::cos(j * t);
    }
    case COSSIN: {
      const double ph = 2 * M_PI * i / nconf;
      const int m = j / 2;
      return (j & 1) ? std::sin(m * ph) : std::cos(m * ph);
    }
    case COS_EVEN: return std::cos(2 * j * (M_PI / 2. * i / (nconf - 1)));
    case COS_ODD: return std::cos((2 * j + 1) * (M_PI / 2. * i / (nconf - 1)));
    case SIN_EVEN: return std::sin(2 * j * (M_PI / 2. * i / (nconf - 1)));
    case SIN_ODD: return std::sin((2 * j + 1) * (M_PI / 2. * i / (nconf - 1)));
    default: not_impl("basis_value", base);
  }
}

Mat coef_i_matrix(int base, int nconf) {
  const int nc = ncoef_of(base, nconf);
  Mat E(nconf, nc);
  for (int i = 0; i < nconf; ++i) {
    if (point_is_zero(base, nconf, i)) continue;
    for (int j = 0; j < nc; ++j)
      if (coef_slot_used(base, nconf, j)) E(i, j) = basis_value(base, nconf, j, i);
  }
  return E;
}

// Gauss-Jordan inverse with partial pivoting (small, well-conditioned systems).
static std::vector<double> invert(std::vector<double> a, int n) {
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)p * n + c])) p = r;
    if (a[(size_t)p * n + c] == 0.0) throw std::runtime_error("singular transform");
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)p * n + k], a[(size_t)c * n + k]);
        std::swap(inv[(size_t)p * n + k], inv[(size_t)c * n + k]);
      }
    const double d = a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] /= d;
      inv[(size_t)c * n + k] /= d;
    }
    for (int r = 0; r < n; ++r)
      if (r != c) {
        const double f = a[(size_t)r * n + c];
        if (f == 0.0) continue;
        for (int k = 0; k < n; ++k) {
          a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
          inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
        }
      }
  }
  return inv;
}

// Analysis matrix.  CHEB and COS_EVEN: DCT-I with the trapezoid weights (the
// discrete orthogonality the reference's folded FFT computes, coef_1d.cpp:57-100,
// 314-363); COSSIN: real DFT in the [a0, 0, a1, b1, ..., a_{n/2}, 0] layout
// (coef_1d.cpp:254-292); COS_ODD / SIN_EVEN / SIN_ODD: exact inverse of the
// synthesis restricted to the points where the basis does not vanish
// identically and to the slots coef_i reads (coef_1d.cpp:414-605).
Mat coef_matrix(int base, int nconf) {
  const int nc = ncoef_of(base, nconf);
  Mat C(nc, nconf);
  if (base == CHEB || base == COS_EVEN) {
    const int N = nconf;
    for (int j = 0; j < N; ++j) {
      const double g = (j == 0 || j == N - 1) ? 2.0 : 1.0;
      for (int i = 0; i < N; ++i) {
        const double w = (i == 0 || i == N - 1) ? 0.5 : 1.0;
        C(j, i) = 2.0 / (N - 1) / g * w * basis_value(base, N, j, i);
      }
    }
    return C;
  }
  if (base == COSSIN) {
    const int n = nconf;
    for (int m = 0; m <= n / 2; ++m) {
      const double f = (m == 0 || 2 * m == n) ? 1.0 / n : 2.0 / n;
      for (int k = 0; k < n; ++k) {
        const double ph = 2 * M_PI * k / n;
        C(2 * m, k) = f * std::cos(m * ph);
        if (m > 0 && 2 * m < n) C(2 * m + 1, k) = 2.0 / n * std::sin(m * ph);
      }
    }
    return C;
  }
  std::vector<int> pts, slots;
  for (int i = 0; i < nconf; ++i)
    if (!point_is_zero(base, nconf, i)) pts.push_back(i);
  for (int j = 0; j < nc; ++j)
    if (coef_slot_used(base, nconf, j)) slots.push_back(j);
  if (pts.size() != slots.size()) throw std::runtime_error("coef_matrix: informative size mismatch");
  const int m = (int)pts.size();
  std::vector<double> a((size_t)m * m);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) a[(size_t)r * m + c] = basis_value(base, nconf, slots[c], pts[r]);
  const std::vector<double> inv = invert(a, m);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) C(slots[r], pts[c]) = inv[(size_t)r * m + c];
  return C;
}

}  // namespace kadhost
