// 1-D spectral bases of the Kadath spherical domains and their coefficient-space
// operators, as dense per-line matrices for the batched GPU kernels.
//
// Basis ids are the reference's (include/base_spectral.hpp:27-42).  The 1-D
// operators restate src/Ope_1d/*.cpp and src/Coef/coef_1d.cpp / coef_i_1d.cpp;
// each builder below cites the routine it follows.
#pragma once
#include <vector>
#include <string>

namespace kadhost {

enum Base1d : int {
  CHEB = 1, CHEB_EVEN = 2, CHEB_ODD = 3, COSSIN = 4, COS = 5, COS_EVEN = 6, COS_ODD = 7,
  SIN = 8, SIN_EVEN = 9, SIN_ODD = 10, LEG = 11, LEG_EVEN = 12, LEG_ODD = 13,
  COSSIN_EVEN = 14, COSSIN_ODD = 15
};

// Dense row-major matrix: y = M x, M is nout x nin.
struct Mat {
  int nout = 0, nin = 0;
  std::vector<double> a;
  Mat() = default;
  Mat(int o, int i) : nout(o), nin(i), a((size_t)o * i, 0.0) {}
  double& operator()(int r, int c) { return a[(size_t)r * nin + c]; }
  double operator()(int r, int c) const { return a[(size_t)r * nin + c]; }
};
Mat matmul(const Mat& A, const Mat& B);
Mat identity(int n);

// Coefficient-space operators on one line.  Each takes the line in place and
// returns the output basis, exactly as the reference's int f(Array<double>&).
typedef int (*Op1d)(int base, std::vector<double>& tab);
int der_1d(int base, std::vector<double>& tab);        // src/Ope_1d/der_1d.cpp
int mult_sin_1d(int base, std::vector<double>& tab);   // src/Ope_1d/mult_sin_1d.cpp
int mult_cos_1d(int base, std::vector<double>& tab);   // src/Ope_1d/mult_cos_1d.cpp
int div_sin_1d(int base, std::vector<double>& tab);    // src/Ope_1d/div_sin_1d.cpp
int mult_x_1d(int base, std::vector<double>& tab);     // src/Ope_1d/mult_x_1d.cpp
int mult_xm1_1d(int base, std::vector<double>& tab);   // src/Ope_1d/mult_xm1_1d.cpp
int div_xm1_1d(int base, std::vector<double>& tab);    // src/Ope_1d/div_xm1_1d.cpp

// Matrix of a linear Op1d for a line of length n in basis `base`; *out_base
// receives the output basis.
Mat op_matrix(Op1d op, int base, int n, int* out_base);

// Collocation points of the 1-D grids of Domain_shell / Domain_compact
// (src/Domain/Spheric/domain_shell.cpp:252-267).
std::vector<double> coloc_r(int n);      // x_i = -cos(pi i/(n-1))
std::vector<double> coloc_theta(int n);  // theta_j = pi/2 j/(n-1)
std::vector<double> coloc_phi(int n);    // phi_k = 2 pi k/n

// Synthesis (coef_i) and analysis (coef) matrices of one 1-D basis.
// nconf = points, ncoef = coefficients (COSSIN: ncoef = nconf + 2).
Mat coef_i_matrix(int base, int nconf);  // ncoef -> nconf (evaluation of the series)
Mat coef_matrix(int base, int nconf);    // nconf -> ncoef
int ncoef_of(int base, int nconf);
// Value of basis function j of `base` at the collocation point i.
double basis_value(int base, int nconf, int j, int i);

}  // namespace kadhost
