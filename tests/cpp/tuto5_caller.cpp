// A Kadath-style program compiled against include/kadath_amd.hpp: the steps of
// tutorials/tuto5.cpp (Laplace equation, Robin inner boundary condition, exact
// solution 1 + a/r) through System_of_eqs::do_newton on the device, with the single
// LU and with the domain-decomposed solver.
//
// Exit status: 0 = converged and within 1e-8 of the exact solution at every
// collocation point (the radial truncation of 13 points is 4e-9); 1 = wrong answer; 3 = no device (the declaration and its
// argument checks ran, the first device call failed cleanly).
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "kadath_amd.hpp"

using namespace kadath_amd;

int main() {
  const int res[3] = {13, 5, 4};
  const double center[3] = {0, 0, 0};
  const std::vector<double> bounds = {1, 2, 4};
  const int ndom = 4;
  Space_spheric space(CHEB_TYPE, center, res, bounds);
  System_of_eqs syst(space, 1, ndom - 1);
  Scalar field(space);
  field = 1;
  field.std_base();
  syst.add_var("F", field);
  const double rad = space.get_radius(1)[0];  // inner radius of the first shell
  std::printf("The inner radius of the first shell is %g\n", rad);
  syst.add_cst("a", rad);
  for (int d = 1; d < ndom; d++) syst.add_eq_inside(d, "Lap(F)=0");
  syst.add_eq_bc(1, INNER_BC, "dn(F) + 0.5*F/a = 0");
  syst.add_eq_bc(ndom - 1, OUTER_BC, "F=1");
  for (int d = 1; d < ndom - 1; d++) {
    syst.add_eq_matching(d, OUTER_BC, "F");
    syst.add_eq_matching(d, OUTER_BC, "dn(F)");
  }
  try {
    std::vector<double> errors = syst.check_equations();
    std::printf("The system has %zu equations.\n", errors.size());
    double prec = 1e-8, error = 0;
    int it = 0;
    bool endloop = false;
    while (!endloop && it < 5) {
      endloop = syst.do_newton(prec, error);
      std::printf("Newton iteration %d error %.3e\n", it++, error);
    }
    if (!endloop) return 1;
    double worst = 0;
    for (int d = 1; d < ndom; d++) {
      const std::vector<double> r = space.get_radius(d);
      for (size_t q = 0; q < r.size(); q++) {
        const double analytic = std::isfinite(r[q]) ? 1 + rad / r[q] : 1.0;
        worst = std::fmax(worst, std::fabs(field(d)[q] - analytic));
      }
    }
    std::printf("max |numerical - analytic| = %.3e\n", worst);
    if (!(worst < 1e-8)) return 1;
    // the same system solved by domains (the domain blocks factored together, the
    // interfaces through the capacitance matrix): the same known answer
    {
      System_of_eqs sdd(space, 1, ndom - 1);
      Scalar f2(space);
      f2 = 1;
      f2.std_base();
      sdd.add_var("F", f2);
      sdd.add_cst("a", rad);
      for (int d = 1; d < ndom; d++) sdd.add_eq_inside(d, "Lap(F)=0");
      sdd.add_eq_bc(1, INNER_BC, "dn(F) + 0.5*F/a = 0");
      sdd.add_eq_bc(ndom - 1, OUTER_BC, "F=1");
      for (int d = 1; d < ndom - 1; d++) {
        sdd.add_eq_matching(d, OUTER_BC, "F");
        sdd.add_eq_matching(d, OUTER_BC, "dn(F)");
      }
      bool done = false;
      for (int k = 0; k < 5 && !done; ++k) {
        done = sdd.do_newton(prec, error, NEWTON_F64_DD);
        std::printf("domain-decomposed Newton iteration %d error %.3e\n", k, error);
      }
      if (!done) return 1;
      double worst2 = 0;
      for (int d = 1; d < ndom; d++) {
        const std::vector<double> r = space.get_radius(d);
        for (size_t q = 0; q < r.size(); q++) {
          const double analytic = std::isfinite(r[q]) ? 1 + rad / r[q] : 1.0;
          worst2 = std::fmax(worst2, std::fabs(f2(d)[q] - analytic));
        }
      }
      std::printf("domain-decomposed: max |numerical - analytic| = %.3e\n", worst2);
      if (!(worst2 < 1e-8)) return 1;
    }
    // a declaration the reference rejects (an operator without a device form) fails
    // to build with an argument error, not a device error
    System_of_eqs bad(space, 1, ndom - 1);
    Scalar g(space);
    g = 1;
    bad.add_var("G", g);
    for (int d = 1; d < ndom; d++) bad.add_eq_inside(d, "sqrt(G) = 1");
    try {
      bad.check_equations();
      std::printf("unsupported operator accepted\n");
      return 1;
    } catch (const std::runtime_error& e) {
      std::printf("rejected: %s\n", e.what());
      if (std::string(e.what()).find("status -1") == std::string::npos) return 1;
    }
    return 0;
  } catch (const std::runtime_error& e) {
    std::printf("no device: %s\n", e.what());
    return 3;
  }
}
