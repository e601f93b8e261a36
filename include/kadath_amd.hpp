/*
 * kadath_amd.hpp -- the reference's C++ front end of the Newton path, over the
 * C-ABI of kadath_amd.h (header only; link with libkadath_amd.so).
 *
 * Mirrors the calls a Kadath program makes around System_of_eqs::do_newton
 * (tutorials/tuto5.cpp, codes/Par_version/Kerr/kerr.cpp):
 *   Space_spheric space(CHEB_TYPE, center, res, bounds);
 *   Scalar field(space); field = 1; field.std_base();
 *   System_of_eqs syst(space, 1, ndom - 1);
 *   syst.add_var("F", field); syst.add_cst("a", rad);
 *   syst.add_eq_inside(d, "Lap(F)=0"); syst.add_eq_bc(1, INNER_BC, "dn(F) + 0.5*F/a = 0");
 *   syst.add_eq_matching(d, OUTER_BC, "F");
 *   while (!syst.do_newton(prec, error)) {}
 * with the reference's semantics (include/system_of_eqs.hpp:134,871,
 * src_par/do_newton.cpp:65-430): do_newton returns true when ||F||_inf < precision at
 * entry (error receives that value) and otherwise takes one Newton step, which updates
 * the registered fields in place.
 *
 * Differences, each forced by the boundary: a Space_spheric here is Chebyshev in r
 * with the nucleus excised (systems are solved from domain 1), tensors are scalars
 * and Cartesian vectors over the flat metric, fields are held as their values at the
 * collocation points (Scalar::set_domain(d)[i + nr (j + nt k)]), and errors throw
 * std::runtime_error where the reference prints and aborts -- except a singular
 * Jacobian, which prints the reference's message and aborts as it does
 * (do_newton.cpp:323-330).
 */
#ifndef KADATH_AMD_HPP
#define KADATH_AMD_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "kadath_amd.h"

namespace kadath_amd {

enum { INNER_BC = 0, OUTER_BC = 1 };
enum { CHEB_TYPE = 1 };
// do_newton's solver argument: the reference's SOLVER (NEWTON_RAPHSON, GAUSS_NEWTON,
// system_of_eqs.hpp:134) plus the mixed-precision solvers of this build; the values are
// the C-ABI's solver codes
enum SOLVER {
  NEWTON_RAPHSON = 0,
  NEWTON_F32_IR = 1,
  NEWTON_F32_GMRES = 2,
  NEWTON_BF16_GMRES = 3,
  GAUSS_NEWTON = 4,
  NEWTON_F64_REUSE = 5,    // fp64 factors of an earlier Jacobian + fp64 flexible GMRES
  NEWTON_F64_DD = 6,       // domain-decomposed fp64 factors + fp64 flexible GMRES
  NEWTON_F64_DD_REUSE = 7  // the same, the factors of an earlier Jacobian kept while they precondition well
};

namespace detail {
inline void check(int rc, const char* what) {
  if (rc < 0) throw std::runtime_error(std::string(what) + " failed with status " + std::to_string(rc));
}
// one device context per process (device 0), created on first use
inline kad_ctx* context() {
  static kad_ctx* ctx = nullptr;
  if (!ctx) check(kad_ctx_create(&ctx, 0), "kad_ctx_create");
  return ctx;
}
}  // namespace detail

// Space_spheric (src/Space/Spheric/space_spheric.cpp:25-47): a nucleus [0, b0],
// shells [b_i, b_i+1] and a compactified domain [b_last, inf), (nr, nt, np) points each.
class Space_spheric {
 public:
  Space_spheric(int type_coloc, const double center[3], const int res[3], const std::vector<double>& bounds)
      : nr_(res[0]), nt_(res[1]), np_(res[2]), bounds_(bounds) {
    if (type_coloc != CHEB_TYPE) throw std::runtime_error("Space_spheric: CHEB_TYPE only");
    if (center && (center[0] != 0 || center[1] != 0 || center[2] != 0))
      throw std::runtime_error("Space_spheric: centred at the origin only");
    if (bounds.empty()) throw std::runtime_error("Space_spheric: at least one bound");
  }
  int get_nbr_domains() const { return (int)bounds_.size() + 1; }
  int nr() const { return nr_; }
  int nt() const { return nt_; }
  int np() const { return np_; }
  int conf_size() const { return nr_ * nt_ * np_; }
  const std::vector<double>& bounds() const { return bounds_; }
  // collocation points of domain d >= 1 (r = +inf at infinity)
  void coloc(int d, std::vector<double>& r, std::vector<double>& theta, std::vector<double>& phi) const {
    kad_sysdef* def = nullptr;
    detail::check(kad_sysdef_create(nr_, nt_, np_, bounds_.data(), (int)bounds_.size(), 1, (int)bounds_.size(), &def),
                  "kad_sysdef_create");
    r.resize(conf_size());
    theta.resize(conf_size());
    phi.resize(conf_size());
    const int rc = kad_sysdef_coloc(def, d, r.data(), theta.data(), phi.data());
    kad_sysdef_destroy(def);
    detail::check(rc, "kad_sysdef_coloc");
  }
  // radius of the collocation points of domain d (Domain::get_radius)
  std::vector<double> get_radius(int d) const {
    std::vector<double> r, t, p;
    coloc(d, r, t, p);
    return r;
  }
  inline void add_eq(class System_of_eqs& sys, const char* eq, const char* m1, const char* m2) const;
  inline void add_inner_bc(class System_of_eqs& sys, const char* eq) const;
  inline void add_outer_bc(class System_of_eqs& sys, const char* eq) const;

 private:
  int nr_, nt_, np_;
  std::vector<double> bounds_;
};

// A scalar field: values at the collocation points of every domain (the nucleus,
// domain 0, is held but never solved for).
class Scalar {
 public:
  explicit Scalar(const Space_spheric& sp) : sp_(&sp), val_(sp.get_nbr_domains(), std::vector<double>(sp.conf_size())) {}
  Scalar& operator=(double x) {
    for (auto& v : val_)
      for (double& e : v) e = x;
    dirty_ = true;
    return *this;
  }
  void std_base() {}  // bases follow from the field's role (Tensor::std_base)
  void annule_hard() { *this = 0.0; }
  std::vector<double>& set_domain(int d) {
    dirty_ = true;
    return val_.at(d);
  }
  const std::vector<double>& operator()(int d) const { return val_.at(d); }
  const Space_spheric& get_space() const { return *sp_; }
  bool dirty_ = true;

 private:
  const Space_spheric* sp_;
  std::vector<std::vector<double>> val_;
};

// A contravariant vector in the Cartesian basis: components 1, 2, 3 = x, y, z.
class Vector {
 public:
  explicit Vector(const Space_spheric& sp) : c_{Scalar(sp), Scalar(sp), Scalar(sp)} {}
  Scalar& set(int i) { return c_[i - 1]; }
  const Scalar& operator()(int i) const { return c_[i - 1]; }
  void std_base() {}

 private:
  Scalar c_[3];
};

class System_of_eqs {
 public:
  System_of_eqs(const Space_spheric& sp, int dmin, int dmax) : sp_(&sp), dmin_(dmin), dmax_(dmax) {
    detail::check(kad_sysdef_create(sp.nr(), sp.nt(), sp.np(), sp.bounds().data(), (int)sp.bounds().size(), dmin, dmax,
                                    &def_),
                  "System_of_eqs");
  }
  System_of_eqs(const System_of_eqs&) = delete;
  System_of_eqs& operator=(const System_of_eqs&) = delete;
  ~System_of_eqs() {
    if (sys_) kad_system_destroy(sys_);
    if (def_) kad_sysdef_destroy(def_);
  }

  void add_var(const char* name, Scalar& s) {
    detail::check(kad_sysdef_add_var(declaring(), name, 0), "add_var");
    fields_.push_back(&s);
  }
  void add_var(const char* name, Vector& v) {
    detail::check(kad_sysdef_add_var(declaring(), name, 1), "add_var");
    for (int i = 1; i <= 3; ++i) fields_.push_back(&v.set(i));
  }
  void add_cst(const char* name, double x) { detail::check(kad_sysdef_add_cst(declaring(), name, x), "add_cst"); }
  void add_cst(const char* name, const Scalar& s) { add_cst_field(name, 0, {&s}); }
  void add_cst(const char* name, const Vector& v) { add_cst_field(name, 1, {&v(1), &v(2), &v(3)}); }
  // a constant that is a parameter of the sensitivities (dF/dp); returns its index
  int add_param(const char* name, double x) {
    const int rc = kad_sysdef_add_param(declaring(), name, x);
    detail::check(rc, "add_param");
    return rc;
  }
  void set_flat_metric(const char* name) { detail::check(kad_sysdef_set_flat_metric(declaring(), name), "metric"); }
  void add_def(const char* def) { detail::check(kad_sysdef_add_def(declaring(), def), "add_def"); }
  void add_eq_inside(int dom, const char* eq) { add_eq_order(dom, 2, eq); }
  void add_eq_order(int dom, int order, const char* eq) {
    detail::check(kad_sysdef_add_eq_inside(declaring(), dom, eq, order), "add_eq_inside");
  }
  void add_eq_bc(int dom, int bound, const char* eq) {
    detail::check(kad_sysdef_add_eq_bc(declaring(), dom, bound, eq), "add_eq_bc");
  }
  void add_eq_matching(int dom, int bound, const char* eq) {
    detail::check(kad_sysdef_add_eq_matching(declaring(), dom, bound, eq), "add_eq_matching");
  }
  int get_dom_min() const { return dmin_; }
  int get_dom_max() const { return dmax_; }
  // unknown order of the device system: 0 the reference's, 1 domain-major (default;
  // a block-tridiagonal Jacobian whose LU skips the structural zeros)
  void set_order(int order) { order_ = order; }

  int get_nbr_unknowns() {
    ready();
    int n = 0;
    detail::check(kad_system_sizes(sys_, &n, nullptr, nullptr), "get_nbr_unknowns");
    return n;
  }
  // max |F| per equation component (System_of_eqs::check_equations, solver.cpp:24-47)
  std::vector<double> check_equations() {
    ready();
    std::vector<double> e(4096);
    const int q = kad_check_equations(sys_, e.data(), (int)e.size());
    detail::check(q, "check_equations");
    e.resize(q);
    return e;
  }
  // One Newton iteration (do_newton.cpp:65-430); true when ||F||_inf < precision at
  // entry.  The registered fields are read at entry and written back after the step.
  bool do_newton(double precision, double& error, SOLVER solver = NEWTON_RAPHSON) {
    ready();
    double times[5];
    const int rc = kad_system_do_newton(sys_, precision, &error, (int)solver, times);
    if (rc == KAD_ERR_SINGULAR) {
      std::fprintf(stderr, "Jacobian is singular! Stopping.\n");
      std::abort();
    }
    detail::check(rc, "do_newton");
    if (rc == 1) return true;
    download();
    return false;
  }
  // Newton iteration with backtracking line search (do_newton.cpp:454-517); true when
  // ||F||_inf < precision at entry.  The reference aborts on a roundoff failure of the
  // search, and so does this call.
  bool do_newton_with_linesearch(double precision, double& error, int ntrymax = 10, double stepmax = 1.0) {
    ready();
    const int rc = kad_system_do_newton_linesearch(sys_, precision, &error, ntrymax, stepmax, NEWTON_RAPHSON,
                                                   nullptr, nullptr);
    if (rc == KAD_ERR_SINGULAR) {
      std::fprintf(stderr, "Jacobian is singular! Stopping.\n");
      std::abort();
    }
    if (rc == KAD_ERR_NOCONV) std::abort();
    detail::check(rc, "do_newton_with_linesearch");
    if (rc == 1) return true;
    download();
    return false;
  }
  kad_system* handle() {
    ready();
    return sys_;
  }

 private:
  const Space_spheric* sp_;
  int dmin_, dmax_;
  int order_ = 1;
  kad_sysdef* def_ = nullptr;
  kad_system* sys_ = nullptr;
  std::vector<Scalar*> fields_;

  kad_sysdef* declaring() {
    if (!def_) throw std::runtime_error("System_of_eqs: the system is already built");
    return def_;
  }
  void add_cst_field(const char* name, int valence, std::vector<const Scalar*> comps) {
    std::vector<double> conf;
    for (const Scalar* c : comps)
      for (int d = dmin_; d <= dmax_; ++d) {
        const std::vector<double>& v = (*c)(d);
        conf.insert(conf.end(), v.begin(), v.end());
      }
    detail::check(kad_sysdef_add_cst_field(declaring(), name, valence, conf.data(), (int64_t)conf.size()), "add_cst");
  }
  // build on first use, then upload the fields the caller changed (vars_to_terms)
  void ready() {
    if (!sys_) {
      kad_sysdef* d = def_;
      def_ = nullptr;
      detail::check(kad_system_build(detail::context(), d, order_, 4096, &sys_), "System_of_eqs build");
      for (Scalar* s : fields_) s->dirty_ = true;
    }
    for (size_t f = 0; f < fields_.size(); ++f) {
      if (!fields_[f]->dirty_) continue;
      for (int d = dmin_; d <= dmax_; ++d)
        detail::check(kad_system_set_field(sys_, (int)f, d, (*fields_[f])(d).data()), "set_field");
      fields_[f]->dirty_ = false;
    }
  }
  void download() {
    for (size_t f = 0; f < fields_.size(); ++f) {
      for (int d = dmin_; d <= dmax_; ++d)
        detail::check(kad_system_get_field(sys_, (int)f, d, fields_[f]->set_domain(d).data()), "get_field");
      fields_[f]->dirty_ = false;
    }
  }
};

// Space_spheric::add_eq / add_inner_bc / add_outer_bc over the system's domain range
// (spheric_add_eq.cpp:26-52)
inline void Space_spheric::add_eq(System_of_eqs& sys, const char* eq, const char* m1, const char* m2) const {
  for (int d = sys.get_dom_min(); d < sys.get_dom_max(); ++d) {
    sys.add_eq_inside(d, eq);
    sys.add_eq_matching(d, OUTER_BC, m1);
    sys.add_eq_matching(d, OUTER_BC, m2);
  }
  sys.add_eq_inside(sys.get_dom_max(), eq);
}
inline void Space_spheric::add_inner_bc(System_of_eqs& sys, const char* eq) const {
  sys.add_eq_bc(sys.get_dom_min(), INNER_BC, eq);
}
inline void Space_spheric::add_outer_bc(System_of_eqs& sys, const char* eq) const {
  sys.add_eq_bc(sys.get_dom_max(), OUTER_BC, eq);
}

}  // namespace kadath_amd

#endif
