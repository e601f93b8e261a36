// A Newton system given the way Kadath users give it: unknowns, constants,
// definitions and equations as strings (System_of_eqs::add_var / add_cst / add_def /
// add_eq_inside / add_eq_bc / add_eq_matching, src/System_of_eqs/system_of_eqs.cpp,
// add_eq.cpp; Space_spheric::add_eq / add_inner_bc / add_outer_bc,
// src/Space/Spheric/spheric_add_eq.cpp), traced into the device programs of
// SystemBase.
//
// Expression grammar (the splitting order of System_of_eqs::find_ope,
// src/System_of_eqs/give_ope.cpp:559-1135): a string is split at its last top-level
// '+', else '-', else a leading '-', else the last '*', else the last '/'; then the
// unary operators Lap / lap / dn / dr / multr / divr / multsint / divsint, the
// derivatives D_i / D^i / partial_i (flat metric, Cartesian basis), X^n with an integer
// n, then names: unknowns, the flat metric, constants, numbers, definitions.
//
// Tensors are Cartesian (CARTESIAN_BASIS) with the flat metric, so upper and lower
// indices coincide; a repeated index name is summed (Einstein convention), as
// Term_eq products and Ope_der do in the reference.  Unknowns and constants are
// scalars or vectors; vector components use the standard bases of
// Tensor::std_base for Cartesian vectors (x, y: std; z: anti, tensor.cpp:609-650).
#pragma once
#include <map>
#include <string>
#include <vector>
#include "system_base.hpp"

namespace kadhost {

class ParsedSystem : public SystemBase {
 public:
  // Space_spheric(CHEB_TYPE, 0, (nr, nt, np), bounds) solved in domains dmin..dmax;
  // order 0: the reference unknown / row order, 1: domain-major.
  ParsedSystem(int nr, int nt, int np, const std::vector<double>& bounds, int dmin, int dmax, int order);

  // unknown scalar (valence 0) or Cartesian vector (valence 1)
  void add_var(const std::string& name, int valence);
  // double constant (System_of_eqs::add_cst(const char*, double))
  void add_cst(const std::string& name, double value);
  // double unknown (System_of_eqs::add_var(const char*, double&), system_of_eqs.cpp:382-400):
  // a global unknown, before every field in the unknown order; returns its index
  int add_var_double(const std::string& name, double value);
  // double constant that carries a tangent: a parameter of the adjoint sensitivities
  // (its value can be changed after build by set_param); index = its parameter number
  int add_param(const std::string& name, double value);
  // constant field (add_cst(const char*, const Tensor&)): conf values for every
  // component (1 or 3) and every solved domain, component-major then domain, each a
  // domain array of conf_size() values
  void add_cst_field(const std::string& name, int valence, const std::vector<double>& conf);
  // flat metric (Metric_flat::set_system, metric_flat.cpp): f_ij = f^ij = delta
  void set_flat_metric(const std::string& name);
  void add_def(const std::string& def);
  void add_eq_inside(int dom, const std::string& eq, int order = 2);
  void add_eq_bc(int dom, int bound, const std::string& eq);
  void add_eq_matching(int dom, int bound, const std::string& expr);
  // integral equation evaluated in domain dom at a boundary, e.g. "integ(intS) - chi = 0"
  // (Eq_int, eq_int.cpp:38-51; Space_adapted_bh::add_eq_int_bh, adapted_bh_add_eq.cpp:88-119);
  // its row comes before every other equation's (sec_member, solver.cpp:49-85)
  void add_eq_int(int dom, int bound, const std::string& eq);
  // global quantity: an integral expression evaluated in dom at bound (M_ADM, J_ADM, ...),
  // reported by the value program and differentiated for the adjoint sensitivities
  int add_global(const std::string& name, int dom, int bound, const std::string& expr);
  // Space_spheric helpers (spheric_add_eq.cpp:26-52)
  void space_add_eq(const std::string& eq, const std::string& m1, const std::string& m2);
  void space_add_inner_bc(const std::string& eq) { add_eq_bc(dmin, 0, eq); }
  void space_add_outer_bc(const std::string& eq) { add_eq_bc(dmax, 1, eq); }

  Base field_base(int f, int d) const override;
  int param_slot(const std::string& name, int d) const override;
  void build() override;
  // initial values of unknown field component f (conf values of domain d)
  int var_of_field(int f) const { return field_var[f]; }

  int order_mode = 0;

 private:
  // One integral of a sum of integrals: r r sin(theta) so of the value and of the
  // tangent, the boundary it is taken at, and its factor.
  struct GlTerm {
    VDp val, der;
    int dom, bound;
    double fac;
    int point = 0;  // 1: the value of a parameter (a constant field), not an integral
  };
  // A scalar component in one domain: a plain number, a Term_eq, or (glob) a global
  // double sum_k fac_k integ(.)_k + x, as Ope_int returns (ope_int.cpp).
  struct Sc {
    bool num = false;
    double x = 0.0;
    TE te;
    bool glob = false;
    std::vector<GlTerm> gl;
  };
  // A Cartesian tensor with named indices: 3^valence components, the first index
  // slowest.
  struct Tv {
    std::string idx;
    std::vector<Sc> c;
  };
  struct Var {
    std::string name;
    int valence;
    int f0;  // first field component
  };
  struct Cst {
    std::string name;
    int valence = 0;
    bool field = false;      // a constant field (else a double)
    bool param = false;      // carries a tangent in its parameter program
    int par = -1;
    int dvar = -1;           // >= 0: a double unknown (its index)
    double value = 0.0;
    std::vector<double> conf;  // field: component-major, then domain, conf values
  };
  struct Def {
    std::string name, idx, rhs;
  };
  struct Ctx {
    int d = 0;
    int bound = 1;
    int seed_f = -1;       // tangent-seeded field component (-1: none)
    VDp seed;              // its seed node
    int seed_par = -1;     // tangent-seeded parameter (-1: none)
    int par_tab = -1;
    VDp par_seed;          // its seed node in this domain (created on first use)
    bool glob_ctx = false; // evaluating an integral expression: parameters are scalars
    std::map<std::string, Tv> defs;  // definitions evaluated in this context
  };

  std::vector<Var> vars;
  std::vector<int> field_var, field_comp;
  std::vector<Cst> csts;
  std::vector<Def> defs;
  struct Glob {
    std::string name, expr;
    int dom, bound;
  };
  std::vector<Glob> globs;
  void trace_globals();
  void export_glob_terms(const Sc& g, int row, bool der, int rowstride_q);
  std::string metric;
  std::vector<std::vector<VDp>> field_nodes;  // [domain][field] value nodes
  std::map<std::string, VDp> const_nodes;     // "name:comp:domain" value nodes of constant fields
  std::map<std::pair<int, int>, int> seed_tab_of;

  Tv eval(const std::string& s, Ctx& cx);
  Tv eval_name(const std::string& s, Ctx& cx);
  Tv tv_add(const Tv& a, const Tv& b, bool sub, Ctx& cx);
  Tv tv_mul(const Tv& a, const Tv& b, Ctx& cx);
  Tv tv_div(const Tv& a, const Tv& b, Ctx& cx);
  Tv tv_neg(const Tv& a, Ctx& cx);
  Tv tv_map(const Tv& a, const std::string& op, Ctx& cx);
  Tv tv_grad(const Tv& a, char name, Ctx& cx);
  Tv tv_pow(const Tv& a, int n, Ctx& cx);
  Tv contract(const Tv& a, Ctx& cx);
  Sc sc_add(const Sc& a, const Sc& b, bool sub, Ctx& cx);
  Sc sc_mul(const Sc& a, const Sc& b, Ctx& cx);
  Sc sc_div(const Sc& a, const Sc& b, Ctx& cx);
  TE as_field(const Sc& a, Ctx& cx);
  VDp const_field(const std::string& key, int d, const std::vector<double>& conf, const Base& b,
                  bool keep_zero = false);
  Tv residual(const std::string& eq, Ctx& cx);
  std::vector<Sc> eq_parts(const EqRec& e, int part, Ctx& cx);
  void trace_value();
  void trace_tangent(int f, int d);
  void trace_param(int p);
  void reorder_rows();
};

}  // namespace kadhost
