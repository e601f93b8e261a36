// What the device runtime needs of a traced Newton system, whichever way it was
// traced (the hand-written XCTS system or a system parsed from equation strings):
// the domains, the value program (residual F), one tangent program per (field
// component, domain) with its seed table (Jacobian columns), the affecte_tau maps
// (update), the row blocks (check_equations), and optional programs for global
// quantities and parameter derivatives (adjoint sensitivities).
//
// Unknown order (system_of_eqs.cpp:597-612) is fixed by `layout`; row order by
// the equations' row0.
#pragma once
#include <array>
#include <string>
#include <vector>
#include "tape.hpp"

namespace kadhost {

// One equation (or equation block) and its rows.
struct EqRec {
  int kind;  // 0 inside, 1 matching d|d+1, 2 outer bc, 3 inner bc, 4 integral (one row)
  int d;
  std::string which;
  int ncomp;
  int row0;
  std::vector<int> nrows;  // per component
  int order = 2;           // tau order of inside equations (rows i < nr - order)
  int bound = 1;           // boundary of bc / matching equations (INNER_BC 0, OUTER_BC 1)
};

struct SeedTab {
  std::vector<int> pos;     // 3 per unknown
  std::vector<double> val;  // 3 per unknown
  std::vector<int> cnt;
};

// affecte_tau of one (field component, domain) as CSR: coefficient <- unknowns.
struct Csr {
  std::vector<int> rowptr, col;
  std::vector<double> w;
};

// Global quantity: value slot of the coefficients of its integrand (r r sin(theta) so
// at the outer boundary), theta basis of phi coefficient 0, factor.
struct ObsRec {
  int slot = -1;  // -1: identically zero
  int bt0 = 0;
  double fac = 0;
};

class SystemBase {
 public:
  std::vector<Domain> doms;
  Tracer tr;
  int dmin = 1, dmax = 1;
  int nfields = 0;                 // scalar field components (unknowns are their coefficients)
  std::vector<std::string> field_names;
  std::vector<std::vector<int>> var_fields;  // field components of each unknown variable, in add_var order
  std::vector<EqRec> eqs;
  int n_unknowns = 0, n_rows = 0;
  std::vector<std::array<int, 4>> layout;   // field, domain, first unknown, count
  std::vector<std::vector<int>> field_slot;  // value slot of each field's coefficients, [field][domain]
  std::vector<Program> tprog;               // one per layout entry
  std::vector<SeedTab> seeds;
  std::vector<Csr> tau_maps;                // indexed like layout
  // global quantities at the outer boundary (may be empty)
  Program obs_prog;
  std::vector<ObsRec> obs;
  std::vector<Program> obs_tprog;                 // d(quantities)/d(unknowns) programs
  std::vector<std::array<int, 4>> obs_tprog_lay;  // field, domain, first unknown, count
  std::vector<int> obs_tprog_tab;                 // seed table of each
  // dF/dp of the parameters (constants that carry a tangent): one program each,
  // seeded by a one-coefficient table (the constant 1), exporting into the rows
  std::vector<std::string> par_names;
  std::vector<Program> par_prog;
  std::vector<int> par_tab;
  // global quantities given as integral expressions (the parsed systems' counterpart of
  // the observables above): a value program writing Q[q], one tangent program per
  // (field, domain) writing dQ_q/du at q * n + column, and per parameter a program
  // writing dQ_q/dp at fixed fields
  std::vector<std::string> glob_names;
  Program glob_prog;
  std::vector<Program> glob_tprog;                 // indexed like layout
  std::vector<Program> glob_par_prog;              // indexed like par_names
  // double unknowns (add_var of a double): unknowns 0 .. ndouble-1, each a parameter
  // whose dF/dp program is its Jacobian column; the fields' unknowns follow
  int ndouble = 0;
  std::vector<int> dvar_par;  // parameter index of each double unknown

  SystemBase() : tr(&doms) {}
  SystemBase(const SystemBase&) = delete;
  SystemBase& operator=(const SystemBase&) = delete;
  virtual ~SystemBase() = default;
  // spectral base of field component f in domain d
  virtual Base field_base(int f, int d) const = 0;
  // value slot of a named constant field in domain d (-1 if absent); set_param writes it
  virtual int param_slot(const std::string& name, int d) const = 0;
  virtual void build() = 0;
};

}  // namespace kadhost
