// Spherical domains of a Kadath Space_spheric (host side): geometry, spectral
// bases and the tau / Galerkin bookkeeping that fixes the unknown and row order
// of the Newton system.
//
// Restates Domain_shell (src/Domain/Spheric/domain_shell.cpp) and Domain_compact
// (src/Domain/Spheric/domain_compact.cpp) for CHEB_TYPE, valence 0 and Cartesian
// vectors (the cases of the XCTS black-hole system).
#pragma once
#include <vector>
#include <string>
#include "spectral.hpp"

namespace kadhost {

enum { INNER_BC = 0, OUTER_BC = 1 };

// Base_spectral of one 3-D domain (src/Base_spectral/base_spectral.cpp:92-104):
// br[k*nt + j] for the r lines, bt[k] for the theta lines, bp for phi.
struct Base {
  bool def = false;
  std::vector<int> br, bt;
  int bp = COSSIN;
  bool operator==(const Base& o) const { return def && o.def && bp == o.bp && bt == o.bt && br == o.br; }
  bool operator!=(const Base& o) const { return !(*this == o); }
};

// One exported row of export_tau (domain_shell_export_tau.cpp:27-120) or
// export_tau_boundary (domain_shell_export_tau_boundary.cpp:27-120):
// value = X(pos) - gw * X(gpos)   (gpos < 0: no Galerkin partner).
// Inside rows: pos/gpos are coefficient indices i + nr*(j + nt*k).
// Boundary rows: pos/gpos are line indices (j + nt*k); X(line) = sum_i s^i cf(i, line).
struct TauRow {
  int pos, gpos;
  double gw;
};

class Domain {
 public:
  int num = 0;
  bool compact = false;
  // a shell of an adapted space (Domain_shell_inner_adapted / _outer_adapted, e.g. the
  // homothetic shells of Space_adapted_bh) whose boundary spheres are round: its
  // surface integrals weight in configuration space (domain_shell_inner_adapted.cpp:1134-1208)
  bool adapted = false;
  int nr = 0, nt = 0, np = 0, npc = 0;
  double alpha = 0, beta = 0;
  std::vector<double> xr, th, ph;  // collocation points
  std::vector<double> R;           // radius on the conf grid, (k, j, i) with i fastest

  Domain() = default;
  // shell [rint, rext] (domain_shell.cpp:28-35) or compactified domain r >= rint (domain_compact.cpp:28-33)
  Domain(int num, bool compact, int nr, int nt, int np, double rint, double rext);

  int conf_size() const { return nr * nt * np; }
  int coef_size() const { return nr * nt * npc; }

  Base std_base() const;   // set_cheb_base  (domain_shell.cpp:288-309)
  Base anti_base() const;  // set_anti_cheb_base (domain_shell.cpp:319-340)
  Base mult_base(const Base& a, const Base& b) const;  // Domain_shell::mult (domain_shell.cpp:685-850)

  // ---- tau bookkeeping (mlim = 0 throughout: scalars and Cartesian components)
  // (k, j) lines carrying unknowns, k outer / j inner (domain_shell_nbr_unknowns.cpp:26-66).
  std::vector<std::pair<int, int>> unknown_slots(const std::vector<int>& bt) const;
  int nbr_unknowns(const std::vector<int>& bt) const { return (int)unknown_slots(bt).size() * nr; }
  // Unit seed of unknown u: coefficient positions and values, with the Galerkin
  // partner (domain_shell_affecte_tau_one_coef.cpp:26-105).
  void seed(const std::vector<int>& bt, int u, int* pos, double* val, int* cnt) const;
  // affecte_tau_val_domain (domain_shell_affecte_tau.cpp:26-130) as a sparse linear map
  // coefficient <- unknowns: for each coefficient index, (unknown, weight) pairs.
  void affecte_tau_map(const std::vector<int>& bt, std::vector<int>& rowptr, std::vector<int>& col,
                       std::vector<double>& w, int* consumed) const;
  // Row recipes (order = 2 for Eq_inside; boundary version for Eq_bc / Eq_matching).
  std::vector<TauRow> tau_rows_inside(const std::vector<int>& bt, int order) const;
  std::vector<TauRow> tau_rows_boundary(const std::vector<int>& bt) const;

 private:
  bool true_tet(int baset, int j, int k, int kmin, bool quirk) const;
  void row_lines(const std::vector<int>& bt, std::vector<std::tuple<int, int, int, double>>& lines) const;
};

}  // namespace kadhost
