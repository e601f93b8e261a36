// The XCTS black-hole system on a spherical space, traced into device programs.
//
// Equations: the Kerr prototype of codes/Par_version/Kerr/kerr.cpp:69-99 in the
// Cartesian tensor basis of the FUKA BH solver (bh_3d_xcts_solver_imp.cpp:119-159):
//   A^ij = (D^i bet^j + D^j bet^i - 2/3 D_k bet^k f^ij) / 2 / N
//   inner: N = n0 ; bet^i = n0 / P^2 s^i + Ome a m^i ;
//          dn(P) + 0.5 P / a + P^3 A_ij s^i s^j / 4 = 0
//   D_i D^i N + 2 D_i P D^i N / P - N P^4 A_ij A^ij = 0   (matching N, dn(N))
//   D_i D^i P + P^5 A_ij A^ij / 8 = 0                    (matching P, dn(P))
//   D_j A^ij + 6 A^ij D_j P / P = 0                       (matching bet^i, dn(bet^i))
//   outer: N = 1, P = 1, bet^i = 0
// Equation, unknown and row order as System_of_eqs (src/System_of_eqs/solver.cpp,
// system_of_eqs.cpp) with Space_spheric::add_eq / add_inner_bc / add_outer_bc
// (src/Space/Spheric/spheric_add_eq.cpp:26-52).
#pragma once
#include <vector>
#include <string>
#include <array>
#include "system_base.hpp"

namespace kadhost {

struct XctsParams {
  int nr = 9, nt = 7, np = 4;
  std::vector<double> bounds;  // Space_spheric bounds: nucleus [0, b0], shells, compact [b_last, inf)
  int dmin = 1, dmax = 3;
  double n0 = 0.5, a = 1.0, ome = 0.0;
  // 0: reference order (unknowns variable-major as system_of_eqs.cpp:597-612, rows in
  //    equation order as kerr.cpp:84-99); 1: domain-major (unknowns by domain, then
  //    field; rows by the domain an equation lives in, matching d|d+1 with d+1), which
  //    makes J block-tridiagonal by domain
  int order = 0;
};

class XctsSystem : public SystemBase {
 public:
  static constexpr int NF = 5;  // P, N, bet^x, bet^y, bet^z
  XctsParams prm;
  std::vector<std::array<int, 2>> tprog_fd;

  explicit XctsSystem(const XctsParams& p);
  Base field_base(int f, int d) const override { return f == 4 ? doms[d].anti_base() : doms[d].std_base(); }
  // value slot of the configuration values of a cached constant field (-1 if absent)
  int const_slot(const std::string& name, int d) const {
    auto it = consts.find(name + ":" + std::to_string(d));
    return it == consts.end() ? -1 : it->second.val->c;
  }
  int param_slot(const std::string& name, int d) const override { return const_slot(name, d); }
  void build() override;

 private:
  struct Defs {
    int d;
    TE P, N, bet[3];
    TE A[3][3];
    TE AA;
    TE ome;
  };
  std::vector<std::array<VDp, NF>> field_nodes;  // value nodes per domain
  std::map<std::string, TE> consts;              // constant fields, created once per domain
  TE cached_const(int d, const std::string& name, const std::vector<double>& v, const Base& b);

  TE t_add(const TE& a, const TE& b);
  TE t_sub(const TE& a, const TE& b);
  TE t_mul(const TE& a, const TE& b);
  TE t_div(const TE& a, const TE& b);
  TE t_scal(const TE& a, double x);
  TE t_div_scal(const TE& a, double x);
  TE t_pow(const TE& a, int n, int d);
  void t_grad(const TE& a, TE out[3]);
  TE t_dn(const TE& a, int bound);
  TE t_const(int d, const std::vector<double>& v, const Base& b);
  TE lap(const TE& a);

  Defs defs(int d, const std::array<VDp, NF>& ders);
  std::vector<TE> expr(const EqRec& e, int part, Defs& D);
  void trace_value();
  void reorder_rows();
  void trace_observables();
  void trace_obs_tangents();
  void trace_ome_tangent();
  VDp obs_integrand(int q, Defs& D, const TE sm[3], const TE mh[3], bool der);
  std::map<std::pair<int, int>, int> seed_tab_of;  // (field, domain) -> seed table
  VDp ome_seed = nullptr;                          // tangent of Ome while tracing ome_prog
  void unit_vectors(int d, TE sm[3], TE mg[3], bool unit_m);
  void trace_tangent(int f, int d);
  std::vector<std::array<VDp, NF>> zero_ders();
};

}  // namespace kadhost
