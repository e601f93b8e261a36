// Spherical domains: geometry, bases, tau bookkeeping.  See domain.hpp.
#include "domain.hpp"
#include <cmath>
#include <stdexcept>
#include <tuple>
#include <map>

namespace kadhost {

Domain::Domain(int num_, bool compact_, int nr_, int nt_, int np_, double rint, double rext)
    : num(num_), compact(compact_), nr(nr_), nt(nt_), np(np_), npc(np_ + 2) {
  if (!compact) {
    alpha = (rext - rint) / 2.0;  // domain_shell.cpp:28-29
    beta = (rext + rint) / 2.0;
  } else {
    alpha = -0.5 / rint;  // domain_compact.cpp:28
  }
  xr = coloc_r(nr);
  th = coloc_theta(nt);
  ph = coloc_phi(np);
  R.resize((size_t)conf_size());
  for (int k = 0; k < np; ++k)
    for (int j = 0; j < nt; ++j)
      for (int i = 0; i < nr; ++i) {
        const double r = compact ? 1.0 / (alpha * (xr[i] - 1.0)) : alpha * xr[i] + beta;
        R[(size_t)i + nr * (j + (size_t)nt * k)] = r;
      }
}

Base Domain::std_base() const {
  Base b;
  b.def = true;
  b.bp = COSSIN;
  b.bt.resize(npc);
  for (int k = 0; k < npc; ++k) b.bt[k] = ((k / 2) % 2 == 0) ? COS_EVEN : SIN_ODD;
  b.br.assign((size_t)npc * nt, CHEB);
  return b;
}

Base Domain::anti_base() const {
  Base b = std_base();
  for (int k = 0; k < npc; ++k) b.bt[k] = ((k / 2) % 2 == 0) ? COS_ODD : SIN_EVEN;
  return b;
}

// theta rule of Domain_shell::mult (domain_shell.cpp:722-798): result for k%4 < 2, else.
static std::pair<int, int> mult_t(int a, int b) {
  static const std::map<std::pair<int, int>, std::pair<int, int>> t = {
      {{COS_EVEN, COS_EVEN}, {COS_EVEN, SIN_ODD}}, {{COS_EVEN, COS_ODD}, {COS_ODD, SIN_EVEN}},
      {{COS_EVEN, SIN_EVEN}, {SIN_EVEN, COS_ODD}}, {{COS_EVEN, SIN_ODD}, {SIN_ODD, COS_EVEN}},
      {{COS_ODD, COS_EVEN}, {COS_ODD, SIN_EVEN}},  {{COS_ODD, COS_ODD}, {COS_EVEN, SIN_ODD}},
      {{COS_ODD, SIN_EVEN}, {SIN_ODD, COS_EVEN}},  {{COS_ODD, SIN_ODD}, {SIN_EVEN, COS_ODD}},
      {{SIN_EVEN, COS_EVEN}, {SIN_EVEN, COS_ODD}}, {{SIN_EVEN, COS_ODD}, {SIN_ODD, COS_EVEN}},
      {{SIN_EVEN, SIN_EVEN}, {COS_EVEN, SIN_ODD}}, {{SIN_EVEN, SIN_ODD}, {COS_ODD, SIN_EVEN}},
      {{SIN_ODD, COS_EVEN}, {SIN_ODD, COS_EVEN}},  {{SIN_ODD, COS_ODD}, {SIN_EVEN, COS_ODD}},
      {{SIN_ODD, SIN_EVEN}, {COS_ODD, SIN_EVEN}},  {{SIN_ODD, SIN_ODD}, {COS_EVEN, SIN_ODD}}};
  auto it = t.find({a, b});
  if (it == t.end()) return {-1, -1};
  return it->second;
}

Base Domain::mult_base(const Base& a, const Base& b) const {
  Base r;
  if (!a.def || !b.def || a.bp != COSSIN || b.bp != COSSIN) return r;
  r.bp = COSSIN;
  r.bt.resize(npc);
  for (int k = 0; k < npc; ++k) {
    auto p = mult_t(a.bt[k], b.bt[k]);
    if (p.first < 0) return Base();
    r.bt[k] = (k % 4 < 2) ? p.first : p.second;
  }
  for (size_t q = 0; q < a.br.size(); ++q)
    if (a.br[q] != CHEB || b.br[q] != CHEB) return Base();
  r.br.assign((size_t)npc * nt, CHEB);
  r.def = true;
  return r;
}

bool Domain::true_tet(int baset, int j, int k, int kmin, bool quirk) const {
  switch (baset) {
    case COS_EVEN: return !(j == 0 && k >= kmin);
    case COS_ODD: return !(j == nt - 1 || (j == 0 && k >= kmin));
    case SIN_EVEN: return !((j == 1 && k >= kmin + 2) || j == 0 || j == nt - 1);
    case SIN_ODD: {
      // affecte_tau uses k > kmin+2 (domain_shell_affecte_tau.cpp:58); nbr_unknowns and
      // affecte_tau_one_coef use k >= kmin+2.
      const bool cond = quirk ? (k > kmin + 2) : (k >= kmin + 2);
      return !((j == 0 && cond) || j == nt - 1);
    }
    default: throw std::runtime_error("unknown theta basis in tau bookkeeping");
  }
}

std::vector<std::pair<int, int>> Domain::unknown_slots(const std::vector<int>& bt) const {
  const int kmin = 2;
  std::vector<std::pair<int, int>> out;
  for (int k = 0; k < npc - 1; ++k) {
    if (k == 1) continue;
    for (int j = 0; j < nt; ++j)
      if (true_tet(bt[k], j, k, kmin, false)) out.emplace_back(k, j);
  }
  return out;
}

void Domain::seed(const std::vector<int>& bt, int u, int* pos, double* val, int* cnt) const {
  const int kmin = 2;
  const auto slots = unknown_slots(bt);
  const int s = u / nr, i = u % nr;
  const int k = slots[s].first, j = slots[s].second;
  const int baset = bt[k];
  auto idx = [&](int ii, int jj, int kk) { return ii + nr * (jj + nt * kk); };
  int c = 0;
  pos[c] = idx(i, j, k);
  val[c++] = 1.0;
  if ((baset == COS_EVEN || baset == COS_ODD) && k >= kmin) {
    pos[c] = idx(i, 0, k);
    val[c++] = -1.0;
  }
  if (baset == SIN_EVEN && k >= kmin + 2) {
    pos[c] = idx(i, 1, k);
    val[c++] = -(double)j;
  }
  if (baset == SIN_ODD && k >= kmin + 2) {
    pos[c] = idx(i, 0, k);
    val[c++] = -(double)(2 * j + 1);
  }
  *cnt = c;
}

void Domain::affecte_tau_map(const std::vector<int>& bt, std::vector<int>& rowptr, std::vector<int>& col,
                             std::vector<double>& w, int* consumed) const {
  const int kmin = 2;
  const int ncf = coef_size();
  std::vector<std::vector<std::pair<int, double>>> rows(ncf);
  auto idx = [&](int ii, int jj, int kk) { return ii + nr * (jj + nt * kk); };
  int pos = 0;
  for (int k = 0; k < npc - 1; ++k) {
    if (k == 1) continue;
    const int baset = bt[k];
    for (int j = 0; j < nt; ++j)
      if (true_tet(baset, j, k, kmin, true))
        for (int i = 0; i < nr; ++i) rows[idx(i, j, k)].push_back({pos++, 1.0});
  }
  // regularisation (domain_shell_affecte_tau.cpp:76-128): the partner coefficient is
  // minus a weighted sum of the others on the same (i, k) line.
  for (int k = 0; k < npc - 1; ++k) {
    const int baset = bt[k];
    int jt = -1, j0 = 0;
    std::vector<double> wj(nt, 0.0);
    if ((baset == COS_EVEN || baset == COS_ODD) && k >= kmin) {
      jt = 0; j0 = 1;
      for (int j = 1; j < nt; ++j) wj[j] = 1.0;
    } else if (baset == SIN_EVEN && k >= kmin + 2) {
      jt = 1; j0 = 2;
      for (int j = 2; j < nt; ++j) wj[j] = j;
    } else if (baset == SIN_ODD && k >= kmin + 2) {
      jt = 0; j0 = 1;
      for (int j = 1; j < nt; ++j) wj[j] = 2 * j + 1;
    }
    if (jt < 0) continue;
    for (int i = 0; i < nr; ++i) {
      std::vector<std::pair<int, double>> acc;
      for (int j = j0; j < nt; ++j)
        for (auto& e : rows[idx(i, j, k)]) acc.push_back({e.first, -wj[j] * e.second});
      rows[idx(i, jt, k)] = acc;  // "so.cf->set(pos_cf) = -sum" overwrites
    }
  }
  rowptr.assign(ncf + 1, 0);
  col.clear();
  w.clear();
  for (int q = 0; q < ncf; ++q) {
    for (auto& e : rows[q]) {
      col.push_back(e.first);
      w.push_back(e.second);
    }
    rowptr[q + 1] = (int)col.size();
  }
  *consumed = pos;
}

void Domain::row_lines(const std::vector<int>& bt, std::vector<std::tuple<int, int, int, double>>& lines) const {
  const int kmin = 2;
  lines.clear();
  for (int k = 0; k < npc - 1; ++k) {
    if (k == 1) continue;
    const int baset = bt[k];
    for (int j = 0; j < nt; ++j) {
      switch (baset) {
        case COS_EVEN:
          if (k < kmin) lines.emplace_back(k, j, -1, 0.0);
          else if (j != 0) lines.emplace_back(k, j, 0, 2.0);
          break;
        case COS_ODD:
          if (j != nt - 1) {
            if (k < kmin) lines.emplace_back(k, j, -1, 0.0);
            else if (j != 0) lines.emplace_back(k, j, 0, 1.0);
          }
          break;
        case SIN_EVEN:
          if (j != 0 && j != nt - 1) {
            if (k < kmin + 2) lines.emplace_back(k, j, -1, 0.0);
            else if (j != 1) lines.emplace_back(k, j, 1, (double)j);
          }
          break;
        case SIN_ODD:
          if (j != nt - 1) {
            if (k < kmin + 2) lines.emplace_back(k, j, -1, 0.0);
            else if (j != 0) lines.emplace_back(k, j, 0, (double)(2 * j + 1));
          }
          break;
        default:
          throw std::runtime_error("unknown theta basis in export_tau");
      }
    }
  }
}

std::vector<TauRow> Domain::tau_rows_inside(const std::vector<int>& bt, int order) const {
  std::vector<std::tuple<int, int, int, double>> lines;
  row_lines(bt, lines);
  std::vector<TauRow> rows;
  for (auto& L : lines) {
    const int k = std::get<0>(L), j = std::get<1>(L), gj = std::get<2>(L);
    const double g = std::get<3>(L);
    for (int i = 0; i < nr - order; ++i) {
      TauRow r;
      r.pos = i + nr * (j + nt * k);
      r.gpos = gj < 0 ? -1 : i + nr * (gj + nt * k);
      r.gw = g;
      rows.push_back(r);
    }
  }
  return rows;
}

std::vector<TauRow> Domain::tau_rows_boundary(const std::vector<int>& bt) const {
  std::vector<std::tuple<int, int, int, double>> lines;
  row_lines(bt, lines);
  std::vector<TauRow> rows;
  for (auto& L : lines) {
    const int k = std::get<0>(L), j = std::get<1>(L), gj = std::get<2>(L);
    TauRow r;
    r.pos = j + nt * k;
    r.gpos = gj < 0 ? -1 : gj + nt * k;
    r.gw = std::get<3>(L);
    rows.push_back(r);
  }
  return rows;
}

}  // namespace kadhost
