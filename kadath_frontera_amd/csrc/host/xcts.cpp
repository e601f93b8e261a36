// XCTS system traced into device programs.  See xcts.hpp.
#include "xcts.hpp"
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <map>

namespace kadhost {

XctsSystem::XctsSystem(const XctsParams& p) : prm(p) {
  const int ndom = (int)p.bounds.size() + 1;
  if (p.dmin < 1 || p.dmax != ndom - 1 || p.dmin > p.dmax)
    throw std::runtime_error("XctsSystem: solve from the first shell to the compactified domain");
  dmin = p.dmin;
  dmax = p.dmax;
  nfields = NF;
  field_names = {"P", "N", "bet^x", "bet^y", "bet^z"};
  var_fields = {{0}, {1}, {2, 3, 4}};
  doms.resize(ndom);
  // Space_spheric (space_spheric.cpp:25-47); the nucleus is excised and not built.
  for (int i = 1; i < ndom - 1; ++i) doms[i] = Domain(i, false, p.nr, p.nt, p.np, p.bounds[i - 1], p.bounds[i]);
  doms[ndom - 1] = Domain(ndom - 1, true, p.nr, p.nt, p.np, p.bounds.back(), 0.0);
}

// ---------------------------------------------------------------------------
// Term_eq rules (src/Term_eq/term_eq_math.cpp)
// ---------------------------------------------------------------------------
TE XctsSystem::t_add(const TE& a, const TE& b) { return {tr.add(a.val, b.val), tr.add(a.der, b.der)}; }
TE XctsSystem::t_sub(const TE& a, const TE& b) { return {tr.sub(a.val, b.val), tr.sub(a.der, b.der)}; }
// der = a' b + a b'   (term_eq_math.cpp:248-342)
TE XctsSystem::t_mul(const TE& a, const TE& b) {
  return {tr.mul(a.val, b.val), tr.add(tr.mul(a.der, b.val), tr.mul(a.val, b.der))};
}
// der = (a' b - a b') / (b b)   (term_eq_math.cpp:344-441)
TE XctsSystem::t_div(const TE& a, const TE& b) {
  return {tr.div(a.val, b.val),
          tr.div(tr.sub(tr.mul(a.der, b.val), tr.mul(a.val, b.der)), tr.mul(b.val, b.val))};
}
TE XctsSystem::t_scal(const TE& a, double x) { return {tr.scal(a.val, x), tr.scal(a.der, x)}; }
TE XctsSystem::t_div_scal(const TE& a, double x) { return {tr.div_scal(a.val, x), tr.div_scal(a.der, x)}; }
// s^i radial unit vector and m^i = (-y, x, 0) (the rotation generator) of domain d as
// constant fields; x/y components standard base, z anti-symmetric, m^z = 0.  With
// unit_m, m / r = sin(theta) (-sin(phi), cos(phi), 0) instead (bounded at infinity).
void XctsSystem::unit_vectors(int d, TE sm[3], TE mg[3], bool unit_m) {
  const Domain& dm = doms[d];
  std::vector<double> sx(dm.conf_size()), sy(dm.conf_size()), sz(dm.conf_size()), mx(dm.conf_size()),
      my(dm.conf_size());
  for (int k = 0; k < dm.np; ++k)
    for (int j = 0; j < dm.nt; ++j)
      for (int i = 0; i < dm.nr; ++i) {
        const size_t q = (size_t)i + dm.nr * (j + (size_t)dm.nt * k);
        const double st = std::sin(dm.th[j]), ct = std::cos(dm.th[j]), cp = std::cos(dm.ph[k]),
                     sp = std::sin(dm.ph[k]);
        sx[q] = st * cp;
        sy[q] = st * sp;
        sz[q] = ct;
        const double rf = unit_m ? 1.0 : dm.R[q];
        mx[q] = -rf * st * sp;
        my[q] = rf * st * cp;
      }
  sm[0] = cached_const(d, "sx", sx, dm.std_base());
  sm[1] = cached_const(d, "sy", sy, dm.std_base());
  sm[2] = cached_const(d, "sz", sz, dm.anti_base());
  mg[0] = cached_const(d, unit_m ? "mhx" : "mx", mx, dm.std_base());
  mg[1] = cached_const(d, unit_m ? "mhy" : "my", my, dm.std_base());
  mg[2] = TE{tr.zero(d, SP_VAL), tr.zero(d, SP_TAN)};
}

TE XctsSystem::t_const(int d, const std::vector<double>& v, const Base& b) {
  return {tr.const_conf(d, v, b), tr.zero(d, SP_TAN)};
}
TE XctsSystem::cached_const(int d, const std::string& name, const std::vector<double>& v, const Base& b) {
  const std::string key = name + ":" + std::to_string(d);
  auto it = consts.find(key);
  if (it != consts.end()) return TE{it->second.val, tr.zero(d, SP_TAN)};
  TE t = t_const(d, v, b);
  consts[key] = t;
  return t;
}
// pow(Term_eq, int) (term_eq_math.cpp:547-567): repeated products for n > 0
TE XctsSystem::t_pow(const TE& a, int n, int d) {
  if (n <= 0) throw std::runtime_error("t_pow: n > 0 only in this system");
  (void)d;
  TE r = a;
  for (int i = 1; i < n; ++i) r = t_mul(r, a);
  return r;
}
// Metric_flat::derive_partial_cart -> Domain::partial_cart (metric_flat.cpp:150-195, domain.cpp:408-451)
void XctsSystem::t_grad(const TE& a, TE out[3]) {
  VDp gv[3], gd[3];
  tr.der_abs(a.val, gv);
  tr.der_abs(a.der, gd);
  for (int i = 0; i < 3; ++i) out[i] = {gv[i], gd[i]};
}
TE XctsSystem::t_dn(const TE& a, int bound) { return {tr.der_normal(a.val, bound), tr.der_normal(a.der, bound)}; }
TE XctsSystem::lap(const TE& a) {
  TE g[3];
  t_grad(a, g);
  TE r;
  for (int i = 0; i < 3; ++i) {
    TE gg[3];
    t_grad(g[i], gg);
    r = (i == 0) ? gg[i] : t_add(r, gg[i]);
  }
  return r;
}

XctsSystem::Defs XctsSystem::defs(int d, const std::array<VDp, NF>& ders) {
  Defs D;
  D.d = d;
  D.P = {field_nodes[d][0], ders[0]};
  D.N = {field_nodes[d][1], ders[1]};
  for (int i = 0; i < 3; ++i) D.bet[i] = {field_nodes[d][2 + i], ders[2 + i]};
  TE Dbet[3][3];  // Dbet[j][i] = d_i bet^j
  for (int j = 0; j < 3; ++j) t_grad(D.bet[j], Dbet[j]);
  TE divb = t_add(t_add(Dbet[0][0], Dbet[1][1]), Dbet[2][2]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      TE s = t_add(Dbet[j][i], Dbet[i][j]);
      if (i == j) s = t_sub(s, t_scal(divb, 2.0 / 3.0));
      D.A[i][j] = t_div(t_div_scal(s, 2.0), D.N);
    }
  bool first = true;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      TE p = t_mul(D.A[i][j], D.A[i][j]);
      D.AA = first ? p : t_add(D.AA, p);
      first = false;
    }
  return D;
}

std::vector<TE> XctsSystem::expr(const EqRec& e, int part, Defs& D) {
  const int d = D.d;
  Domain& dm = doms[d];
  const std::string& w = e.which;
  if (e.kind == 1) {  // matching: part 0 at OUTER_BC of d, part 1 at INNER_BC of d+1
    const int bound = part == 0 ? OUTER_BC : INNER_BC;
    if (w == "N") return {D.N};
    if (w == "dnN") return {t_dn(D.N, bound)};
    if (w == "P") return {D.P};
    if (w == "dnP") return {t_dn(D.P, bound)};
    if (w == "bet") return {D.bet[0], D.bet[1], D.bet[2]};
    if (w == "dnbet") return {t_dn(D.bet[0], bound), t_dn(D.bet[1], bound), t_dn(D.bet[2], bound)};
    throw std::runtime_error("unknown matching " + w);
  }
  if (e.kind == 0) {
    if (w == "N") {
      TE gP[3], gN[3];
      t_grad(D.P, gP);
      t_grad(D.N, gN);
      TE s = t_add(t_add(t_mul(gP[0], gN[0]), t_mul(gP[1], gN[1])), t_mul(gP[2], gN[2]));
      TE t2 = t_div(t_scal(s, 2.0), D.P);
      TE t3 = t_mul(t_mul(D.N, t_pow(D.P, 4, d)), D.AA);
      return {t_sub(t_add(lap(D.N), t2), t3)};
    }
    if (w == "P") return {t_add(lap(D.P), t_div_scal(t_mul(t_pow(D.P, 5, d), D.AA), 8.0))};
    TE gP[3];
    t_grad(D.P, gP);
    std::vector<TE> out;
    TE gA[3][3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) t_grad(D.A[i][j], gA[i][j]);
    for (int i = 0; i < 3; ++i) {
      TE divA = t_add(t_add(gA[i][0][0], gA[i][1][1]), gA[i][2][2]);
      TE s = t_add(t_add(t_mul(D.A[i][0], gP[0]), t_mul(D.A[i][1], gP[1])), t_mul(D.A[i][2], gP[2]));
      TE t2 = t_div(t_scal(s, 6.0), D.P);
      out.push_back(t_add(divA, t2));
    }
    return out;
  }
  if (e.kind == 2) {  // outer boundary: N = 1, P = 1, bet^i = 0
    if (w == "N") return {TE{tr.add_scal(D.N.val, -1.0), D.N.der}};
    if (w == "P") return {TE{tr.add_scal(D.P.val, -1.0), D.P.der}};
    return {D.bet[0], D.bet[1], D.bet[2]};
  }
  // inner boundary (excision)
  if (w == "N") return {TE{tr.add_scal(D.N.val, -prm.n0), D.N.der}};
  TE sm[3], mg[3];
  unit_vectors(d, sm, mg, false);
  if (w == "bet") {
    TE ome = cached_const(d, "ome", std::vector<double>(dm.conf_size(), prm.ome), dm.std_base());
    if (ome_seed && ome_seed->dom == d) ome.der = ome_seed;
    TE n0 = cached_const(d, "n0", std::vector<double>(dm.conf_size(), prm.n0), dm.std_base());
    TE nP2 = t_div(n0, t_pow(D.P, 2, d));
    std::vector<TE> out;
    for (int i = 0; i < 3; ++i) {
      TE rhs = t_mul(nP2, sm[i]);
      rhs = t_add(rhs, t_scal(t_mul(ome, mg[i]), prm.a));
      out.push_back(t_sub(D.bet[i], rhs));
    }
    return out;
  }
  // dn(P) + 0.5 * P / a + P^3 * A_ij * s^i * s^j / 4 = 0
  TE t1 = t_dn(D.P, INNER_BC);
  TE t2 = t_div_scal(t_scal(D.P, 0.5), prm.a);
  TE t3;
  bool first = true;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      TE p = t_mul(t_mul(D.A[i][j], sm[i]), sm[j]);
      t3 = first ? p : t_add(t3, p);
      first = false;
    }
  t3 = t_div_scal(t_mul(t_pow(D.P, 3, d), t3), 4.0);
  return {t_add(t_add(t1, t2), t3)};
}

std::vector<std::array<VDp, XctsSystem::NF>> XctsSystem::zero_ders() {
  std::vector<std::array<VDp, NF>> z(doms.size());
  for (size_t d = 0; d < doms.size(); ++d)
    for (int f = 0; f < NF; ++f) z[d][f] = tr.zero((int)d, SP_TAN);
  return z;
}

void XctsSystem::trace_value() {
  Program dummy;
  tr.tan = &dummy;
  field_nodes.assign(doms.size(), {});
  field_slot.assign(NF, std::vector<int>(doms.size(), -1));
  for (int d = prm.dmin; d <= prm.dmax; ++d)
    for (int f = 0; f < NF; ++f) field_nodes[d][f] = tr.field_coef(d, field_base(f, d), &field_slot[f][d]);
  auto zd = zero_ders();
  std::vector<Defs> D(doms.size());
  for (int d = prm.dmin; d <= prm.dmax; ++d) D[d] = defs(d, zd[d]);
  // equation list (kerr.cpp:84-99)
  eqs.clear();
  auto add_eq = [&](int kind, int d, const std::string& w, int nc) { eqs.push_back(EqRec{kind, d, w, nc, 0, {}}); };
  add_eq(3, prm.dmin, "N", 1);
  add_eq(3, prm.dmin, "bet", 3);
  add_eq(3, prm.dmin, "AH", 1);
  const char* names[3][3] = {{"N", "N", "dnN"}, {"P", "P", "dnP"}, {"bet", "bet", "dnbet"}};
  for (int q = 0; q < 3; ++q) {
    const int nc = q == 2 ? 3 : 1;
    for (int d = prm.dmin; d < prm.dmax; ++d) {
      add_eq(0, d, names[q][0], nc);
      add_eq(1, d, names[q][1], nc);
      add_eq(1, d, names[q][2], nc);
    }
    add_eq(0, prm.dmax, names[q][0], nc);
  }
  add_eq(2, prm.dmax, "N", 1);
  add_eq(2, prm.dmax, "P", 1);
  add_eq(2, prm.dmax, "bet", 3);

  int row = 0;
  for (auto& e : eqs) {
    e.row0 = row;
    std::vector<TE> p0 = expr(e, 0, D[e.d]);
    std::vector<TE> p1;
    if (e.kind == 1) p1 = expr(e, 1, D[e.d + 1]);
    for (int c = 0; c < e.ncomp; ++c) {
      int nr = 0;
      if (e.kind == 0) {
        if (tr.export_inside(p0[c].val, 2, row, 1.0, 0, &nr) < 0) throw std::runtime_error("zero residual");
      } else {
        const int bound = (e.kind == 3) ? INNER_BC : OUTER_BC;
        if (tr.export_boundary(p0[c].val, bound, row, 1.0, 0, &nr) < 0) throw std::runtime_error("zero residual");
        if (e.kind == 1) {
          int nr2 = 0;
          if (tr.export_boundary(p1[c].val, INNER_BC, row, -1.0, 1, &nr2) < 0 || nr2 != nr)
            throw std::runtime_error("matching row count mismatch");
        }
      }
      e.nrows.push_back(nr);
      row += nr;
    }
  }
  n_rows = row;
  tr.tan = nullptr;
}

void XctsSystem::trace_tangent(int f, int d) {
  tprog.emplace_back();
  tprog_fd.push_back({f, d});
  const int tab = (int)seeds.size();
  {
    SeedTab st;
    const auto bt = field_base(f, d).bt;
    const int cnt = doms[d].nbr_unknowns(bt);
    st.pos.assign((size_t)cnt * 3, 0);
    st.val.assign((size_t)cnt * 3, 0.0);
    st.cnt.assign(cnt, 0);
    for (int u = 0; u < cnt; ++u) doms[d].seed(bt, u, &st.pos[(size_t)u * 3], &st.val[(size_t)u * 3], &st.cnt[u]);
    seeds.push_back(st);
  }
  seed_tab_of[{f, d}] = tab;
  tr.begin_tangent(&tprog.back());
  auto zd = zero_ders();
  zd[d][f] = tr.seed_node(d, field_base(f, d), tab);
  std::vector<Defs> D(doms.size());
  D[d] = defs(d, zd[d]);
  for (auto& e : eqs) {
    int part = -1;
    if ((e.kind == 0 || e.kind == 2 || e.kind == 3) && e.d == d) part = 0;
    if (e.kind == 1 && e.d == d) part = 0;
    if (e.kind == 1 && e.d + 1 == d) part = 1;
    if (part < 0) continue;
    std::vector<TE> p = expr(e, part, D[d]);
    int row = e.row0;
    for (int c = 0; c < e.ncomp; ++c) {
      int nr = 0;
      if (e.kind == 0) {
        tr.export_inside(p[c].der, 2, row, 1.0, 0, &nr);
      } else {
        const int bound = (e.kind == 3 || part == 1) ? INNER_BC : OUTER_BC;
        tr.export_boundary(p[c].der, bound, row, part == 1 ? -1.0 : 1.0, 0, &nr);
      }
      if (nr != 0 && nr != e.nrows[c]) throw std::runtime_error("tangent row count differs from the residual's");
      row += e.nrows[c];
    }
  }
  tr.tan = nullptr;
}

// Domain-major rows: equation blocks move as a whole (their components stay
// contiguous), ordered by the domain they live in and, inside a domain, as in the
// reference order.  The value program's export rows are shifted accordingly;
// tangent programs are traced afterwards from the updated row0.
void XctsSystem::reorder_rows() {
  struct Blk {
    int old0, len, new0;
  };
  std::vector<Blk> blks;
  int row = 0;
  for (int dom = prm.dmin; dom <= prm.dmax; ++dom)
    for (auto& e : eqs) {
      const int rd = e.kind == 1 ? e.d + 1 : e.d;
      if (rd != dom) continue;
      int len = 0;
      for (int c : e.nrows) len += c;
      blks.push_back({e.row0, len, row});
      e.row0 = row;
      row += len;
    }
  if (row != n_rows) throw std::runtime_error("row reordering lost rows");
  std::sort(blks.begin(), blks.end(), [](const Blk& x, const Blk& y) { return x.old0 < y.old0; });
  for (auto& in : tr.val.code) {
    if (in.kind != I_EXPORT_IN && in.kind != I_EXPORT_BD) continue;
    auto it = std::upper_bound(blks.begin(), blks.end(), in.row0, [](int r, const Blk& b) { return r < b.old0; });
    if (it == blks.begin()) throw std::runtime_error("export row outside every equation block");
    --it;
    if (in.row0 >= it->old0 + it->len) throw std::runtime_error("export row outside its equation block");
    in.row0 += it->new0 - it->old0;
  }
}

void XctsSystem::build() {
  trace_value();
  if (prm.order == 1) reorder_rows();
  // unknown layout (system_of_eqs.cpp:597-612): P, N by domain; bet by domain with x, y, z
  // together -- or, domain-major, all five fields of a domain together
  layout.clear();
  tau_maps.clear();
  int off = 0;
  std::vector<std::pair<int, int>> order_fd;  // (field, domain) in column order
  if (prm.order == 1) {
    for (int d = prm.dmin; d <= prm.dmax; ++d)
      for (int f = 0; f < NF; ++f) order_fd.push_back({f, d});
  } else {
    const std::vector<std::vector<int>> vars = {{0}, {1}, {2, 3, 4}};
    for (auto& var : vars)
      for (int d = prm.dmin; d <= prm.dmax; ++d)
        for (int f : var) order_fd.push_back({f, d});
  }
  for (auto& fd : order_fd) {
    const int f = fd.first, d = fd.second;
    const auto bt = field_base(f, d).bt;
    const int cnt = doms[d].nbr_unknowns(bt);
    layout.push_back({f, d, off, cnt});
    Csr m;
    int used = 0;
    doms[d].affecte_tau_map(bt, m.rowptr, m.col, m.w, &used);
    if (used != cnt) throw std::runtime_error("affecte_tau consumes a different count than nbr_unknowns");
    tau_maps.push_back(m);
    off += cnt;
  }
  n_unknowns = off;
  if (n_unknowns != n_rows) throw std::runtime_error("Number of unknowns is different from the number of equations");
  for (auto& L : layout) trace_tangent(L[0], L[1]);
  trace_observables();
}

// Surface integrands of the global quantities at the outer boundary of the
// compactified domain (codes/Par_version/Kerr/reader.cpp:114-125), in the Cartesian
// basis: M_ADM = -integ(d_r P)/2pi, M_Komar = integ(d_r N)/4pi,
// J = -integ(A^ij s_i m_j)/8pi, where integ (Domain_compact::integ,
// domain_compact.cpp:89-121) takes r r sin(theta) so, evaluates it on the boundary and
// weighs the theta coefficients.  Traced into a program of its own (value space),
// run after the residual program on the current fields.
void XctsSystem::trace_observables() {
  const int d = prm.dmax;
  if (!doms[d].compact) throw std::runtime_error("observables need a compactified outer domain");
  Program saved = std::move(tr.val);
  tr.val = Program();
  auto zd = zero_ders();
  Defs D = defs(d, zd[d]);
  TE sm[3], mh[3];
  unit_vectors(d, sm, mh, true);
  const double fac[3] = {-1.0 / (2 * M_PI), 1.0 / (4 * M_PI), -1.0 / (8 * M_PI)};
  obs.clear();
  for (int q = 0; q < 3; ++q) {
    ObsRec r;
    r.fac = fac[q];
    VDp rr = obs_integrand(q, D, sm, mh, false);
    if (!rr->zero) {
      tr.coef(rr);
      r.slot = rr->cf;
      r.bt0 = rr->base.bt[0];
    }
    obs.push_back(r);
  }
  obs_prog = std::move(tr.val);
  tr.val = std::move(saved);
  trace_obs_tangents();
  trace_ome_tangent();
}

// r r sin(theta) so for quantity q (0 M_ADM: d_r P, 1 M_Komar: d_r N, 2 J_ADM:
// A^ij s_i m_j = r A^ij s_i mh_j) of the value (der = false) or tangent (der = true)
// fields.  d_r of the compactified domain is der_normal there; its mult_r is div_xm1 /
// alpha (domain_compact_ope.cpp:173-189).
VDp XctsSystem::obs_integrand(int q, Defs& D, const TE sm[3], const TE mh[3], bool der) {
  const int d = prm.dmax;
  Domain& dm = doms[d];
  VDp so;
  if (q == 0) {
    so = tr.der_normal(der ? D.P.der : D.P.val, OUTER_BC);
  } else if (q == 1) {
    so = tr.der_normal(der ? D.N.der : D.N.val, OUTER_BC);
  } else {
    so = nullptr;
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 2; ++k) {
        VDp t = tr.mul(tr.mul(der ? D.A[i][k].der : D.A[i][k].val, sm[i].val), mh[k].val);
        so = so ? tr.add(so, t) : t;
      }
    if (!so->zero) so = tr.scal(tr.div_xm1(so), 1.0 / dm.alpha);
  }
  if (so->zero) return so;
  VDp rr = tr.mult_sin_theta(so);
  for (int t = 0; t < 2; ++t) rr = tr.scal(tr.div_xm1(rr), 1.0 / dm.alpha);
  return rr;
}

// d(M_ADM, M_Komar, J_ADM)/d(unknowns) of the outer domain's fields (the quantities
// read no other domain): the Jacobian's seeds, the observables' tangents, one
// integral per quantity and column.
void XctsSystem::trace_obs_tangents() {
  const int d = prm.dmax;
  obs_tprog.clear();
  obs_tprog_lay.clear();
  obs_tprog_tab.clear();
  TE sm[3], mh[3];
  unit_vectors(d, sm, mh, true);
  const double fac[3] = {-1.0 / (2 * M_PI), 1.0 / (4 * M_PI), -1.0 / (8 * M_PI)};
  for (auto& L : layout) {
    if (L[1] != d) continue;
    const int f = L[0];
    obs_tprog.emplace_back();
    obs_tprog_lay.push_back(L);
    const int tab = seed_tab_of.at({f, d});
    obs_tprog_tab.push_back(tab);
    tr.begin_tangent(&obs_tprog.back());
    auto zd = zero_ders();
    zd[d][f] = tr.seed_node(d, field_base(f, d), tab);
    Defs D = defs(d, zd[d]);
    for (int q = 0; q < 3; ++q) tr.export_integral(obs_integrand(q, D, sm, mh, true), q, fac[q]);
    tr.tan = nullptr;
  }
}

// dF/dOme: Ome enters the inner boundary condition of the shift,
// bet^i = n0 / P^2 s^i + a Ome m^i (kerr.cpp:91-95); its tangent is the constant 1,
// a single coefficient (T_0 cos(0 theta) a_0) seeded by a one-entry table.
void XctsSystem::trace_ome_tangent() {
  const int d = prm.dmin;
  SeedTab st;
  st.pos = {0, 0, 0};
  st.val = {1.0, 0.0, 0.0};
  st.cnt = {1};
  par_names = {"ome"};
  par_tab = {(int)seeds.size()};
  seeds.push_back(st);
  par_prog.assign(1, Program());
  tr.begin_tangent(&par_prog[0]);
  ome_seed = tr.seed_node(d, doms[d].std_base(), par_tab[0]);
  auto zd = zero_ders();
  Defs D = defs(d, zd[d]);
  for (auto& e : eqs) {
    if (e.kind != 3 || e.d != d || e.which != "bet") continue;
    std::vector<TE> p = expr(e, 0, D);
    int row = e.row0;
    for (int c = 0; c < e.ncomp; ++c) {
      int nr = 0;
      tr.export_boundary(p[c].der, INNER_BC, row, 1.0, 0, &nr);
      row += e.nrows[c];
    }
  }
  ome_seed = nullptr;
  tr.tan = nullptr;
}

}  // namespace kadhost
