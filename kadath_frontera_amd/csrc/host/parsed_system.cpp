// Newton systems parsed from Kadath equation strings.  See parsed_system.hpp.
#include "parsed_system.hpp"
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace kadhost {

namespace {

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

bool balanced(const std::string& s) {
  int depth = 0;
  for (char c : s) {
    if (c == '(') ++depth;
    if (c == ')' && --depth < 0) return false;
  }
  return depth == 0;
}

// Split at the last occurrence of op whose both sides are non-empty and balanced
// (System_of_eqs::is_ope_bin with get_parts, give_ope.cpp:276-291,
// src/Utilities/name_tools.cpp:150-184).
bool split_last(const std::string& s, char op, std::string& l, std::string& r) {
  for (size_t p = s.size(); p-- > 0;) {
    if (s[p] != op) continue;
    std::string a = trim(s.substr(0, p)), b = trim(s.substr(p + 1));
    if (!a.empty() && !b.empty() && balanced(a) && balanced(b)) {
      l = a;
      r = b;
      return true;
    }
  }
  return false;
}

// "name(arg)" with the parentheses enclosing the rest of the string
bool unary(const std::string& s, const std::string& name, std::string& arg) {
  if (s.size() < name.size() + 2 || s.compare(0, name.size(), name) != 0) return false;
  size_t p = name.size();
  while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
  if (p >= s.size() || s[p] != '(' || s.back() != ')') return false;
  std::string in = s.substr(p + 1, s.size() - p - 2);
  if (!balanced(in)) return false;
  arg = trim(in);
  return true;
}

bool is_number(const std::string& s, double& v) {
  if (s.empty()) return false;
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == '\0' && (std::isdigit((unsigned char)s[0]) || s[0] == '.');
}

int ipow3(int v) {
  int r = 1;
  for (int i = 0; i < v; ++i) r *= 3;
  return r;
}

// "name^ij" / "name_ij" / "name": base name and index letters (positions ignored:
// flat metric, Cartesian basis)
void split_indices(const std::string& s, std::string& base, std::string& idx) {
  size_t p = s.find_first_of("^_");
  base = s.substr(0, p);
  idx.clear();
  if (p == std::string::npos) return;
  for (size_t q = p; q < s.size(); ++q) {
    const char c = s[q];
    if (c == '^' || c == '_') continue;
    if (!std::isalpha((unsigned char)c)) throw std::runtime_error("bad index in '" + s + "'");
    idx.push_back(c);
  }
}

}  // namespace

ParsedSystem::ParsedSystem(int nr, int nt, int np, const std::vector<double>& bounds, int dmin_, int dmax_,
                           int order) {
  const int ndom = (int)bounds.size() + 1;
  if (nr < 3 || nt < 3 || np < 4 || np % 2 || dmin_ < 1 || dmax_ != ndom - 1 || dmin_ > dmax_)
    throw std::runtime_error("ParsedSystem: Space_spheric solved from a shell to the compactified domain");
  for (size_t i = 1; i < bounds.size(); ++i)
    if (!(bounds[i] > bounds[i - 1])) throw std::runtime_error("ParsedSystem: bounds must increase");
  if (order != 0 && order != 1) throw std::runtime_error("ParsedSystem: order 0 or 1");
  dmin = dmin_;
  dmax = dmax_;
  order_mode = order;
  doms.resize(ndom);
  for (int i = 1; i < ndom - 1; ++i) doms[i] = Domain(i, false, nr, nt, np, bounds[i - 1], bounds[i]);
  doms[ndom - 1] = Domain(ndom - 1, true, nr, nt, np, bounds.back(), 0.0);
}

void ParsedSystem::add_var(const std::string& name, int valence) {
  if (valence != 0 && valence != 1) throw std::runtime_error("add_var: scalars and Cartesian vectors only");
  for (auto& v : vars)
    if (v.name == name) throw std::runtime_error("add_var: '" + name + "' already declared");
  Var v{name, valence, nfields};
  std::vector<int> comps;
  const int nc = ipow3(valence);
  for (int c = 0; c < nc; ++c) {
    field_var.push_back((int)vars.size());
    field_comp.push_back(c);
    field_names.push_back(valence == 0 ? name : name + "^" + "xyz"[c]);
    comps.push_back(nfields++);
  }
  var_fields.push_back(comps);
  vars.push_back(v);
}

void ParsedSystem::add_cst(const std::string& name, double value) {
  Cst c;
  c.name = name;
  c.value = value;
  csts.push_back(c);
}

int ParsedSystem::add_param(const std::string& name, double value) {
  Cst c;
  c.name = name;
  c.value = value;
  c.param = true;
  c.par = (int)par_names.size();
  par_names.push_back(name);
  csts.push_back(c);
  return c.par;
}

int ParsedSystem::add_var_double(const std::string& name, double value) {
  for (auto& c : csts)
    if (c.name == name) throw std::runtime_error("add_var: '" + name + "' already declared");
  Cst c;
  c.name = name;
  c.value = value;
  c.param = true;
  c.par = (int)par_names.size();
  c.dvar = ndouble++;
  par_names.push_back(name);
  dvar_par.push_back(c.par);
  csts.push_back(c);
  return c.dvar;
}

void ParsedSystem::add_eq_int(int dom, int bound, const std::string& eq) {
  if (dom < dmin || dom > dmax || (bound != INNER_BC && bound != OUTER_BC))
    throw std::runtime_error("add_eq_int: bad domain or boundary");
  EqRec e{4, dom, eq, 1, 0, {}};
  e.bound = bound;
  // integral equations keep their order among themselves and come first
  auto it = eqs.begin();
  while (it != eqs.end() && it->kind == 4) ++it;
  eqs.insert(it, e);
}

int ParsedSystem::add_global(const std::string& name, int dom, int bound, const std::string& expr) {
  if (dom < dmin || dom > dmax || (bound != INNER_BC && bound != OUTER_BC))
    throw std::runtime_error("add_global: bad domain or boundary");
  globs.push_back(Glob{name, expr, dom, bound});
  glob_names.push_back(name);
  return (int)globs.size() - 1;
}

void ParsedSystem::add_cst_field(const std::string& name, int valence, const std::vector<double>& conf) {
  if (valence != 0 && valence != 1) throw std::runtime_error("add_cst: scalar or Cartesian vector fields only");
  size_t need = 0;
  for (int d = dmin; d <= dmax; ++d) need += doms[d].conf_size();
  need *= ipow3(valence);
  if (conf.size() != need) throw std::runtime_error("add_cst: '" + name + "' needs " + std::to_string(need) + " values");
  Cst c;
  c.name = name;
  c.valence = valence;
  c.field = true;
  c.conf = conf;
  csts.push_back(c);
}

void ParsedSystem::set_flat_metric(const std::string& name) { metric = name; }

void ParsedSystem::add_def(const std::string& def) {
  std::string l, r;
  if (!split_last(def, '=', l, r)) throw std::runtime_error("add_def: no '=' in '" + def + "'");
  Def d;
  split_indices(trim(l), d.name, d.idx);
  d.rhs = r;
  defs.push_back(d);
}

void ParsedSystem::add_eq_inside(int dom, const std::string& eq, int order) {
  if (dom < dmin || dom > dmax || order < 0) throw std::runtime_error("add_eq_inside: bad domain or order");
  EqRec e{0, dom, eq, 0, 0, {}};
  e.order = order;
  eqs.push_back(e);
}

void ParsedSystem::add_eq_bc(int dom, int bound, const std::string& eq) {
  if (dom < dmin || dom > dmax || (bound != INNER_BC && bound != OUTER_BC))
    throw std::runtime_error("add_eq_bc: bad domain or boundary");
  EqRec e{bound == INNER_BC ? 3 : 2, dom, eq, 0, 0, {}};
  e.bound = bound;
  eqs.push_back(e);
}

void ParsedSystem::add_eq_matching(int dom, int bound, const std::string& expr) {
  // Eq_matching between dom and its outer neighbour (the Space_spheric case)
  if (bound != OUTER_BC || dom < dmin || dom >= dmax)
    throw std::runtime_error("add_eq_matching: OUTER_BC of a domain below dmax only");
  EqRec e{1, dom, expr, 0, 0, {}};
  e.bound = bound;
  eqs.push_back(e);
}

void ParsedSystem::space_add_eq(const std::string& eq, const std::string& m1, const std::string& m2) {
  for (int d = dmin; d <= dmax; ++d) {
    add_eq_inside(d, eq);
    if (d != dmax) {
      add_eq_matching(d, OUTER_BC, m1);
      add_eq_matching(d, OUTER_BC, m2);
    }
  }
}

Base ParsedSystem::field_base(int f, int d) const {
  const Var& v = vars[field_var[f]];
  if (v.valence == 1 && field_comp[f] == 2) return doms[d].anti_base();
  return doms[d].std_base();
}

int ParsedSystem::param_slot(const std::string& name, int d) const {
  auto it = const_nodes.find(name + ":0:" + std::to_string(d));
  return it == const_nodes.end() ? -1 : it->second->c;
}

VDp ParsedSystem::const_field(const std::string& key, int d, const std::vector<double>& conf, const Base& b,
                              bool keep_zero) {
  auto it = const_nodes.find(key);
  if (it != const_nodes.end()) return it->second;
  bool zero = !keep_zero;  // a parameter's field stays a field (set_param rewrites it)
  for (double x : conf) zero = zero && x == 0.0;
  VDp n = zero ? tr.zero(d, SP_VAL) : tr.const_conf(d, conf, b);
  const_nodes[key] = n;
  return n;
}

// ---------------------------------------------------------------------------
// scalar algebra: numbers stay numbers; a number meeting a field is applied with
// the scalar forms (Val_domain operator+(double) etc., val_domain_math.cpp)
// ---------------------------------------------------------------------------
TE ParsedSystem::as_field(const Sc& a, Ctx& cx) {
  if (!a.num) return a.te;
  const Domain& dm = doms[cx.d];
  char key[64];
  std::snprintf(key, sizeof(key), "#%.17g:%d", a.x, cx.d);
  VDp v = const_field(key, cx.d, std::vector<double>(dm.conf_size(), a.x), dm.std_base());
  return TE{v, tr.zero(cx.d, SP_TAN)};
}

ParsedSystem::Sc ParsedSystem::sc_add(const Sc& a, const Sc& b, bool sub, Ctx& cx) {
  Sc r;
  if (a.glob || b.glob) {
    // sums of integrals and numbers (Term_eq doubles, term_eq_math.cpp)
    if ((!a.glob && !a.num) || (!b.glob && !b.num))
      throw std::runtime_error("an integral added to a field");
    r.glob = true;
    r.x = sub ? a.x - b.x : a.x + b.x;
    r.gl = a.glob ? a.gl : std::vector<GlTerm>();
    if (b.glob)
      for (GlTerm t : b.gl) {
        if (sub) t.fac = -t.fac;
        r.gl.push_back(t);
      }
    return r;
  }
  if (a.num && b.num) {
    r.num = true;
    r.x = sub ? a.x - b.x : a.x + b.x;
    return r;
  }
  if (b.num) {  // field +- number
    r.te = TE{b.x == 0.0 ? a.te.val : tr.add_scal(a.te.val, sub ? -b.x : b.x), a.te.der};
    return r;
  }
  if (a.num) {  // number +- field
    VDp v = sub ? tr.neg(b.te.val) : b.te.val;
    r.te = TE{a.x == 0.0 ? v : tr.add_scal(v, a.x), sub ? tr.neg(b.te.der) : b.te.der};
    return r;
  }
  // Term_eq + / - (term_eq_math.cpp)
  r.te = sub ? TE{tr.sub(a.te.val, b.te.val), tr.sub(a.te.der, b.te.der)}
             : TE{tr.add(a.te.val, b.te.val), tr.add(a.te.der, b.te.der)};
  (void)cx;
  return r;
}

ParsedSystem::Sc ParsedSystem::sc_mul(const Sc& a, const Sc& b, Ctx& cx) {
  Sc r;
  if (a.glob || b.glob) {
    if (!(a.num || b.num)) throw std::runtime_error("integrals combine linearly with numbers only");
    r = a.glob ? a : b;
    const double x = a.glob ? b.x : a.x;
    r.x *= x;
    for (auto& t : r.gl) t.fac *= x;
    return r;
  }
  if (a.num && b.num) {
    r.num = true;
    r.x = a.x * b.x;
    return r;
  }
  if (a.num || b.num) {
    const Sc& f = a.num ? b : a;
    const double x = a.num ? a.x : b.x;
    if (x == 0.0) {
      r.num = true;
      r.x = 0.0;
      return r;
    }
    if (x == 1.0) return f;
    r.te = TE{tr.scal(f.te.val, x), tr.scal(f.te.der, x)};
    return r;
  }
  // der = a' b + a b'   (term_eq_math.cpp:248-342)
  r.te = TE{tr.mul(a.te.val, b.te.val), tr.add(tr.mul(a.te.der, b.te.val), tr.mul(a.te.val, b.te.der))};
  (void)cx;
  return r;
}

ParsedSystem::Sc ParsedSystem::sc_div(const Sc& a, const Sc& b, Ctx& cx) {
  Sc r;
  if (a.glob || b.glob) {
    if (!a.glob || !b.num || b.x == 0.0) throw std::runtime_error("integrals divide by non-zero numbers only");
    r = a;
    r.x /= b.x;
    for (auto& t : r.gl) t.fac /= b.x;
    return r;
  }
  if (b.num) {
    if (b.x == 0.0) throw std::runtime_error("division by the number 0");
    if (a.num) {
      r.num = true;
      r.x = a.x / b.x;
      return r;
    }
    r.te = TE{tr.div_scal(a.te.val, b.x), tr.div_scal(a.te.der, b.x)};
    return r;
  }
  const TE A = as_field(a, cx);
  const TE& B = b.te;
  // der = (a' b - a b') / (b b)   (term_eq_math.cpp:344-441)
  r.te = TE{tr.div(A.val, B.val), tr.div(tr.sub(tr.mul(A.der, B.val), tr.mul(A.val, B.der)), tr.mul(B.val, B.val))};
  return r;
}

// ---------------------------------------------------------------------------
// tensor algebra over named indices
// ---------------------------------------------------------------------------
// Sum over repeated index names (each name at most twice).
ParsedSystem::Tv ParsedSystem::contract(const Tv& a, Ctx& cx) {
  const int v = (int)a.idx.size();
  int p = -1, q = -1;
  for (int i = 0; i < v && p < 0; ++i)
    for (int j = i + 1; j < v; ++j)
      if (a.idx[i] == a.idx[j]) {
        p = i;
        q = j;
        break;
      }
  if (p < 0) return a;
  for (int j = q + 1; j < v; ++j)
    if (a.idx[j] == a.idx[p]) throw std::runtime_error(std::string("index '") + a.idx[p] + "' used more than twice");
  Tv r;
  for (int i = 0; i < v; ++i)
    if (i != p && i != q) r.idx.push_back(a.idx[i]);
  const int rv = v - 2;
  r.c.resize(ipow3(rv));
  for (int rc = 0; rc < (int)r.c.size(); ++rc) {
    // digits of the result component, then the full index with i at p and q
    std::vector<int> dig(rv);
    for (int t = rv - 1, x = rc; t >= 0; --t, x /= 3) dig[t] = x % 3;
    Sc acc;
    acc.num = true;
    acc.x = 0.0;
    for (int i = 0; i < 3; ++i) {
      int full = 0, t = 0;
      for (int s = 0; s < v; ++s) full = full * 3 + ((s == p || s == q) ? i : dig[t++]);
      acc = (i == 0) ? a.c[full] : sc_add(acc, a.c[full], false, cx);
    }
    r.c[rc] = acc;
  }
  return contract(r, cx);
}

ParsedSystem::Tv ParsedSystem::tv_add(const Tv& a, const Tv& b, bool sub, Ctx& cx) {
  // a number added to (or subtracted from) every component of a tensor
  if (b.idx.empty() && !a.idx.empty() && b.c[0].num) {
    Tv r = a;
    for (auto& x : r.c) x = sc_add(x, b.c[0], sub, cx);
    return r;
  }
  if (a.idx.empty() && !b.idx.empty() && a.c[0].num) {
    Tv r = b;
    for (auto& x : r.c) x = sc_add(a.c[0], x, sub, cx);
    return r;
  }
  std::string sa = a.idx, sb = b.idx;
  std::sort(sa.begin(), sa.end());
  std::sort(sb.begin(), sb.end());
  if (sa != sb) throw std::runtime_error("sum of tensors with different free indices ('" + a.idx + "', '" + b.idx + "')");
  Tv r;
  r.idx = a.idx;
  const int v = (int)a.idx.size();
  r.c.resize(a.c.size());
  for (int rc = 0; rc < (int)a.c.size(); ++rc) {
    std::vector<int> dig(v);
    for (int t = v - 1, x = rc; t >= 0; --t, x /= 3) dig[t] = x % 3;
    int bc = 0;
    for (int s = 0; s < v; ++s) bc = bc * 3 + dig[a.idx.find(b.idx[s])];
    r.c[rc] = sc_add(a.c[rc], b.c[bc], sub, cx);
  }
  return r;
}

ParsedSystem::Tv ParsedSystem::tv_mul(const Tv& a, const Tv& b, Ctx& cx) {
  Tv r;
  r.idx = a.idx + b.idx;
  r.c.resize(a.c.size() * b.c.size());
  for (size_t i = 0; i < a.c.size(); ++i)
    for (size_t j = 0; j < b.c.size(); ++j) r.c[i * b.c.size() + j] = sc_mul(a.c[i], b.c[j], cx);
  return contract(r, cx);
}

ParsedSystem::Tv ParsedSystem::tv_div(const Tv& a, const Tv& b, Ctx& cx) {
  if (!b.idx.empty()) throw std::runtime_error("division by a tensor");
  Tv r;
  r.idx = a.idx;
  for (auto& x : a.c) r.c.push_back(sc_div(x, b.c[0], cx));
  return r;
}

ParsedSystem::Tv ParsedSystem::tv_neg(const Tv& a, Ctx& cx) {
  Tv r = a;
  for (auto& x : r.c) {
    if (x.glob) {
      x.x = -x.x;
      for (auto& t : x.gl) t.fac = -t.fac;
    } else if (x.num) x.x = -x.x;
    else x.te = TE{tr.neg(x.te.val), tr.neg(x.te.der)};
  }
  (void)cx;
  return r;
}

// integer power by repeated products (pow(Term_eq, int), term_eq_math.cpp:547-567)
ParsedSystem::Tv ParsedSystem::tv_pow(const Tv& a, int n, Ctx& cx) {
  if (!a.idx.empty()) throw std::runtime_error("power of a tensor");
  if (n < 1) throw std::runtime_error("powers with a positive integer exponent only");
  Tv r = a;
  for (int i = 1; i < n; ++i) r.c[0] = sc_mul(r.c[0], a.c[0], cx);
  return r;
}

// D_i X: flat Cartesian gradient of every component (Metric_flat::derive_partial_cart
// -> Domain::partial_cart, metric_flat.cpp:150-195, domain.cpp:408-451); the new
// index comes first.
ParsedSystem::Tv ParsedSystem::tv_grad(const Tv& a, char name, Ctx& cx) {
  Tv r;
  r.idx = std::string(1, name) + a.idx;
  const size_t nc = a.c.size();
  r.c.resize(3 * nc);
  for (size_t q = 0; q < nc; ++q) {
    const Sc& x = a.c[q];
    if (x.num) {
      for (int i = 0; i < 3; ++i) {
        r.c[i * nc + q].num = true;
        r.c[i * nc + q].x = 0.0;
      }
      continue;
    }
    VDp gv[3], gd[3];
    tr.der_abs(x.te.val, gv);
    tr.der_abs(x.te.der, gd);
    for (int i = 0; i < 3; ++i) r.c[i * nc + q].te = TE{gv[i], gd[i]};
  }
  return contract(r, cx);
}

// Component-wise linear operators of one domain.
ParsedSystem::Tv ParsedSystem::tv_map(const Tv& a, const std::string& op, Ctx& cx) {
  Tv r = a;
  auto lin = [&](const VDp& v) -> VDp {
    if (v->zero) return v;
    if (op == "dn") return tr.der_normal(v, cx.bound);
    if (op == "dr") return tr.der_normal(v, OUTER_BC);
    if (op == "multsint") return tr.mult_sin_theta(v);
    if (op == "divsint") return tr.div_sin_theta(v);
    if (op == "divr") return tr.div_r(v);
    if (op == "multr") return tr.mult_r(v);
    throw std::runtime_error("unknown operator " + op);
  };
  for (auto& x : r.c) {
    if (x.glob) throw std::runtime_error("operator " + op + " applied to an integral");
    if (x.num && (op == "dn" || op == "dr")) {
      x.x = 0.0;  // derivative of a number
      continue;
    }
    const TE t = as_field(x, cx);
    x.num = false;
    x.te = TE{lin(t.val), lin(t.der)};
  }
  return r;
}

// ---------------------------------------------------------------------------
// the expression evaluator
// ---------------------------------------------------------------------------
ParsedSystem::Tv ParsedSystem::eval(const std::string& s0, Ctx& cx) {
  std::string s = trim(s0);
  if (s.empty()) throw std::runtime_error("empty expression");
  // enclosing parentheses
  while (s.front() == '(' && s.back() == ')' && balanced(s.substr(1, s.size() - 2))) s = trim(s.substr(1, s.size() - 2));
  std::string l, r, arg;
  if (split_last(s, '+', l, r)) return tv_add(eval(l, cx), eval(r, cx), false, cx);
  if (split_last(s, '-', l, r)) return tv_add(eval(l, cx), eval(r, cx), true, cx);
  if (s[0] == '-') return tv_neg(eval(s.substr(1), cx), cx);
  if (split_last(s, '*', l, r)) return tv_mul(eval(l, cx), eval(r, cx), cx);
  if (split_last(s, '/', l, r)) return tv_div(eval(l, cx), eval(r, cx), cx);
  if (unary(s, "Lap", arg) || unary(s, "lap", arg)) {
    // flat Laplacian D_i D^i of every component (Ope_lap, Domain::laplacian)
    Tv a = eval(arg, cx);
    Tv g = tv_grad(tv_grad(a, '\x01', cx), '\x01', cx);
    return g;
  }
  for (const char* op : {"dn", "dr", "multr", "divr", "multsint", "divsint"})
    if (unary(s, op, arg)) return tv_map(eval(arg, cx), op, cx);
  if (unary(s, "integ", arg)) {
    // surface integral at the equation's boundary (Ope_int -> Domain::integ,
    // domain_shell.cpp:862-897, domain_compact.cpp:89-121): r r sin(theta) so, then the
    // theta weights of the boundary values
    const bool saved = cx.glob_ctx;
    cx.glob_ctx = false;
    Tv a = eval(arg, cx);
    cx.glob_ctx = saved;
    if (!a.idx.empty()) throw std::runtime_error("integ of a tensor");
    if (a.c[0].glob) throw std::runtime_error("integ of an integral");
    const TE t = as_field(a.c[0], cx);
    auto rr = [&](const VDp& v) { return tr.mult_r(tr.mult_r(tr.mult_sin_theta(v))); };
    if (doms[cx.d].adapted) {
      // adapted shell: so times sqrt(det q) = r^2 sin(theta) (round boundary), formed in
      // configuration space and given the base of sin(theta) (one.mult_sin_theta()),
      // then the same theta weights (domain_shell_inner_adapted.cpp:1134-1208)
      const Domain& dm = doms[cx.d];
      Base wb = dm.std_base();
      for (int k = 0; k < dm.npc; ++k) (void)op_matrix(mult_sin_1d, wb.bt[k], dm.nt, &wb.bt[k]);
      std::vector<double> w(dm.conf_size());
      for (int k = 0; k < dm.np; ++k)
        for (int j = 0; j < dm.nt; ++j)
          for (int i = 0; i < dm.nr; ++i) {
            const size_t q = (size_t)i + dm.nr * (j + (size_t)dm.nt * k);
            w[q] = dm.R[q] * dm.R[q] * std::sin(dm.th[j]);
          }
      VDp wn = const_field("integ-weight:" + std::to_string(cx.d), cx.d, w, wb);
      Tv r;
      r.c.resize(1);
      r.c[0].glob = true;
      r.c[0].gl.push_back(GlTerm{tr.mul(wn, t.val), tr.mul(wn, t.der), cx.d, cx.bound, 1.0});
      return r;
    }
    Tv r;
    r.c.resize(1);
    r.c[0].glob = true;
    r.c[0].gl.push_back(GlTerm{rr(t.val), rr(t.der), cx.d, cx.bound, 1.0});
    return r;
  }
  for (const char* fn : {"sqrt", "exp", "log", "cos", "sin", "cosh", "sinh", "atanh", "integvolume"})
    if (unary(s, fn, arg)) throw std::runtime_error(std::string("operator ") + fn + " is not supported on the device");
  // D_i X, D^i X, partial_i X
  for (const char* pre : {"D_", "D^", "partial_"}) {
    const size_t L = std::strlen(pre);
    if (s.size() > L + 2 && s.compare(0, L, pre) == 0 && std::isalpha((unsigned char)s[L]) &&
        std::isspace((unsigned char)s[L + 1]))
      return tv_grad(eval(s.substr(L + 2), cx), s[L], cx);
  }
  // X^n (is_ope_pow): '^' followed by a number
  {
    size_t p = s.rfind('^');
    double e;
    if (p != std::string::npos && p > 0 && is_number(trim(s.substr(p + 1)), e)) {
      const int n = (int)std::lround(e);
      if ((double)n != e) throw std::runtime_error("non-integer power in '" + s + "'");
      return tv_pow(eval(s.substr(0, p), cx), n, cx);
    }
  }
  return eval_name(s, cx);
}

ParsedSystem::Tv ParsedSystem::eval_name(const std::string& s, Ctx& cx) {
  Tv r;
  double x;
  if (is_number(s, x)) {
    r.c.resize(1);
    r.c[0].num = true;
    r.c[0].x = x;
    return r;
  }
  std::string base, idx;
  split_indices(s, base, idx);
  for (char c : base)
    if (!(std::isalnum((unsigned char)c) || c == '_')) throw std::runtime_error("cannot parse '" + s + "'");
  const int d = cx.d;
  // unknowns
  for (size_t vi = 0; vi < vars.size(); ++vi) {
    const Var& v = vars[vi];
    if (v.name != base) continue;
    if ((int)idx.size() != v.valence) throw std::runtime_error("wrong number of indices on '" + s + "'");
    r.idx = idx;
    for (int c = 0; c < ipow3(v.valence); ++c) {
      const int f = v.f0 + c;
      Sc sc;
      sc.te.val = field_nodes[d][f];
      sc.te.der = (cx.seed_f == f && cx.seed && cx.seed->dom == d) ? cx.seed : tr.zero(d, SP_TAN);
      r.c.push_back(sc);
    }
    return contract(r, cx);
  }
  // the flat metric
  if (!metric.empty() && base == metric) {
    if (idx.size() != 2) throw std::runtime_error("the metric takes two indices");
    r.idx = idx;
    r.c.resize(9);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        r.c[3 * i + j].num = true;
        r.c[3 * i + j].x = (i == j) ? 1.0 : 0.0;
      }
    return contract(r, cx);
  }
  // constants
  for (const Cst& c : csts) {
    if (c.name != base) continue;
    if ((int)idx.size() != c.valence) throw std::runtime_error("wrong number of indices on '" + s + "'");
    r.idx = idx;
    if (!c.field && !c.param) {
      r.c.resize(1);
      r.c[0].num = true;
      r.c[0].x = c.value;
      return r;
    }
    const Domain& dm = doms[d];
    if (c.param) {
      // a constant field of the parameter's value, rewritable by set_param; its tangent
      // is the seed (constant 1) in the parameter's own program
      VDp v = const_field(c.name + ":0:" + std::to_string(d), d, std::vector<double>(dm.conf_size(), c.value),
                          dm.std_base(), true);
      Sc sc;
      sc.te.val = v;
      if (cx.seed_par == c.par) {
        if (!cx.par_seed) cx.par_seed = tr.seed_node(d, dm.std_base(), cx.par_tab);
        sc.te.der = cx.par_seed;
      } else {
        sc.te.der = tr.zero(d, SP_TAN);
      }
      if (cx.glob_ctx) {
        // in an integral expression a parameter is a scalar: its value read from the field
        // at run time (so set_param reaches it) and its tangent from the seed
        Sc g;
        g.glob = true;
        g.gl.push_back(GlTerm{sc.te.val, sc.te.der, d, cx.bound, 1.0, 1});
        r.c.push_back(g);
        return r;
      }
      r.c.push_back(sc);
      return r;
    }
    size_t off = 0;
    for (int dd = dmin; dd < d; ++dd) off += doms[dd].conf_size();
    size_t per = 0;
    for (int dd = dmin; dd <= dmax; ++dd) per += doms[dd].conf_size();
    for (int q = 0; q < ipow3(c.valence); ++q) {
      std::vector<double> conf(c.conf.begin() + q * per + off, c.conf.begin() + q * per + off + dm.conf_size());
      const Base b = (c.valence == 1 && q == 2) ? dm.anti_base() : dm.std_base();
      Sc sc;
      sc.te.val = const_field(c.name + ":" + std::to_string(q) + ":" + std::to_string(d), d, conf, b);
      sc.te.der = tr.zero(d, SP_TAN);
      r.c.push_back(sc);
    }
    return contract(r, cx);
  }
  // definitions (evaluated once per context)
  for (const Def& df : defs) {
    if (df.name != base) continue;
    if (idx.size() != df.idx.size()) throw std::runtime_error("wrong number of indices on '" + s + "'");
    auto it = cx.defs.find(df.name);
    if (it == cx.defs.end()) it = cx.defs.emplace(df.name, eval(df.rhs, cx)).first;
    const Tv& v = it->second;
    std::string sv = v.idx, sd = df.idx;
    std::sort(sv.begin(), sv.end());
    std::sort(sd.begin(), sd.end());
    if (sv != sd) throw std::runtime_error("definition '" + df.name + "' has free indices '" + v.idx + "'");
    // rename: the definition's index letters -> the use's letters
    r.idx.clear();
    for (char c : v.idx) r.idx.push_back(idx[df.idx.find(c)]);
    r.c = v.c;
    return contract(r, cx);
  }
  throw std::runtime_error("unknown name '" + base + "' in '" + s + "'");
}

// residual of "L = R" as L - R (Equation, add_eq.cpp)
ParsedSystem::Tv ParsedSystem::residual(const std::string& eq, Ctx& cx) {
  std::string l, r;
  if (!split_last(eq, '=', l, r)) throw std::runtime_error("no '=' in equation '" + eq + "'");
  return tv_add(eval(l, cx), eval(r, cx), true, cx);
}

// The expression of equation e evaluated in its domain (part 0) or, for matching,
// in the outer neighbour at its inner boundary (part 1); components in storage order.
std::vector<ParsedSystem::Sc> ParsedSystem::eq_parts(const EqRec& e, int part, Ctx& cx) {
  cx.bound = (e.kind == 1) ? (part == 0 ? OUTER_BC : INNER_BC)
                            : (e.kind == 4 ? e.bound : (e.kind == 3 ? INNER_BC : OUTER_BC));
  cx.glob_ctx = (e.kind == 4);
  Tv t = (e.kind == 1) ? eval(e.which, cx) : residual(e.which, cx);
  cx.glob_ctx = false;
  return t.c;
}

void ParsedSystem::trace_value() {
  Program dummy;
  tr.tan = &dummy;
  field_nodes.assign(doms.size(), std::vector<VDp>(nfields));
  field_slot.assign(nfields, std::vector<int>(doms.size(), -1));
  for (int d = dmin; d <= dmax; ++d)
    for (int f = 0; f < nfields; ++f) field_nodes[d][f] = tr.field_coef(d, field_base(f, d), &field_slot[f][d]);
  std::vector<Ctx> cxs(doms.size());
  for (int d = dmin; d <= dmax; ++d) cxs[d].d = d;
  int row = 0;
  for (auto& e : eqs) {
    e.row0 = row;
    e.nrows.clear();
    std::vector<Sc> p0 = eq_parts(e, 0, cxs[e.d]), p1;
    if (e.kind == 1) p1 = eq_parts(e, 1, cxs[e.d + 1]);
    e.ncomp = (int)p0.size();
    if (e.kind == 4) {
      if (e.ncomp != 1 || !p0[0].glob) throw std::runtime_error("integral equation '" + e.which + "' is not an integral");
      export_glob_terms(p0[0], row, false, 0);
      e.nrows.push_back(1);
      row += 1;
      continue;
    }
    for (int c = 0; c < e.ncomp; ++c) {
      if (p0[c].num) throw std::runtime_error("equation '" + e.which + "' has a constant component");
      if (p0[c].glob) throw std::runtime_error("equation '" + e.which + "' mixes an integral with fields");
      int nr = 0;
      if (e.kind == 0) {
        if (tr.export_inside(p0[c].te.val, e.order, row, 1.0, 0, &nr) < 0)
          throw std::runtime_error("equation '" + e.which + "' has an identically zero component");
      } else {
        const int bound = (e.kind == 3) ? INNER_BC : OUTER_BC;
        if (tr.export_boundary(p0[c].te.val, bound, row, 1.0, 0, &nr) < 0)
          throw std::runtime_error("equation '" + e.which + "' has an identically zero component");
        if (e.kind == 1) {
          int nr2 = 0;
          if (p1[c].num || tr.export_boundary(p1[c].te.val, INNER_BC, row, -1.0, 1, &nr2) < 0 || nr2 != nr)
            throw std::runtime_error("matching row count mismatch in '" + e.which + "'");
        }
      }
      e.nrows.push_back(nr);
      row += nr;
    }
  }
  n_rows = row;
  tr.tan = nullptr;
}

void ParsedSystem::trace_tangent(int f, int d) {
  tprog.emplace_back();
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
  Ctx cx;
  cx.d = d;
  cx.seed_f = f;
  cx.seed = tr.seed_node(d, field_base(f, d), tab);
  for (auto& e : eqs) {
    int part = -1;
    if (e.kind != 1 && e.d == d) part = 0;
    if (e.kind == 1 && e.d == d) part = 0;
    if (e.kind == 1 && e.d + 1 == d) part = 1;
    if (part < 0) continue;
    std::vector<Sc> p = eq_parts(e, part, cx);
    if (e.kind == 4) {
      export_glob_terms(p[0], e.row0, true, 0);
      continue;
    }
    int row = e.row0;
    for (int c = 0; c < e.ncomp; ++c) {
      int nr = 0;
      if (!p[c].num) {
        if (e.kind == 0) {
          tr.export_inside(p[c].te.der, e.order, row, 1.0, 0, &nr);
        } else {
          const int bound = (e.kind == 3 || part == 1) ? INNER_BC : OUTER_BC;
          tr.export_boundary(p[c].te.der, bound, row, part == 1 ? -1.0 : 1.0, 0, &nr);
        }
      }
      if (nr != 0 && nr != e.nrows[c]) throw std::runtime_error("tangent row count differs from the residual's");
      row += e.nrows[c];
    }
  }
  tr.tan = nullptr;
}

// dF/dp: every equation of every domain with the parameter's tangent seeded to the
// constant 1 (a one-coefficient seed table: T_0 of the std base).
void ParsedSystem::trace_param(int p) {
  tr.begin_tangent(&par_prog[p]);
  for (int d = dmin; d <= dmax; ++d) {
    Ctx cx;
    cx.d = d;
    cx.seed_par = p;
    cx.par_tab = par_tab[p];
    for (auto& e : eqs) {
      int part = -1;
      if (e.kind != 1 && e.d == d) part = 0;
      if (e.kind == 1 && e.d == d) part = 0;
      if (e.kind == 1 && e.d + 1 == d) part = 1;
      if (part < 0) continue;
      std::vector<Sc> v = eq_parts(e, part, cx);
      if (e.kind == 4) {
        export_glob_terms(v[0], e.row0, true, 0);
        continue;
      }
      int row = e.row0;
      for (int c = 0; c < e.ncomp; ++c) {
        int nr = 0;
        if (!v[c].num) {
          // a matching equation's two parts land on the same rows: accumulate the second
          const int acc = (e.kind == 1 && part == 1) ? 1 : 0;
          if (e.kind == 0) {
            tr.export_inside(v[c].te.der, e.order, row, 1.0, acc, &nr);
          } else {
            const int bound = (e.kind == 3 || part == 1) ? INNER_BC : OUTER_BC;
            tr.export_boundary(v[c].te.der, bound, row, part == 1 ? -1.0 : 1.0, acc, &nr);
          }
        }
        row += e.nrows[c];
      }
    }
  }
  tr.tan = nullptr;
}

// Rows of a global scalar: value programs write the row (the first term sets it, with
// the constant), tangent programs add every term's tangent (their output is cleared).
void ParsedSystem::export_glob_terms(const Sc& g, int row, bool der, int rowstride_q) {
  (void)rowstride_q;
  if (!g.glob) throw std::runtime_error("not an integral expression");
  if (!der) {
    if (g.gl.empty()) tr.export_global(tr.zero(dmin, SP_VAL), dmin, OUTER_BC, row, 0.0, g.x, 0);
    for (size_t k = 0; k < g.gl.size(); ++k)
      tr.export_global(g.gl[k].val, g.gl[k].dom, g.gl[k].bound, row, g.gl[k].fac, k == 0 ? g.x : 0.0, k > 0,
                       g.gl[k].point);
    return;
  }
  for (const GlTerm& t : g.gl) tr.export_global(t.der, t.dom, t.bound, row, t.fac, 0.0, 1, t.point);
}

// Global quantities: the value program (Q[q]), then per (field, domain) the tangent
// program of the quantities evaluated there (dQ_q/du at q * n + column), then per
// parameter the quantities' own dependence on it at fixed fields.
void ParsedSystem::trace_globals() {
  if (globs.empty()) return;
  Program saved = std::move(tr.val);
  tr.val = Program();
  Program dummy;
  tr.tan = &dummy;
  for (size_t q = 0; q < globs.size(); ++q) {
    Ctx cx;
    cx.d = globs[q].dom;
    cx.bound = globs[q].bound;
    cx.glob_ctx = true;
    Tv t = eval(globs[q].expr, cx);
    if (!t.idx.empty() || !t.c[0].glob) throw std::runtime_error("global '" + globs[q].name + "' is not an integral");
    export_glob_terms(t.c[0], (int)q, false, 0);
  }
  glob_prog = std::move(tr.val);
  tr.val = std::move(saved);
  tr.tan = nullptr;
  glob_tprog.assign(layout.size(), Program());
  for (size_t L = 0; L < layout.size(); ++L) {
    const int f = layout[L][0], d = layout[L][1];
    tr.begin_tangent(&glob_tprog[L]);
    Ctx cx;
    cx.d = d;
    cx.seed_f = f;
    cx.seed = tr.seed_node(d, field_base(f, d), seed_tab_of.at({f, d}));
    for (size_t q = 0; q < globs.size(); ++q) {
      if (globs[q].dom != d) continue;
      cx.bound = globs[q].bound;
      cx.glob_ctx = true;
      Tv t = eval(globs[q].expr, cx);
      // dQ_q / du lands at q * n + column: row q * n_unknowns with a unit column stride
      export_glob_terms(t.c[0], (int)q * n_unknowns, true, 0);
    }
    tr.tan = nullptr;
  }
  glob_par_prog.assign(par_names.size(), Program());
  for (size_t p = 0; p < par_names.size(); ++p) {
    tr.begin_tangent(&glob_par_prog[p]);
    for (size_t q = 0; q < globs.size(); ++q) {
      Ctx cx;
      cx.d = globs[q].dom;
      cx.bound = globs[q].bound;
      cx.seed_par = (int)p;
      cx.par_tab = par_tab[p];
      cx.glob_ctx = true;
      Tv t = eval(globs[q].expr, cx);
      export_glob_terms(t.c[0], (int)q, true, 0);
    }
    tr.tan = nullptr;
  }
}

// Domain-major rows: equation blocks keep their components together and are ordered
// by the domain they live in (matching d|d+1 in d+1), the reference order inside a
// domain.  The value program's export rows are shifted accordingly.
void ParsedSystem::reorder_rows() {
  struct Blk {
    int old0, len, new0;
  };
  std::vector<Blk> blks;
  int row = 0;
  for (int dom = dmin; dom <= dmax; ++dom)
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

void ParsedSystem::build() {
  if (nfields == 0) throw std::runtime_error("the system has no unknown");
  if (eqs.empty()) throw std::runtime_error("the system has no equation");
  trace_value();
  if (order_mode == 1) reorder_rows();
  // unknown layout (system_of_eqs.cpp:597-612): variables, then domains, then
  // components -- or, domain-major, domains first
  layout.clear();
  tau_maps.clear();
  std::vector<std::pair<int, int>> order_fd;
  if (order_mode == 1) {
    for (int d = dmin; d <= dmax; ++d)
      for (auto& vf : var_fields)
        for (int f : vf) order_fd.push_back({f, d});
  } else {
    for (auto& vf : var_fields)
      for (int d = dmin; d <= dmax; ++d)
        for (int f : vf) order_fd.push_back({f, d});
  }
  int off = ndouble;  // the double unknowns come first (System_of_eqs::xx_to_vars)
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
  // A system need not be square to be evaluated (sec_member, check_equations, the
  // Jacobian's columns); the solvers check it (do_newton.cpp:113-118).
  n_unknowns = off;
  for (auto& L : layout) trace_tangent(L[0], L[1]);
  // parameter programs
  par_prog.assign(par_names.size(), Program());
  par_tab.clear();
  for (size_t p = 0; p < par_names.size(); ++p) {
    SeedTab st;
    st.pos = {0, 0, 0};
    st.val = {1.0, 0.0, 0.0};
    st.cnt = {1};
    par_tab.push_back((int)seeds.size());
    seeds.push_back(st);
  }
  for (size_t p = 0; p < par_names.size(); ++p) trace_param((int)p);
  trace_globals();
}

}  // namespace kadhost
