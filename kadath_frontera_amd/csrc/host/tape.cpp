// Device Term_eq tracer.  See tape.hpp.
#include "tape.hpp"
#include <sstream>
#include <stdexcept>
#include <cmath>

namespace kadhost {

int Tracer::new_slot(int space) {
  if (space == SP_VAL) return nval_slots++;
  return tan->nslots++;
}

VDp Tracer::new_node(int dom, int space) {
  auto n = std::make_shared<Node>();
  n->id = next_id++;
  n->dom = dom;
  n->space = space;
  return n;
}

VDp Tracer::clone_with_base(const VDp& v, const Base& b) {
  auto n = new_node(v->dom, v->space);
  n->zero = v->zero;
  n->c = v->c;
  n->cf = v->cf;
  n->base = b;
  return n;
}

std::string Tracer::mkey(const char* op, const VDp& a, const VDp& b, double s) const {
  std::ostringstream o;
  o.precision(17);
  o << op << ":" << (a ? a->id : 0) << ":" << (b ? b->id : 0) << ":" << s;
  return o.str();
}

bool Tracer::memo_get(const std::string& k, VDp& out) const {
  auto it = memo.find(k);
  if (it == memo.end()) return false;
  out = it->second;
  return true;
}

void Tracer::memo_put(const std::string& k, const VDp& r) { memo[k] = r; }

void Tracer::begin_tangent(Program* p) {
  tan = p;
  for (auto it = memo.begin(); it != memo.end();) {
    if (it->second->space == SP_TAN) it = memo.erase(it);
    else ++it;
  }
}

VDp Tracer::seed_node(int dom, const Base& base, int seed_tab) {
  auto n = new_node(dom, SP_TAN);
  n->cf = new_slot(SP_TAN);
  n->base = base;
  Instr in;
  in.kind = I_SEED;
  in.dom = dom;
  in.dst = {n->cf, SP_TAN};
  in.seed_tab = seed_tab;
  in.size = D(dom).coef_size();
  tan->code.push_back(in);
  return n;
}

VDp Tracer::zero(int dom, int space) {
  auto n = new_node(dom, space);
  n->zero = true;
  return n;
}

VDp Tracer::const_conf(int dom, const std::vector<double>& values, const Base& base) {
  auto n = new_node(dom, SP_VAL);
  n->c = new_slot(SP_VAL);
  n->base = base;
  const_slots.emplace_back(n->c, values);
  return n;
}

VDp Tracer::radius(int dom) {
  // Domain::get_radius(): conf values, base not defined (domain_shell.cpp:113-124)
  auto it = radius_slot.find(dom);
  if (it != radius_slot.end()) {
    auto n = new_node(dom, SP_VAL);
    n->c = it->second;
    return n;
  }
  auto n = const_conf(dom, D(dom).R, Base());
  radius_slot[dom] = n->c;
  return n;
}

VDp Tracer::field_coef(int dom, const Base& base, int* slot_out) {
  auto n = new_node(dom, SP_VAL);
  n->cf = new_slot(SP_VAL);
  n->base = base;
  *slot_out = n->cf;
  return n;
}

int Tracer::plan_for(const std::string& key, const HostPlan& p) {
  auto it = plan_ids.find(key);
  if (it != plan_ids.end()) return it->second;
  const int id = (int)plans.size();
  plans.push_back(p);
  plan_ids[key] = id;
  return id;
}

static std::string base_key(const Base& b) {
  std::ostringstream o;
  o << b.bp << ":";
  for (int x : b.bt) o << x << ",";
  o << ":";
  for (int x : b.br) o << x << ",";
  return o.str();
}

// Forward (dir 0) / inverse (dir 1) 3-D transform plan for a base: phi, theta, r
// (coef.cpp:96-104) or r, theta, phi (coef_i.cpp:96-104).
static HostPlan transform_plan(const Domain& d, const Base& b, int dir) {
  HostPlan p;
  std::map<std::pair<int, int>, int> where;
  auto add = [&](int axis, int base, int n) {
    auto k = std::make_pair(axis, base);
    auto f = where.find(k);
    if (f != where.end()) return f->second;
    const Mat M = dir == 0 ? coef_matrix(base, n) : coef_i_matrix(base, n);
    const int off = (int)p.pool.size();
    p.pool.insert(p.pool.end(), M.a.begin(), M.a.end());
    where[k] = off;
    return off;
  };
  HostPass pp{2, dir == 0 ? d.npc : d.np, {add(2, b.bp, d.np)}};
  HostPass pt{1, d.nt, {}};
  for (int k = 0; k < d.npc; ++k) pt.line_mat.push_back(add(1, b.bt[k], d.nt));
  HostPass pr{0, d.nr, {}};
  for (int q = 0; q < d.npc * d.nt; ++q) pr.line_mat.push_back(add(0, b.br[q], d.nr));
  if (dir == 0) p.passes = {pp, pt, pr};
  else p.passes = {pr, pt, pp};
  p.maxsize = d.coef_size();
  return p;
}

void Tracer::coef(const VDp& v) {
  if (v->zero || v->cf >= 0) return;
  if (!v->base.def) throw std::runtime_error("Base not defined in Val_domain::coef");
  Domain& d = D(v->dom);
  const int pid = plan_for("F" + std::to_string(v->dom) + ":" + base_key(v->base), transform_plan(d, v->base, 0));
  Instr in;
  in.kind = I_LIN;
  in.dom = v->dom;
  in.plan = pid;
  in.a = {v->c, v->space};
  v->cf = new_slot(v->space);
  in.dst = {v->cf, v->space};
  in.din[0] = d.nr; in.din[1] = d.nt; in.din[2] = d.np;
  prog(v->space).code.push_back(in);
}

void Tracer::coef_i(const VDp& v) {
  if (v->zero || v->c >= 0) return;
  if (!v->base.def) throw std::runtime_error("Base not defined in Val_domain::coef_i");
  Domain& d = D(v->dom);
  const int pid = plan_for("I" + std::to_string(v->dom) + ":" + base_key(v->base), transform_plan(d, v->base, 1));
  Instr in;
  in.kind = I_LIN;
  in.dom = v->dom;
  in.plan = pid;
  in.a = {v->cf, v->space};
  v->c = new_slot(v->space);
  in.dst = {v->c, v->space};
  in.din[0] = d.nr; in.din[1] = d.nt; in.din[2] = d.npc;
  prog(v->space).code.push_back(in);
}

int Tracer::space_of(const VDp& a, const VDp& b) const {
  return (a->space == SP_TAN || (b && b->space == SP_TAN)) ? SP_TAN : SP_VAL;
}

void Tracer::emit_ew(int kind, const VDp& dst, int dslot, const VDp& a, int aslot, const VDp& b, int bslot, double s,
                     int size) {
  Instr in;
  in.kind = I_EW;
  in.dom = dst->dom;
  in.ew = kind;
  in.dst = {dslot, dst->space};
  in.a = {aslot, a->space};
  if (b) in.b = {bslot, b->space};
  in.s = s;
  in.size = size;
  prog(dst->space).code.push_back(in);
}

// operator+ (val_domain_math.cpp:150-195)
VDp Tracer::add(const VDp& a, const VDp& b) {
  if (!a->zero && !b->zero) {
    VDp m;
    const std::string k = mkey("add", a, b, 0.0);
    if (memo_get(k, m)) return m;
    VDp r = add_impl(a, b);
    memo_put(k, r);
    return r;
  }
  return add_impl(a, b);
}

VDp Tracer::add_impl(const VDp& a, const VDp& b) {
  if (a->zero) return b;
  if (b->zero) return a;
  if (a->base.def && b->base.def && a->base != b->base)
    throw std::runtime_error("base mismatch in Val_domain operator+");
  auto r = new_node(a->dom, space_of(a, b));
  if (a->base.def && b->base.def) r->base = a->base;
  Domain& d = D(a->dom);
  if (a->c >= 0 || b->c >= 0) {
    coef_i(a);
    coef_i(b);
    r->c = new_slot(r->space);
    emit_ew(EW_ADD, r, r->c, a, a->c, b, b->c, 0.0, d.conf_size());
  } else {
    r->cf = new_slot(r->space);
    emit_ew(EW_ADD, r, r->cf, a, a->cf, b, b->cf, 0.0, d.coef_size());
  }
  return r;
}

// operator- (val_domain_math.cpp:220-262)
VDp Tracer::sub(const VDp& a, const VDp& b) {
  if (!a->zero && !b->zero) {
    VDp m;
    const std::string k = mkey("sub", a, b, 0.0);
    if (memo_get(k, m)) return m;
    VDp r = sub_impl(a, b);
    memo_put(k, r);
    return r;
  }
  return sub_impl(a, b);
}

VDp Tracer::sub_impl(const VDp& a, const VDp& b) {
  if (a->zero) return neg(b);
  if (b->zero) return a;
  if (a->base.def && b->base.def && a->base != b->base)
    throw std::runtime_error("base mismatch in Val_domain operator-");
  auto r = new_node(a->dom, space_of(a, b));
  if (a->base.def && b->base.def) r->base = a->base;
  Domain& d = D(a->dom);
  if (a->c >= 0 || b->c >= 0) {
    coef_i(a);
    coef_i(b);
    r->c = new_slot(r->space);
    emit_ew(EW_SUB, r, r->c, a, a->c, b, b->c, 0.0, d.conf_size());
  } else {
    r->cf = new_slot(r->space);
    emit_ew(EW_SUB, r, r->cf, a, a->cf, b, b->cf, 0.0, d.coef_size());
  }
  return r;
}

// unary - (val_domain_math.cpp:140-147): both representations are negated
VDp Tracer::neg(const VDp& a) {
  if (a->zero) return a;
  return scal(a, -1.0);
}

// operator* (val_domain_math.cpp:302-318)
VDp Tracer::mul(const VDp& a, const VDp& b) {
  if (!a->zero && !b->zero) {
    VDp m;
    const std::string k = mkey("mul", a, b, 0.0);
    if (memo_get(k, m)) return m;
    VDp r = mul_impl(a, b);
    memo_put(k, r);
    return r;
  }
  return mul_impl(a, b);
}

VDp Tracer::mul_impl(const VDp& a, const VDp& b) {
  if (a->zero) return a;
  if (b->zero) return b;
  coef_i(a);
  coef_i(b);
  auto r = new_node(a->dom, space_of(a, b));
  r->base = D(a->dom).mult_base(a->base, b->base);
  r->c = new_slot(r->space);
  emit_ew(EW_MUL, r, r->c, a, a->c, b, b->c, 0.0, D(a->dom).conf_size());
  return r;
}

// operator/ (val_domain_math.cpp:382-399)
VDp Tracer::div(const VDp& a, const VDp& b) {
  if (!a->zero && !b->zero) {
    VDp m;
    const std::string k = mkey("div", a, b, 0.0);
    if (memo_get(k, m)) return m;
    VDp r = div_impl(a, b);
    memo_put(k, r);
    return r;
  }
  return div_impl(a, b);
}

VDp Tracer::div_impl(const VDp& a, const VDp& b) {
  if (a->zero) return a;
  if (b->zero) throw std::runtime_error("Division by zero");
  coef_i(a);
  coef_i(b);
  auto r = new_node(a->dom, space_of(a, b));
  r->base = D(a->dom).mult_base(a->base, b->base);
  r->c = new_slot(r->space);
  emit_ew(EW_DIV, r, r->c, a, a->c, b, b->c, 0.0, D(a->dom).conf_size());
  return r;
}

// operator*(Val_domain, double) (val_domain_math.cpp:320-331): every stored representation scaled
VDp Tracer::scal(const VDp& a, double x) {
  if (a->zero) return a;
  VDp m;
  const std::string k = mkey("scal", a, nullptr, x);
  if (memo_get(k, m)) return m;
  VDp r = scal_impl(a, x);
  memo_put(k, r);
  return r;
}

VDp Tracer::scal_impl(const VDp& a, double x) {
  if (a->zero) return a;
  auto r = new_node(a->dom, a->space);
  r->base = a->base;
  Domain& d = D(a->dom);
  if (a->c >= 0) {
    r->c = new_slot(r->space);
    emit_ew(EW_SCAL, r, r->c, a, a->c, nullptr, -1, x, d.conf_size());
  }
  if (a->cf >= 0) {
    r->cf = new_slot(r->space);
    emit_ew(EW_SCAL, r, r->cf, a, a->cf, nullptr, -1, x, d.coef_size());
  }
  return r;
}

// operator/(Val_domain, double) (val_domain_math.cpp:401-412)
VDp Tracer::scal_div(const VDp& a, double x) {
  if (a->zero) return a;
  VDp m;
  const std::string k = mkey("sdiv", a, nullptr, x);
  if (memo_get(k, m)) return m;
  VDp r = scal_div_impl(a, x);
  memo_put(k, r);
  return r;
}

VDp Tracer::scal_div_impl(const VDp& a, double x) {
  if (a->zero) return a;
  auto r = new_node(a->dom, a->space);
  r->base = a->base;
  Domain& d = D(a->dom);
  if (a->c >= 0) {
    r->c = new_slot(r->space);
    emit_ew(EW_SDIV, r, r->c, a, a->c, nullptr, -1, x, d.conf_size());
  }
  if (a->cf >= 0) {
    r->cf = new_slot(r->space);
    emit_ew(EW_SDIV, r, r->cf, a, a->cf, nullptr, -1, x, d.coef_size());
  }
  return r;
}

// operator+(Val_domain, double) (val_domain_math.cpp:197-214)
VDp Tracer::add_scal(const VDp& a, double x) {
  if (!a->zero) {
    VDp m;
    const std::string k = mkey("adds", a, nullptr, x);
    if (memo_get(k, m)) return m;
    VDp r = add_scal_impl(a, x);
    memo_put(k, r);
    return r;
  }
  return add_scal_impl(a, x);
}

VDp Tracer::add_scal_impl(const VDp& a, double x) {
  if (a->zero) {
    Domain& d = D(a->dom);
    return const_conf(a->dom, std::vector<double>(d.conf_size(), x), d.std_base());
  }
  coef_i(a);
  auto r = new_node(a->dom, a->space);
  r->base = a->base;
  r->c = new_slot(r->space);
  emit_ew(EW_ADDS, r, r->c, a, a->c, nullptr, -1, x, D(a->dom).conf_size());
  return r;
}

// ---------------------------------------------------------------------------
// coefficient-space 1-D operators (Base_spectral::ope_1d, ope_1d.cpp:25-82)
// ---------------------------------------------------------------------------
VDp Tracer::lin_theta(const VDp& v, Op1d op, const char* name) {
  if (v->zero) return v;
  {
    VDp m;
    if (memo_get(mkey(name, v, nullptr, 0.0), m)) return m;
  }
  coef(v);
  Domain& d = D(v->dom);
  HostPlan p;
  HostPass pt{1, d.nt, {}};
  Base nb = v->base;
  std::map<int, int> where;
  for (int k = 0; k < d.npc; ++k) {
    const int b = v->base.bt[k];
    int ob;
    auto f = where.find(b);
    Mat M = op_matrix(op, b, d.nt, &ob);
    if (f == where.end()) {
      where[b] = (int)p.pool.size();
      p.pool.insert(p.pool.end(), M.a.begin(), M.a.end());
    }
    pt.line_mat.push_back(where[b]);
    nb.bt[k] = ob;
  }
  p.passes = {pt};
  p.maxsize = d.coef_size();
  const int pid = plan_for(std::string(name) + std::to_string(v->dom) + ":" + base_key(v->base), p);
  auto r = new_node(v->dom, v->space);
  r->base = nb;
  r->cf = new_slot(r->space);
  Instr in;
  in.kind = I_LIN;
  in.dom = v->dom;
  in.plan = pid;
  in.a = {v->cf, v->space};
  in.dst = {r->cf, r->space};
  in.din[0] = d.nr; in.din[1] = d.nt; in.din[2] = d.npc;
  prog(r->space).code.push_back(in);
  memo_put(mkey(name, v, nullptr, 0.0), r);
  return r;
}

VDp Tracer::lin_r(const VDp& v, Op1d op, const char* name) {
  if (v->zero) return v;
  {
    VDp m;
    if (memo_get(mkey(name, v, nullptr, 0.0), m)) return m;
  }
  coef(v);
  Domain& d = D(v->dom);
  for (int b : v->base.br)
    if (b != CHEB) throw std::runtime_error("r operator implemented for CHEB lines only");
  int ob;
  Mat M = op_matrix(op, CHEB, d.nr, &ob);
  HostPlan p;
  p.pool = M.a;
  p.passes = {HostPass{0, d.nr, std::vector<int>((size_t)d.npc * d.nt, 0)}};
  p.maxsize = d.coef_size();
  const int pid = plan_for(std::string(name) + std::to_string(v->dom), p);
  auto r = new_node(v->dom, v->space);
  r->base = v->base;
  for (auto& b : r->base.br) b = ob;
  r->cf = new_slot(r->space);
  Instr in;
  in.kind = I_LIN;
  in.dom = v->dom;
  in.plan = pid;
  in.a = {v->cf, v->space};
  in.dst = {r->cf, r->space};
  in.din[0] = d.nr; in.din[1] = d.nt; in.din[2] = d.npc;
  prog(r->space).code.push_back(in);
  memo_put(mkey(name, v, nullptr, 0.0), r);
  return r;
}

// mult_cos_phi / mult_sin_phi (domain_shell_ope.cpp:31-107): phi operator plus the
// theta-basis inversion; der_var(3) (val_domain.cpp:771-782) keeps the base.
VDp Tracer::phi_flip(const VDp& v, Op1d op, const char* name) {
  if (v->zero) return v;
  {
    VDp m;
    if (memo_get(mkey(name, v, nullptr, 0.0), m)) return m;
  }
  coef(v);
  Domain& d = D(v->dom);
  int ob;
  Mat M = op_matrix(op, COSSIN, d.npc, &ob);
  HostPlan p;
  p.pool = M.a;
  p.passes = {HostPass{2, d.npc, {0}}};
  p.maxsize = d.coef_size();
  const int pid = plan_for(std::string(name) + std::to_string(v->dom), p);
  auto r = new_node(v->dom, v->space);
  r->base = v->base;
  const bool flip = (op != der_1d);
  if (flip)
    for (auto& b : r->base.bt) {
      switch (b) {
        case COS_EVEN: b = SIN_ODD; break;
        case COS_ODD: b = SIN_EVEN; break;
        case SIN_EVEN: b = COS_ODD; break;
        case SIN_ODD: b = COS_EVEN; break;
        default: throw std::runtime_error("Unknown case in mult_cos_phi");
      }
    }
  r->cf = new_slot(r->space);
  Instr in;
  in.kind = I_LIN;
  in.dom = v->dom;
  in.plan = pid;
  in.a = {v->cf, v->space};
  in.dst = {r->cf, r->space};
  in.din[0] = d.nr; in.din[1] = d.nt; in.din[2] = d.npc;
  prog(r->space).code.push_back(in);
  memo_put(mkey(name, v, nullptr, 0.0), r);
  return r;
}

VDp Tracer::der_var(const VDp& v, int var) {
  if (v->zero) return v;
  if (var == 1) return lin_r(v, der_1d, "der_r");
  if (var == 2) return lin_theta(v, der_1d, "der_t");
  return phi_flip(v, der_1d, "der_p");
}
VDp Tracer::mult_cos_phi(const VDp& v) { return phi_flip(v, mult_cos_1d, "mcp"); }
VDp Tracer::mult_sin_phi(const VDp& v) { return phi_flip(v, mult_sin_1d, "msp"); }
VDp Tracer::mult_cos_theta(const VDp& v) { return lin_theta(v, mult_cos_1d, "mct"); }
VDp Tracer::mult_sin_theta(const VDp& v) { return lin_theta(v, mult_sin_1d, "mst"); }
VDp Tracer::div_sin_theta(const VDp& v) { return lin_theta(v, div_sin_1d, "dst"); }
VDp Tracer::mult_xm1(const VDp& v) { return lin_r(v, mult_xm1_1d, "mxm1"); }
VDp Tracer::div_xm1(const VDp& v) { return lin_r(v, div_xm1_1d, "dxm1"); }

VDp Tracer::mult_r(const VDp& v) {
  if (v->zero) return v;
  Domain& d = D(v->dom);
  if (d.compact) return scal_div(div_xm1(v), d.alpha);
  return add(scal(lin_r(v, mult_x_1d, "mx"), d.alpha), scal(v, d.beta));
}

VDp Tracer::div_r(const VDp& v) {
  if (v->zero) return v;
  Domain& d = D(v->dom);
  if (d.compact) return scal(mult_xm1(v), d.alpha);
  // conf-space division by the radius; the result keeps the operand's base
  VDp r = div(v, radius(v->dom));
  return clone_with_base(r, v->base);
}

// do_der_abs_from_der_var: shell domain_shell.cpp:665-683, compact domain_compact.cpp:674-693
void Tracer::der_abs(const VDp& v, VDp out[3]) {
  if (v->zero) {
    out[0] = out[1] = out[2] = v;
    return;
  }
  if (memo_get(mkey("dabs0", v, nullptr, 0.0), out[0])) {
    memo_get(mkey("dabs1", v, nullptr, 0.0), out[1]);
    memo_get(mkey("dabs2", v, nullptr, 0.0), out[2]);
    return;
  }
  der_abs_impl(v, out);
  memo_put(mkey("dabs0", v, nullptr, 0.0), out[0]);
  memo_put(mkey("dabs1", v, nullptr, 0.0), out[1]);
  memo_put(mkey("dabs2", v, nullptr, 0.0), out[2]);
}

void Tracer::der_abs_impl(const VDp& v, VDp out[3]) {
  Domain& d = D(v->dom);
  VDp d0 = der_var(v, 1), d1 = der_var(v, 2), d2 = der_var(v, 3);
  if (!d.compact) {
    VDp sintdr = scal_div(mult_sin_theta(d0), d.alpha);
    VDp rad = radius(v->dom);
    VDp dtsr = clone_with_base(div(d1, rad), d1->base);
    VDp dpsr = clone_with_base(div(d2, rad), d2->base);
    VDp costdtsr = mult_cos_theta(dtsr);
    VDp dpsrssint = div_sin_theta(dpsr);
    VDp s = add(sintdr, costdtsr);
    out[0] = sub(mult_cos_phi(s), mult_sin_phi(dpsrssint));
    out[1] = add(mult_sin_phi(s), mult_cos_phi(dpsrssint));
    out[2] = sub(scal_div(mult_cos_theta(d0), d.alpha), mult_sin_theta(dtsr));
    return;
  }
  VDp dr = neg(mult_xm1(d0));
  VDp sintdr = mult_sin_theta(dr);
  VDp costdt = mult_cos_theta(d1);
  VDp dpssint = div_sin_theta(d2);
  VDp s = add(sintdr, costdt);
  out[0] = scal(mult_xm1(sub(mult_cos_phi(s), mult_sin_phi(dpssint))), d.alpha);
  out[1] = scal(mult_xm1(add(mult_sin_phi(s), mult_cos_phi(dpssint))), d.alpha);
  out[2] = scal(mult_xm1(sub(mult_cos_theta(dr), mult_sin_theta(d1))), d.alpha);
}

// der_normal: shell domain_shell.cpp:71-87, compact domain_compact.cpp:68-87
VDp Tracer::der_normal(const VDp& v, int bound) {
  (void)bound;
  if (v->zero) return v;
  Domain& d = D(v->dom);
  VDp r = der_var(v, 1);
  if (!d.compact) return scal_div(r, d.alpha);
  r = mult_xm1(mult_xm1(r));
  return scal(r, -d.alpha);
}

// ---------------------------------------------------------------------------
// exports
// ---------------------------------------------------------------------------
int Tracer::export_rows(const VDp& v, const std::vector<TauRow>& rows, int kind, int bsign, int row0, double sign,
                        int accumulate) {
  std::ostringstream key;
  key << "rows" << kind << ":" << v->dom << ":" << bsign << ":" << base_key(v->base);
  int tab;
  auto it = plan_ids.find(key.str());
  if (it != plan_ids.end()) {
    tab = it->second;
  } else {
    tab = (int)row_tabs.size();
    row_tabs.push_back(rows);
    plan_ids[key.str()] = tab;
  }
  Instr in;
  in.kind = kind;
  in.dom = v->dom;
  in.a = {v->cf, v->space};
  in.rows = tab;
  in.nrows = (int)rows.size();
  in.row0 = row0;
  in.sign = sign;
  in.accumulate = accumulate;
  in.bsign = bsign;
  prog(v->space).code.push_back(in);
  return (int)rows.size();
}

int Tracer::export_inside(const VDp& v, int order, int row0, double sign, int accumulate, int* nrows_out) {
  if (v->zero) return -1;
  coef(v);
  auto rows = D(v->dom).tau_rows_inside(v->base.bt, order);
  *nrows_out = (int)rows.size();
  return export_rows(v, rows, I_EXPORT_IN, 1, row0, sign, accumulate);
}

void Tracer::export_integral(const VDp& v, int q, double fac) {
  if (v->zero) return;
  coef(v);
  Instr in;
  in.kind = I_EXPORT_INT;
  in.dom = v->dom;
  in.a = {v->cf, v->space};
  in.row0 = q;
  in.s = fac;
  in.aux = v->base.bt[0];
  prog(v->space).code.push_back(in);
}

void Tracer::export_global(const VDp& v, int dom, int bound, int row, double fac, double c0, int accumulate,
                           int point) {
  Instr in;
  in.kind = I_EXPORT_GL;
  in.pt = point;
  in.dom = dom;
  in.row0 = row;
  in.s = fac;
  in.c0 = c0;
  in.accumulate = accumulate;
  in.bsign = bound == INNER_BC ? -1 : 1;
  if (v->zero) {
    if (v->space != SP_VAL) return;  // a zero tangent adds nothing
    in.a = {-1, SP_VAL};
    val.code.push_back(in);
    return;
  }
  if (point) {
    // a constant field: its value is conf[0] (conf slot) or coefficient 0 (coef slot, the
    // T_0 cos(0) term of the std base)
    in.a = {v->c >= 0 ? v->c : v->cf, v->space};
    prog(v->space).code.push_back(in);
    return;
  }
  coef(v);
  in.a = {v->cf, v->space};
  in.aux = v->base.bt[0];
  prog(v->space).code.push_back(in);
}

int Tracer::export_boundary(const VDp& v, int bound, int row0, double sign, int accumulate, int* nrows_out) {
  if (v->zero) return -1;
  coef(v);
  auto rows = D(v->dom).tau_rows_boundary(v->base.bt);
  *nrows_out = (int)rows.size();
  return export_rows(v, rows, I_EXPORT_BD, bound == INNER_BC ? -1 : 1, row0, sign, accumulate);
}

}  // namespace kadhost
