// Device Term_eq: a tracer that turns the Kadath operator algebra (Val_domain,
// Term_eq) into a tape of batched GPU instructions.
//
// Reference semantics restated (each op cites its source in tape.cpp):
//   Val_domain lazy conf / coef storage and base bookkeeping
//     (src/Val_domain/val_domain.cpp:617-653, val_domain_math.cpp),
//   coefficient-space 1-D operators through Base_spectral::ope_1d
//     (include/implementation/ope_1d.cpp:25-82),
//   Cartesian derivatives do_der_abs_from_der_var (domain_shell.cpp:665-683,
//     domain_compact.cpp:674-693),
//   Term_eq forward-mode rules (src/Term_eq/term_eq_math.cpp).
//
// Values live in a value arena (batch 1); tangents in a tangent arena (batch B,
// one tangent per Jacobian column).  A Term_eq whose tangent is identically zero
// for a given seed type carries a zero node: no instruction is emitted for it,
// exactly as Kadath skips zero Val_domains.
#pragma once
#include <memory>
#include <vector>
#include <map>
#include <string>
#include "domain.hpp"

namespace kadhost {

enum Space { SP_VAL = 0, SP_TAN = 1 };

struct Node {
  int id = 0;
  int dom = 0;
  bool zero = false;
  int space = SP_VAL;
  int c = -1, cf = -1;  // slots holding the conf / coef representation (-1: not computed)
  Base base;
};
typedef std::shared_ptr<Node> VDp;

// Instruction kinds.
enum InstrKind {
  I_LIN = 0, I_EW = 1, I_EXPORT_IN = 2, I_EXPORT_BD = 3, I_SEED = 4, I_ZERO = 5, I_EXPORT_INT = 6,
  I_EXPORT_GL = 7  // one row: fac * surface integral of a (r r sin(theta) so) at a boundary (+ c0)
};
enum EwKind { EW_ADD = 0, EW_SUB = 1, EW_MUL = 2, EW_DIV = 3, EW_SCAL = 4, EW_ADDS = 5, EW_COPY = 6, EW_SDIV = 7 };

struct Opnd {
  int slot = -1;
  int space = SP_VAL;
};

struct Instr {
  int kind = 0;
  int dom = 0;
  Opnd dst, a, b;
  // LIN
  int plan = -1;
  int din[3] = {0, 0, 0};
  double scale = 1.0;
  // EW
  int ew = 0;
  int size = 0;  // elements per array
  double s = 0.0;
  // EXPORT
  int rows = -1;   // row table id
  int nrows = 0;
  int row0 = 0;
  double sign = 1.0;
  int accumulate = 0;
  int bsign = 1;   // boundary: +1 (x = 1) or -1 (x = -1)
  // SEED
  int seed_tab = -1;
  // EXPORT_INT: theta basis of phi coefficient 0 (row0 = quantity index, s = factor)
  int aux = 0;
  // LIN or EW: the number of following element-wise instructions this one absorbs (each
  // the only reader of the previous result; set on the device side, which then applies
  // them as the result is stored)
  int fuse = 0;
  // EXPORT_GL: constant added to the row (value programs); a.slot < 0: the constant alone;
  // pt = 1: the point value a[0] of a constant field instead of an integral
  double c0 = 0.0;
  int pt = 0;
};

// A chain of 1-D passes with host matrices (uploaded once per plan).
struct HostPass {
  int axis, nout;
  std::vector<int> line_mat;  // offset of the line's matrix in the pool
};
struct HostPlan {
  std::vector<HostPass> passes;
  std::vector<double> pool;
  int maxsize = 0;
};

struct Program {
  std::vector<Instr> code;
  int nslots = 0;  // virtual slots (tangent programs: before reuse)
};

class Tracer {
 public:
  std::vector<Domain>* doms = nullptr;
  Program val;                 // value program
  Program* tan = nullptr;      // tangent program being traced
  int nval_slots = 0;
  std::vector<HostPlan> plans;
  std::map<std::string, int> plan_ids;
  std::vector<std::vector<TauRow>> row_tabs;
  std::vector<std::pair<int, std::vector<double>>> const_slots;  // constant value slots uploaded at setup
  std::map<std::string, VDp> memo;       // common-subexpression table (value ops persist, tangent ops per trace)
  std::map<int, int> radius_slot;         // domain -> value slot of the radius
  int next_id = 1;

  // start tracing a new tangent program: tangent memo entries are dropped
  void begin_tangent(Program* p);
  VDp seed_node(int dom, const Base& base, int seed_tab);

  explicit Tracer(std::vector<Domain>* d) : doms(d) {}
  Domain& D(int d) { return (*doms)[d]; }

  int new_slot(int space);
  Program& prog(int space) { return space == SP_VAL ? val : *tan; }

  // ---- node constructors
  VDp zero(int dom, int space = SP_TAN);
  VDp const_conf(int dom, const std::vector<double>& values, const Base& base);  // value space
  VDp radius(int dom);                                                            // base undefined
  VDp field_coef(int dom, const Base& base, int* slot_out);                      // value space input

  // ---- lazy transforms
  void coef(const VDp& v);
  void coef_i(const VDp& v);

  // ---- Val_domain arithmetic
  VDp add(const VDp& a, const VDp& b);
  VDp sub(const VDp& a, const VDp& b);
  VDp neg(const VDp& a);
  VDp mul(const VDp& a, const VDp& b);
  VDp div(const VDp& a, const VDp& b);
  VDp scal(const VDp& a, double x);
  VDp div_scal(const VDp& a, double x) { return scal_div(a, x); }
  VDp add_scal(const VDp& a, double x);

  // ---- domain operators
  VDp der_var(const VDp& v, int var);
  VDp mult_cos_phi(const VDp& v);
  VDp mult_sin_phi(const VDp& v);
  VDp mult_cos_theta(const VDp& v);
  VDp mult_sin_theta(const VDp& v);
  VDp div_sin_theta(const VDp& v);
  VDp mult_xm1(const VDp& v);
  VDp div_xm1(const VDp& v);
  // multiplication / division by r (Domain_shell: mult_x alpha + beta in coefficient
  // space and a conf-space division keeping the base, domain_shell_ope.cpp:155-164,
  // 215-219; Domain_compact: div_xm1 / alpha and mult_xm1 alpha,
  // domain_compact_ope.cpp:173-185)
  VDp mult_r(const VDp& v);
  VDp div_r(const VDp& v);
  void der_abs(const VDp& v, VDp out[3]);
  VDp der_normal(const VDp& v, int bound);

  // ---- exports
  // returns the number of rows; emits nothing for a zero node
  int export_inside(const VDp& v, int order, int row0, double sign, int accumulate, int* nrows_out);
  int export_boundary(const VDp& v, int bound, int row0, double sign, int accumulate, int* nrows_out);
  // surface integral at the outer boundary of the coefficients of v (Domain_compact::integ
  // weights, times fac) into quantity q of the column
  void export_integral(const VDp& v, int q, double fac);
  // integral-equation row: row (+)= fac * integ(v) at the boundary (v already holds
  // r r sin(theta) so, as Domain_shell::integ forms it) + c0; a zero v in the value
  // program still writes c0
  void export_global(const VDp& v, int dom, int bound, int row, double fac, double c0, int accumulate,
                     int point = 0);
  int nrows_inside(const Base& b, int dom, int order) { return (int)D(dom).tau_rows_inside(b.bt, order).size(); }
  int nrows_boundary(const Base& b, int dom) { return (int)D(dom).tau_rows_boundary(b.bt).size(); }

 private:
  VDp scal_div(const VDp& a, double x);
  VDp add_impl(const VDp& a, const VDp& b);
  VDp sub_impl(const VDp& a, const VDp& b);
  VDp mul_impl(const VDp& a, const VDp& b);
  VDp div_impl(const VDp& a, const VDp& b);
  VDp scal_impl(const VDp& a, double x);
  VDp scal_div_impl(const VDp& a, double x);
  VDp add_scal_impl(const VDp& a, double x);
  void der_abs_impl(const VDp& v, VDp out[3]);
  VDp lin_theta(const VDp& v, Op1d op, const char* name);
  VDp lin_r(const VDp& v, Op1d op, const char* name);
  VDp phi_flip(const VDp& v, Op1d op, const char* name);
  int plan_for(const std::string& key, const HostPlan& p);
  VDp new_node(int dom, int space);
  VDp clone_with_base(const VDp& v, const Base& b);
  std::string mkey(const char* op, const VDp& a, const VDp& b, double s) const;
  bool memo_get(const std::string& k, VDp& out) const;
  void memo_put(const std::string& k, const VDp& r);
  void emit_ew(int kind, const VDp& dst, int dslot, const VDp& a, int aslot, const VDp& b, int bslot, double s,
               int size);
  int space_of(const VDp& a, const VDp& b) const;
  int export_rows(const VDp& v, const std::vector<TauRow>& rows, int kind, int bsign, int row0, double sign,
                  int accumulate);
};

// Term_eq over one domain: value node and tangent node.
struct TE {
  VDp val, der;
};

}  // namespace kadhost
