// C-ABI of the batched spectral transforms (Base_spectral::coef / coef_i,
// src/Coef/coef.cpp:96-104 and src/Coef/coef_i.cpp:96-104) and host-side access
// to the 1-D operator matrices (for CPU-side checks of the tables).
#include "kad_common.hpp"
#include "kad_ctx.hpp"
#include "kad_spectral.hpp"
#include "host/spectral.hpp"
#include "../../include/kadath_amd.h"
#include <map>
#include <sstream>
#include <cstring>
#include <algorithm>
#include <stdexcept>

using kad::Ctx;
using kad::DevPlan;
using namespace kadhost;

namespace {

// Builds (or fetches) the device plan of a 3-D forward (dir=0) or inverse
// (dir=1) transform of one domain array with the given basis tables.
int get_plan(Ctx* c, int dir, int nr, int nt, int np, const int* base_r, const int* base_t, int base_p,
             DevPlan** out) {
  const int npc = np + 2;  // phi coefficients (domain_shell.cpp:256-257)
  std::ostringstream key;
  key << "T" << dir << ":" << nr << "," << nt << "," << np << ":" << base_p << ":";
  for (int k = 0; k < npc; ++k) key << base_t[k] << ",";
  key << ":";
  for (int q = 0; q < nt * npc; ++q) key << base_r[q] << ",";
  auto it = c->plans.find(key.str());
  if (it != c->plans.end()) {
    *out = it->second;
    return KAD_OK;
  }
  std::vector<double> pool;
  std::map<std::pair<int, int>, int> where;  // (axis, base) -> pool offset
  auto add = [&](int axis, int base, int nconf) -> int {
    auto k = std::make_pair(axis, base);
    auto f = where.find(k);
    if (f != where.end()) return f->second;
    const Mat M = dir == 0 ? coef_matrix(base, nconf) : coef_i_matrix(base, nconf);
    const int off = (int)pool.size();
    pool.insert(pool.end(), M.a.begin(), M.a.end());
    where[k] = off;
    return off;
  };
  std::vector<int> lp(1), lt(npc), lr(nt * npc);
  try {
    lp[0] = add(2, base_p, np);
    for (int k = 0; k < npc; ++k) lt[k] = add(1, base_t[k], nt);
    for (int q = 0; q < nt * npc; ++q) lr[q] = add(0, base_r[q], nr);
  } catch (const std::exception& e) {
    fprintf(stderr, "[kadath_amd] transform plan: %s\n", e.what());
    return KAD_ERR_ARG;
  }
  DevPlan* p = new DevPlan();
  std::vector<int> lines;
  lines.insert(lines.end(), lp.begin(), lp.end());
  lines.insert(lines.end(), lt.begin(), lt.end());
  lines.insert(lines.end(), lr.begin(), lr.end());
  int rc = c->upload(pool.data(), pool.size() * sizeof(double), (void**)&p->pool);
  if (!rc) rc = c->upload(lines.data(), lines.size() * sizeof(int), (void**)&p->lines);
  if (rc) {
    delete p;
    return rc;
  }
  const int* dlp = p->lines;
  const int* dlt = p->lines + 1;
  const int* dlr = p->lines + 1 + npc;
  p->P.npass = 3;
  p->P.maxsize = nr * nt * npc;
  p->P.pool_size = (int)pool.size();
  if (dir == 0) {  // phi, theta, r  (coef.cpp:100-101: d = ndim-1 .. 0)
    p->P.axis[0] = 2; p->P.nout[0] = npc; p->P.line_mat[0] = dlp; p->P.nlines[0] = 1;
    p->P.axis[1] = 1; p->P.nout[1] = nt;  p->P.line_mat[1] = dlt; p->P.nlines[1] = npc;
    p->P.axis[2] = 0; p->P.nout[2] = nr;  p->P.line_mat[2] = dlr; p->P.nlines[2] = nt * npc;
    p->d_in[0] = nr; p->d_in[1] = nt; p->d_in[2] = np;
    p->d_out[0] = nr; p->d_out[1] = nt; p->d_out[2] = npc;
  } else {  // r, theta, phi  (coef_i.cpp:100-101: d = 0 .. ndim-1)
    p->P.axis[0] = 0; p->P.nout[0] = nr; p->P.line_mat[0] = dlr; p->P.nlines[0] = nt * npc;
    p->P.axis[1] = 1; p->P.nout[1] = nt; p->P.line_mat[1] = dlt; p->P.nlines[1] = npc;
    p->P.axis[2] = 2; p->P.nout[2] = np; p->P.line_mat[2] = dlp; p->P.nlines[2] = 1;
    p->d_in[0] = nr; p->d_in[1] = nt; p->d_in[2] = npc;
    p->d_out[0] = nr; p->d_out[1] = nt; p->d_out[2] = np;
  }
  c->plans[key.str()] = p;
  *out = p;
  return KAD_OK;
}

int get_line_matrix(Ctx* c, int dir, int base, int nconf, const double** dM, int* nin, int* nout) {
  std::ostringstream key;
  key << "L" << dir << ":" << base << ":" << nconf;
  auto it = c->plans.find(key.str());
  DevPlan* p;
  if (it == c->plans.end()) {
    Mat M;
    try {
      M = dir == 0 ? coef_matrix(base, nconf) : coef_i_matrix(base, nconf);
    } catch (const std::exception& e) {
      fprintf(stderr, "[kadath_amd] line plan: %s\n", e.what());
      return KAD_ERR_ARG;
    }
    p = new DevPlan();
    int rc = c->upload(M.a.data(), M.a.size() * sizeof(double), (void**)&p->pool);
    if (rc) {
      delete p;
      return rc;
    }
    p->d_in[0] = M.nin;
    p->d_out[0] = M.nout;
    c->plans[key.str()] = p;
  } else {
    p = it->second;
  }
  *dM = p->pool;
  *nin = p->d_in[0];
  *nout = p->d_out[0];
  return KAD_OK;
}

bool shell_sizes_ok(int nr, int nt, int np) {
  return nr >= 4 && nt >= 4 && np >= 4 && (np % 2 == 0) && nr <= 64 && nt <= 64 && np <= 62 &&
         (size_t)2 * nr * nt * (np + 2) * sizeof(double) <= 160 * 1024;
}

}  // namespace

extern "C" {

int kad_coef_3d(kad_ctx* ctx, int nr, int nt, int np, const int* base_r, const int* base_t, int base_p, int nbatch,
                const double* conf, double* coef) {
  if (!ctx || !base_r || !base_t || !conf || !coef || nbatch < 0 || !shell_sizes_ok(nr, nt, np)) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  DevPlan* p;
  int rc = get_plan(c, 0, nr, nt, np, base_r, base_t, base_p, &p);
  if (rc) return rc;
  return kad::apply3d_launch(conf, (int64_t)nr * nt * np, coef, (int64_t)nr * nt * (np + 2), nbatch, nr, nt, np, p->P,
                             p->pool, 0, 1.0, c->stream);
}

int kad_coef_i_3d(kad_ctx* ctx, int nr, int nt, int np, const int* base_r, const int* base_t, int base_p, int nbatch,
                  const double* coef, double* conf) {
  if (!ctx || !base_r || !base_t || !conf || !coef || nbatch < 0 || !shell_sizes_ok(nr, nt, np)) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  DevPlan* p;
  int rc = get_plan(c, 1, nr, nt, np, base_r, base_t, base_p, &p);
  if (rc) return rc;
  return kad::apply3d_launch(coef, (int64_t)nr * nt * (np + 2), conf, (int64_t)nr * nt * np, nbatch, nr, nt, np + 2,
                             p->P, p->pool, 0, 1.0, c->stream);
}

int kad_coef_1d(kad_ctx* ctx, int base, int nconf, int64_t nlines, const double* conf, double* coef) {
  if (!ctx || !conf || !coef || nlines < 0 || nconf < 4 || nconf > 64) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const double* M;
  int nin, nout;
  int rc = get_line_matrix(c, 0, base, nconf, &M, &nin, &nout);
  if (rc) return rc;
  return kad::apply_lines_launch(conf, coef, nlines, nin, nout, M, c->stream);
}

int kad_coef_i_1d(kad_ctx* ctx, int base, int nconf, int64_t nlines, const double* coef, double* conf) {
  if (!ctx || !conf || !coef || nlines < 0 || nconf < 4 || nconf > 64) return KAD_ERR_ARG;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const double* M;
  int nin, nout;
  int rc = get_line_matrix(c, 1, base, nconf, &M, &nin, &nout);
  if (rc) return rc;
  return kad::apply_lines_launch(coef, conf, nlines, nin, nout, M, c->stream);
}

// Host-side tables (no GPU needed).  kind: 0 = coef (analysis), 1 = coef_i
// (synthesis).  out: row-major nout x nin doubles; returns nout*nin or < 0.
int kad_host_transform_matrix(int kind, int base, int nconf, double* out, int cap) {
  try {
    const Mat M = kind == 0 ? coef_matrix(base, nconf) : coef_i_matrix(base, nconf);
    if ((int)M.a.size() > cap) return KAD_ERR_ARG;
    std::copy(M.a.begin(), M.a.end(), out);
    return (int)M.a.size();
  } catch (const std::exception&) {
    return KAD_ERR_ARG;
  }
}

// op: 0 der, 1 mult_sin, 2 mult_cos, 3 div_sin, 4 mult_x, 5 mult_xm1, 6 div_xm1.
int kad_host_op1d_matrix(int op, int base, int n, double* out, int cap, int* out_base) {
  static const Op1d ops[] = {der_1d, mult_sin_1d, mult_cos_1d, div_sin_1d, mult_x_1d, mult_xm1_1d, div_xm1_1d};
  if (op < 0 || op > 6 || n < 4 || n * n > cap) return KAD_ERR_ARG;
  try {
    const Mat M = op_matrix(ops[op], base, n, out_base);
    std::copy(M.a.begin(), M.a.end(), out);
    return n * n;
  } catch (const std::exception&) {
    return KAD_ERR_ARG;
  }
}

}  // extern "C"
