// forms.cpp -- host mirror of Camellia's symbolic-form surface (VarFactory,
// BF::addTerm, BF::graphNorm, IP::addTerm, RHS::addTerm) compiled into the
// dpg_form_desc term tables of the C ABI, and the formulations used by the
// workloads.
//
// Vector-valued products are expanded into scalar products per component
// (sigma . tau = sum_d sigma_d tau_d), which is how the contraction
// Intrepid::FunctionSpaceTools::integrate sums ranks (LinearTerm.cpp:576).
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dpg_capi.h"
#include "errors.hpp"

namespace dpg {

struct TrialRef {
  int var, comp;
};
struct TestRef {
  int var, op;
};

inline dpg_coef coef(double scale, int slot = -1, int fn = DPG_FN_VALUE, int normal = -1) {
  dpg_coef c;
  c.scale = scale;
  c.coef_slot = slot;
  c.coef_fn = fn;
  c.normal_comp = normal;
  c.pad = 0;
  return c;
}

// VarFactory::fieldVar/traceVar/fluxVar/testVar (VarFactory.cpp:62-141): IDs
// in creation order, trial and test numbered separately.
class VarFactory {
 public:
  int fieldVar(bool vector = false) { return addTrial(vector ? DPG_FIELD_VECTOR : DPG_FIELD); }
  int traceVar() { return addTrial(DPG_TRACE); }
  int fluxVar() { return addTrial(DPG_FLUX); }
  int testVar(int space) {
    if ((int)test_.size() >= DPG_MAX_VARS) throw std::invalid_argument("too many test variables");
    test_.push_back(space);
    return (int)test_.size() - 1;
  }
  const std::vector<int>& trialKinds() const { return trial_; }
  const std::vector<int>& testSpaces() const { return test_; }

 private:
  int addTrial(int kind) {
    if ((int)trial_.size() >= DPG_MAX_VARS) throw std::invalid_argument("too many trial variables");
    trial_.push_back(kind);
    return (int)trial_.size() - 1;
  }
  std::vector<int> trial_, test_;
};

class IP {
 public:
  void addTerm(const std::vector<dpg_lt_summand>& lt) {
    if (lt.empty()) return;
    if (lt.size() > DPG_MAX_LT_SUMMANDS) throw std::invalid_argument("IP term has too many summands");
    terms_.push_back(lt);
  }
  const std::vector<std::vector<dpg_lt_summand>>& terms() const { return terms_; }

 private:
  std::vector<std::vector<dpg_lt_summand>> terms_;
};

class RHS {
 public:
  void addTerm(int loadComp, TestRef v, double scale = 1.0) { terms_.push_back({v.var, v.op, loadComp, 0, scale}); }
  const std::vector<dpg_rhsterm>& terms() const { return terms_; }

 private:
  std::vector<dpg_rhsterm> terms_;
};

class BF {
 public:
  BF(const VarFactory& vf, int spaceDim) : vf_(vf), dim_(spaceDim) {}
  void addTerm(TrialRef u, TestRef v, dpg_coef c = coef(1.0)) { terms_.push_back({u.var, u.comp, v.var, v.op, c}); }

  // TBF::graphNorm (BF.cpp:1039-1141) with unit weights: for every field
  // variable (and component), the sum of the non-boundary test summands it
  // multiplies, then the L2 term of every test variable.
  IP graphNorm() const {
    IP ip;
    for (int v = 0; v < (int)vf_.trialKinds().size(); ++v) {
      int kind = vf_.trialKinds()[v];
      if (kind != DPG_FIELD && kind != DPG_FIELD_VECTOR) continue;
      int nc = kind == DPG_FIELD_VECTOR ? dim_ : 1;
      for (int comp = 0; comp < nc; ++comp) {
        std::vector<dpg_lt_summand> lt;
        for (const auto& t : terms_) {
          if (t.trial_var != v || t.trial_comp != comp) continue;
          if (t.test_op == DPG_OP_DOT_NORMAL || t.coef.normal_comp >= 0) continue;
          dpg_lt_summand s;
          s.test_var = t.test_var;
          s.test_op = t.test_op;
          s.coef = t.coef;
          lt.push_back(s);
        }
        ip.addTerm(lt);
      }
    }
    for (int t = 0; t < (int)vf_.testSpaces().size(); ++t) {
      int nc = vf_.testSpaces()[t] == DPG_HGRAD ? 1 : dim_;
      for (int comp = 0; comp < nc; ++comp) {
        dpg_lt_summand s;
        s.test_var = t;
        s.test_op = vf_.testSpaces()[t] == DPG_HGRAD ? DPG_OP_VALUE : DPG_OP_X + comp;
        s.coef = coef(1.0);
        ip.addTerm({s});
      }
    }
    return ip;
  }

  dpg_form_desc compile(const IP& ip, const RHS& rhs, int nCoefs) const {
    dpg_form_desc f;
    std::memset(&f, 0, sizeof(f));
    f.spatial_dim = dim_;
    f.n_coefs = nCoefs;
    f.n_trial = (int)vf_.trialKinds().size();
    f.n_test = (int)vf_.testSpaces().size();
    for (int i = 0; i < f.n_trial; ++i) f.trial_kind[i] = vf_.trialKinds()[i];
    for (int i = 0; i < f.n_test; ++i) f.test_space[i] = vf_.testSpaces()[i];
    if (terms_.size() > DPG_MAX_BTERMS) throw std::invalid_argument("too many bilinear terms");
    f.n_bterms = (int)terms_.size();
    for (size_t i = 0; i < terms_.size(); ++i) f.bterms[i] = terms_[i];
    if (ip.terms().size() > DPG_MAX_IPTERMS) throw std::invalid_argument("too many IP terms");
    f.n_ipterms = (int)ip.terms().size();
    for (size_t i = 0; i < ip.terms().size(); ++i) {
      f.ipterms[i].n = (int)ip.terms()[i].size();
      for (size_t s = 0; s < ip.terms()[i].size(); ++s) f.ipterms[i].s[s] = ip.terms()[i][s];
    }
    if (rhs.terms().size() > DPG_MAX_RHS_TERMS) throw std::invalid_argument("too many rhs terms");
    f.n_rhs = (int)rhs.terms().size();
    for (size_t i = 0; i < rhs.terms().size(); ++i) f.rhs[i] = rhs.terms()[i];
    return f;
  }

 private:
  const VarFactory& vf_;
  int dim_;
  std::vector<dpg_bterm> terms_;
};

// PoissonFormulation(spaceDim, useConformingTraces, ULTRAWEAK), PoissonFormulation.cpp:30-88
//   b = (u, div tau) + (sigma, tau) - <u_hat, tau.n> - (sigma, grad v) + <sigma_n_hat, v>
dpg_form_desc poisson_uw(int dim) {
  VarFactory vf;
  int u = vf.fieldVar(), sigma = vf.fieldVar(true), uhat = vf.traceVar(), sn = vf.fluxVar();
  int v = vf.testVar(DPG_HGRAD), tau = vf.testVar(DPG_HDIV);
  BF bf(vf, dim);
  bf.addTerm({u, 0}, {tau, DPG_OP_DIV});
  for (int d = 0; d < dim; ++d) bf.addTerm({sigma, d}, {tau, DPG_OP_X + d});
  bf.addTerm({uhat, 0}, {tau, DPG_OP_DOT_NORMAL}, coef(-1.0));
  for (int d = 0; d < dim; ++d) bf.addTerm({sigma, d}, {v, DPG_OP_DX + d}, coef(-1.0));
  bf.addTerm({sn, 0}, {v, DPG_OP_VALUE});
  RHS rhs;
  rhs.addTerm(0, {v, DPG_OP_VALUE});  // PoissonFormulation::rhs: f * v
  return bf.compile(bf.graphNorm(), rhs, 0);
}

// ConvectionDiffusionReactionFormulation ULTRAWEAK, ConvectionDiffusionReactionFormulation.cpp:71-111,
// with per-cell coefficients [eps, beta_1..beta_dim, alpha]:
//   b = (sigma, tau) + (sqrt(eps) u, div tau) - <sqrt(eps) uhat, tau.n>
//     + (sqrt(eps) sigma - beta u, grad v) - <sigma_n, v> + (alpha u, v)
dpg_form_desc cdr_uw(int dim) {
  VarFactory vf;
  int u = vf.fieldVar(), sigma = vf.fieldVar(true), uhat = vf.traceVar(), sn = vf.fluxVar();
  int v = vf.testVar(DPG_HGRAD), tau = vf.testVar(DPG_HDIV);
  const int EPS = 0, BETA = 1, ALPHA = 1 + dim;
  BF bf(vf, dim);
  for (int d = 0; d < dim; ++d) bf.addTerm({sigma, d}, {tau, DPG_OP_X + d});
  bf.addTerm({u, 0}, {tau, DPG_OP_DIV}, coef(1.0, EPS, DPG_FN_SQRT));
  bf.addTerm({uhat, 0}, {tau, DPG_OP_DOT_NORMAL}, coef(-1.0, EPS, DPG_FN_SQRT));
  for (int d = 0; d < dim; ++d) bf.addTerm({sigma, d}, {v, DPG_OP_DX + d}, coef(1.0, EPS, DPG_FN_SQRT));
  for (int d = 0; d < dim; ++d) bf.addTerm({u, 0}, {v, DPG_OP_DX + d}, coef(-1.0, BETA + d));
  bf.addTerm({sn, 0}, {v, DPG_OP_VALUE}, coef(-1.0));
  bf.addTerm({u, 0}, {v, DPG_OP_VALUE}, coef(1.0, ALPHA));
  RHS rhs;
  rhs.addTerm(0, {v, DPG_OP_VALUE});
  return bf.compile(bf.graphNorm(), rhs, dim + 2);
}

// ConvectionDiffusionFormulation, ConvectionDiffusionFormulation.cpp:23-96, per-cell [eps, beta_1..beta_d]:
//   b = (sigma/eps, tau) + (u, div tau) - <uhat, tau.n> + (-beta u + sigma, grad v) + <tc, v>
dpg_form_desc convection_diffusion_uw(int dim) {
  VarFactory vf;
  int u = vf.fieldVar(), sigma = vf.fieldVar(true), uhat = vf.traceVar(), tc = vf.fluxVar();
  int v = vf.testVar(DPG_HGRAD), tau = vf.testVar(DPG_HDIV);
  const int EPS = 0, BETA = 1;
  BF bf(vf, dim);
  for (int d = 0; d < dim; ++d) bf.addTerm({sigma, d}, {tau, DPG_OP_X + d}, coef(1.0, EPS, DPG_FN_INV));
  bf.addTerm({u, 0}, {tau, DPG_OP_DIV});
  bf.addTerm({uhat, 0}, {tau, DPG_OP_DOT_NORMAL}, coef(-1.0));
  for (int d = 0; d < dim; ++d) bf.addTerm({u, 0}, {v, DPG_OP_DX + d}, coef(-1.0, BETA + d));
  for (int d = 0; d < dim; ++d) bf.addTerm({sigma, d}, {v, DPG_OP_DX + d});
  bf.addTerm({tc, 0}, {v, DPG_OP_VALUE});
  RHS rhs;
  rhs.addTerm(0, {v, DPG_OP_VALUE});
  return bf.compile(bf.graphNorm(), rhs, dim + 1);
}

// StokesVGPFormulation (steady, conforming traces), StokesVGPFormulation.cpp:156-408; rhs :1015-1028
dpg_form_desc stokes_vgp_uw(int dim, double mu) {
  if (dim != 2) throw std::invalid_argument("stokes: only spaceDim 2 is built");
  VarFactory vf;
  int U[2], SIG[2][2], UH[2], TN[2];
  U[0] = vf.fieldVar();
  U[1] = vf.fieldVar();
  int P = vf.fieldVar();
  SIG[0][0] = vf.fieldVar();
  SIG[0][1] = vf.fieldVar();
  SIG[1][0] = vf.fieldVar();
  SIG[1][1] = vf.fieldVar();
  UH[0] = vf.traceVar();
  UH[1] = vf.traceVar();
  TN[0] = vf.fluxVar();
  TN[1] = vf.fluxVar();
  int V[2], TAU[2];
  V[0] = vf.testVar(DPG_HGRAD);
  V[1] = vf.testVar(DPG_HGRAD);
  TAU[0] = vf.testVar(DPG_HDIV);
  TAU[1] = vf.testVar(DPG_HDIV);
  int Q = vf.testVar(DPG_HGRAD);
  BF bf(vf, dim);
  for (int d1 = 0; d1 < 2; ++d1) {
    bf.addTerm({U[d1], 0}, {TAU[d1], DPG_OP_DIV});
    for (int d2 = 0; d2 < 2; ++d2) bf.addTerm({SIG[d1][d2], 0}, {TAU[d1], DPG_OP_X + d2}, coef(1.0 / mu));
    bf.addTerm({UH[d1], 0}, {TAU[d1], DPG_OP_DOT_NORMAL}, coef(-1.0));
  }
  for (int d1 = 0; d1 < 2; ++d1) {
    for (int d2 = 0; d2 < 2; ++d2) bf.addTerm({SIG[d1][d2], 0}, {V[d1], DPG_OP_DX + d2});
    bf.addTerm({P, 0}, {V[d1], DPG_OP_DX + d1}, coef(-1.0));
    bf.addTerm({TN[d1], 0}, {V[d1], DPG_OP_VALUE});
  }
  for (int d1 = 0; d1 < 2; ++d1) {
    bf.addTerm({U[d1], 0}, {Q, DPG_OP_DX + d1}, coef(-1.0));
    bf.addTerm({UH[d1], 0}, {Q, DPG_OP_VALUE}, coef(1.0, -1, DPG_FN_VALUE, d1));
  }
  RHS rhs;
  rhs.addTerm(0, {V[0], DPG_OP_VALUE});
  rhs.addTerm(1, {V[1], DPG_OP_VALUE});
  return bf.compile(bf.graphNorm(), rhs, 0);
}

}  // namespace dpg

extern "C" {

int dpg_form_poisson_uw(int spatial_dim, dpg_form_desc* out) {
  return dpg::guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (spatial_dim != 2 && spatial_dim != 3) throw std::invalid_argument("spatial_dim must be 2 or 3");
    *out = dpg::poisson_uw(spatial_dim);
  });
}

int dpg_form_cdr_uw(int spatial_dim, dpg_form_desc* out) {
  return dpg::guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (spatial_dim != 2 && spatial_dim != 3) throw std::invalid_argument("spatial_dim must be 2 or 3");
    *out = dpg::cdr_uw(spatial_dim);
  });
}

int dpg_form_convection_diffusion_uw(int spatial_dim, dpg_form_desc* out) {
  return dpg::guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (spatial_dim != 2 && spatial_dim != 3) throw std::invalid_argument("spatial_dim must be 2 or 3");
    *out = dpg::convection_diffusion_uw(spatial_dim);
  });
}

int dpg_form_stokes_vgp_uw(int spatial_dim, double mu, dpg_form_desc* out) {
  return dpg::guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (!(mu > 0)) throw std::invalid_argument("mu must be positive");
    *out = dpg::stokes_vgp_uw(spatial_dim, mu);
  });
}

int dpg_form_graph_norm(dpg_form_desc* form) {
  return dpg::guarded([&] {
    if (!form) throw std::invalid_argument("form is NULL");
    dpg::VarFactory vf;
    for (int i = 0; i < form->n_trial; ++i) {
      int k = form->trial_kind[i];
      if (k == DPG_FIELD) vf.fieldVar();
      else if (k == DPG_FIELD_VECTOR) vf.fieldVar(true);
      else if (k == DPG_TRACE) vf.traceVar();
      else vf.fluxVar();
    }
    for (int i = 0; i < form->n_test; ++i) vf.testVar(form->test_space[i]);
    dpg::BF bf(vf, form->spatial_dim);
    for (int i = 0; i < form->n_bterms; ++i) {
      const dpg_bterm& t = form->bterms[i];
      bf.addTerm({t.trial_var, t.trial_comp}, {t.test_var, t.test_op}, t.coef);
    }
    dpg::IP ip = bf.graphNorm();
    dpg_form_desc g = bf.compile(ip, dpg::RHS(), form->n_coefs);
    form->n_ipterms = g.n_ipterms;
    std::memcpy(form->ipterms, g.ipterms, sizeof(g.ipterms));
  });
}

}  // extern "C"
