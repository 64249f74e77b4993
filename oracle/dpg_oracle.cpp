// dpg_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of Camellia's DPG local assembly, used as the parity
// reference by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
// bench.py. Nothing in the product library links, loads or calls this file.
//
// What it restates (reference file:line):
//   bases        Intrepid HGRAD_{LINE,QUAD,HEX}_Cn (GLL nodes, x fastest,
//                trilinos-fixes/.../Intrepid_HGRAD_HEX_Cn_FEMDef.hpp:102-138,248-330),
//                HDIV_{QUAD,HEX}_In (closed GLL x open Gauss Lagrange),
//                VectorizedBasis component-major (src/include/VectorizedBasisDef.h:219-223),
//                field HVOL = Cn(k-1), trace Cn(k), flux Cn(k-1) (BasisFactory.cpp:122-420,493-620)
//   cubature     Gauss-Legendre, floor(d/2)+1 points per direction; stiffness
//                degree maxTrial+maxTest, Gram degree 2*maxTest, side degree
//                maxTrialOnSide+maxTest (BasisCache.cpp:68-84,236-267,438-470,696-731)
//   geometry     vertex-basis Jacobian, w*detJ, side measure w*|J t|, unit outward
//                normals, affine side parametrisation (CamelliaCellTools.cpp:134-222,1043-1122,1615-1649;
//                BasisCache.cpp:1798-1894)
//   transforms   HGRAD grad J^-T, HDIV value J/detJ, div/detJ, HVOL identity
//                (BasisEvaluation.cpp:391-650)
//   B, G, l      per-term contractions (LinearTerm.cpp:428-723, IP.cpp:215-292, RHS.cpp:68-82),
//                flux parity (LinearTerm.cpp:1086-1107)
//   K, rhs       factoredCholeskySolve: POEQU scaling, POTRF, unscale, TRSM, SYRK, TRSM, GEMV
//                (BF.cpp:923-1018)
//   Schur        CondensedDofInterpreter.cpp:1063-1192 (LU with partial pivoting)
//   energy error Solution.cpp:2653-2798 (r = l - B u, Cholesky solve G e = r, r^T e)
//   graph norm   BF.cpp:1039-1141
//   formulations PoissonFormulation.cpp:30-88, ConvectionDiffusionReactionFormulation.cpp:71-111,
//                ConvectionDiffusionFormulation.cpp:23-96,
//                StokesVGPFormulation.cpp:372-408,1015-1028
//
// Loops are written for clarity first; OpenMP spreads cells over threads (the
// reference itself is single-threaded and scales by MPI rank over cells).

#include "../include/dpg_capi.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using std::vector;

// ---------------------------------------------------------------- 1D rules
// Legendre P_n and P_{n-1} at x by the three-term recurrence.
void legendre(int n, double x, double& pn, double& pnm1) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { pn = 1.0; pnm1 = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  pn = p1;
  pnm1 = p0;
}

// n-point Gauss-Legendre rule on [-1,1] (Newton on P_n), increasing order.
void gauss_rule(int n, vector<double>& x, vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double xi = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double pn, pnm1;
      legendre(n, xi, pn, pnm1);
      double dp = n * (xi * pn - pnm1) / (xi * xi - 1.0);
      double dx = pn / dp;
      xi -= dx;
      if (std::fabs(dx) < 1e-16) break;
    }
    double pn, pnm1;
    legendre(n, xi, pn, pnm1);
    double dp = n * (xi * pn - pnm1) / (xi * xi - 1.0);
    x[i] = xi;
    w[i] = 2.0 / ((1.0 - xi * xi) * dp * dp);
  }
}

// n+1 Gauss-Lobatto-Legendre nodes (roots of (1-x^2) P_n'), increasing.
void gll_nodes(int n, vector<double>& x) {
  x.assign(n + 1, 0.0);
  if (n == 1) { x[0] = -1.0; x[1] = 1.0; return; }
  for (int i = 0; i <= n; ++i) {
    double xi = -std::cos(M_PI * i / n);
    if (i == 0 || i == n) { x[i] = xi; continue; }
    for (int it = 0; it < 200; ++it) {
      // Newton on q(x) = P_{n+1} - P_{n-1} (proportional to (1-x^2) P_n')
      double pn, pnm1;
      legendre(n, xi, pn, pnm1);
      double pnp1 = ((2.0 * n + 1.0) * xi * pn - n * pnm1) / (n + 1.0);
      double q = pnp1 - pnm1;
      double dq = (2.0 * n + 1.0) * pn;  // d/dx (P_{n+1} - P_{n-1}) = (2n+1) P_n
      double dx = q / dq;
      xi -= dx;
      if (std::fabs(dx) < 1e-16) break;
    }
    x[i] = xi;
  }
}

// Lagrange basis on the given nodes: values and first derivatives at t.
struct Lagrange1D {
  vector<double> nodes;
  int size() const { return (int)nodes.size(); }
  void eval(double t, double* val, double* der) const {
    int n = size();
    for (int i = 0; i < n; ++i) {
      double v = 1.0, d = 0.0;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        double den = nodes[i] - nodes[j];
        // product rule: d/dt prod (t - x_j)/den
        d = d * (t - nodes[j]) / den + v / den;
        v *= (t - nodes[j]) / den;
      }
      val[i] = v;
      if (der) der[i] = d;
    }
  }
};

Lagrange1D closed_basis(int n) {  // Cn(n) on GLL nodes (POINTTYPE_SPECTRAL)
  Lagrange1D b;
  gll_nodes(n, b.nodes);
  return b;
}
Lagrange1D open_basis(int n) {  // degree n-1 on n Gauss points (POINTTYPE_SPECTRAL_OPEN)
  Lagrange1D b;
  vector<double> w;
  gauss_rule(n, b.nodes, w);
  return b;
}

// ------------------------------------------------------------- reference cells
struct RefCell {
  int dim, nverts, nsides;
  vector<double> verts;            // (V, dim)
  vector<vector<int>> side_nodes;  // per side
};

RefCell ref_cell(int topo) {
  RefCell c;
  if (topo == DPG_TRI) {  // shards Triangle<3>
    c.dim = 2; c.nverts = 3; c.nsides = 3;
    c.verts = {0, 0, 1, 0, 0, 1};
    c.side_nodes = {{0, 1}, {1, 2}, {2, 0}};
  } else if (topo == DPG_QUAD) {
    c.dim = 2; c.nverts = 4; c.nsides = 4;
    c.verts = {-1, -1, 1, -1, 1, 1, -1, 1};
    c.side_nodes = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  } else {
    c.dim = 3; c.nverts = 8; c.nsides = 6;
    c.verts = {-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1};
    c.side_nodes = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};
  }
  return c;
}

// vertex (C1) basis values and reference gradients at xi
void vertex_basis(const RefCell& rc, const double* xi, double* N, double* dN /*(V,dim)*/) {
  if (rc.nverts == 3) {  // linear triangle
    N[0] = 1.0 - xi[0] - xi[1]; N[1] = xi[0]; N[2] = xi[1];
    dN[0] = -1.0; dN[1] = -1.0; dN[2] = 1.0; dN[3] = 0.0; dN[4] = 0.0; dN[5] = 1.0;
    return;
  }
  for (int v = 0; v < rc.nverts; ++v) {
    const double* p = &rc.verts[v * rc.dim];
    double f[3], g[3];
    for (int d = 0; d < rc.dim; ++d) {
      f[d] = 0.5 * (1.0 + p[d] * xi[d]);
      g[d] = 0.5 * p[d];
    }
    double val = 1.0;
    for (int d = 0; d < rc.dim; ++d) val *= f[d];
    N[v] = val;
    for (int e = 0; e < rc.dim; ++e) {
      double gr = g[e];
      for (int d = 0; d < rc.dim; ++d)
        if (d != e) gr *= f[d];
      dN[v * rc.dim + e] = gr;
    }
  }
}

// affine side parametrisation xi = c0 + sum_k c_{k+1} s_k  (least squares over the side nodes)
void side_param(const RefCell& rc, int side, double c[3][3] /*[coef][dim]*/) {
  const vector<int>& sn = rc.side_nodes[side];
  for (int d = 0; d < rc.dim; ++d) {
    if (rc.dim == 2) {
      double a = rc.verts[sn[0] * 2 + d], b = rc.verts[sn[1] * 2 + d];
      c[0][d] = 0.5 * (a + b);
      c[1][d] = 0.5 * (b - a);
      c[2][d] = 0.0;
    } else {
      double n0 = rc.verts[sn[0] * 3 + d], n1 = rc.verts[sn[1] * 3 + d];
      double n2 = rc.verts[sn[2] * 3 + d], n3 = rc.verts[sn[3] * 3 + d];
      c[0][d] = 0.25 * (n0 + n1 + n2 + n3);
      c[1][d] = 0.25 * (-n0 + n1 + n2 - n3);
      c[2][d] = 0.25 * (-n0 - n1 + n2 + n3);
    }
  }
}

// ------------------------------------------------------------------ spaces
int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

struct Element {
  int topo, dim, k, kt;
  RefCell rc;
  // trial layout
  vector<int> trial_offset, trial_card;  // per var: first dof, cardinality per side (or total for fields)
  int M = 0, F = 0, N = 0;
  vector<int> test_offset, test_card;
  int field_deg, trace_deg, flux_deg;
  int Q1, Qs1, QG1;  // points per direction
  int maxTrialDeg, maxTestDeg;
};

int scalar_card(int dim, int n) { return ipow(n + 1, dim); }
int hdiv_card(int dim, int n) { return dim == 2 ? 2 * n * (n + 1) : 3 * n * n * (n + 1); }
int tri_field_card(int n) { return (n + 1) * (n + 2) / 2; }

bool build_element(const dpg_form_desc* f, const dpg_bucket_desc* b, Element& e) {
  e.topo = b->topology;
  e.rc = ref_cell(e.topo);
  e.dim = e.rc.dim;
  if (f->spatial_dim != e.dim) return false;
  e.k = b->h1_order;
  e.kt = b->test_degree;
  e.field_deg = e.k - 1;
  e.trace_deg = e.k;
  e.flux_deg = e.k - 1;
  int sideDim = e.dim - 1;
  int off = 0;
  bool hasTrace = false, hasFlux = false, hasField = false;
  int nf = 0;
  for (int v = 0; v < f->n_trial; ++v) {
    int kind = f->trial_kind[v];
    e.trial_offset.push_back(off);
    int card;
    int fc = e.topo == DPG_TRI ? tri_field_card(e.field_deg) : scalar_card(e.dim, e.field_deg);
    if (kind == DPG_FIELD) { card = fc; off += card; hasField = true; nf += card; }
    else if (kind == DPG_FIELD_VECTOR) { card = fc; off += e.dim * card; hasField = true; nf += e.dim * card; }
    else if (kind == DPG_TRACE) { card = scalar_card(sideDim, e.trace_deg); off += e.rc.nsides * card; hasTrace = true; }
    else { card = scalar_card(sideDim, e.flux_deg); off += e.rc.nsides * card; hasFlux = true; }
    e.trial_card.push_back(card);
  }
  e.M = off;
  e.F = nf;
  off = 0;
  for (int v = 0; v < f->n_test; ++v) {
    int card = f->test_space[v] == DPG_HGRAD ? scalar_card(e.dim, e.kt) : hdiv_card(e.dim, e.kt);
    if (e.topo == DPG_TRI) card = f->test_space[v] == DPG_HGRAD ? tri_field_card(e.kt) : e.kt * (e.kt + 2);
    e.test_offset.push_back(off);
    e.test_card.push_back(card);
    off += card;
  }
  e.N = off;
  // DofOrdering::maxBasisDegree over trial bases
  int mt = 0;
  if (hasField) mt = std::max(mt, e.field_deg);
  if (hasFlux) mt = std::max(mt, e.flux_deg);
  if (hasTrace) mt = std::max(mt, e.trace_deg);
  e.maxTrialDeg = mt;
  e.maxTestDeg = e.kt;
  e.Q1 = (e.maxTrialDeg + e.maxTestDeg) / 2 + 1;
  e.QG1 = (2 * e.maxTestDeg) / 2 + 1;
  int sideTrial = hasTrace ? e.trace_deg : (hasFlux ? e.flux_deg : e.maxTrialDeg);
  e.Qs1 = (sideTrial + e.maxTestDeg) / 2 + 1;
  if (e.topo == DPG_TRI) {  // collapsed rule: one more point per direction for the (1-b) Jacobian
    e.Q1 = (e.maxTrialDeg + e.maxTestDeg) / 2 + 2;
    e.QG1 = (2 * e.maxTestDeg) / 2 + 2;
  }
  // fields must precede traces for the condensed (Schur) output
  return true;
}

// tensor cubature on [-1,1]^dim with n points per direction, x fastest
void tensor_rule(int dim, int n, vector<double>& pts, vector<double>& wts) {
  vector<double> x, w;
  gauss_rule(n, x, w);
  int np = ipow(n, dim);
  pts.assign(np * dim, 0.0);
  wts.assign(np, 0.0);
  for (int q = 0; q < np; ++q) {
    int r = q;
    double wt = 1.0;
    for (int d = 0; d < dim; ++d) {
      int id = r % n;
      r /= n;
      pts[q * dim + d] = x[id];
      wt *= w[id];
    }
    wts[q] = wt;
  }
}

// scalar tensor Lagrange basis Cn(n) at a point: values (card) and reference gradient (card, dim)
void scalar_tensor(int dim, const Lagrange1D& L, const double* xi, double* val, double* grad) {
  int n1 = L.size();
  double v[3][16], d[3][16];
  for (int a = 0; a < dim; ++a) L.eval(xi[a], v[a], d[a]);
  int card = ipow(n1, dim);
  for (int idx = 0; idx < card; ++idx) {
    int i = idx % n1, j = (idx / n1) % n1, k = idx / (n1 * n1);
    int ii[3] = {i, j, k};
    double val_ = 1.0;
    for (int a = 0; a < dim; ++a) val_ *= v[a][ii[a]];
    if (val) val[idx] = val_;
    if (grad) {
      for (int e = 0; e < dim; ++e) {
        double g = 1.0;
        for (int a = 0; a < dim; ++a) g *= (a == e) ? d[a][ii[a]] : v[a][ii[a]];
        grad[idx * dim + e] = g;
      }
    }
  }
}

// HDIV In(n): reference vector values (card, dim) and reference divergence (card)
void hdiv_tensor(int dim, const Lagrange1D& C, const Lagrange1D& O, const double* xi, double* val, double* div) {
  double cv[3][16], cd[3][16], ov[3][16];
  for (int a = 0; a < dim; ++a) {
    C.eval(xi[a], cv[a], cd[a]);
    O.eval(xi[a], ov[a], nullptr);
  }
  int nc = C.size(), no = O.size();
  int bf = 0;
  for (int comp = 0; comp < dim; ++comp) {
    // component `comp` is closed in direction comp and open elsewhere; x fastest
    int n[3];
    for (int a = 0; a < dim; ++a) n[a] = (a == comp) ? nc : no;
    int cnt = 1;
    for (int a = 0; a < dim; ++a) cnt *= n[a];
    for (int idx = 0; idx < cnt; ++idx) {
      int ii[3] = {idx % n[0], (idx / n[0]) % n[1], dim == 3 ? idx / (n[0] * n[1]) : 0};
      double v = 1.0, dv = 1.0;
      for (int a = 0; a < dim; ++a) {
        if (a == comp) { v *= cv[a][ii[a]]; dv *= cd[a][ii[a]]; }
        else { v *= ov[a][ii[a]]; dv *= ov[a][ii[a]]; }
      }
      for (int a = 0; a < dim; ++a) val[bf * dim + a] = (a == comp) ? v : 0.0;
      div[bf] = dv;
      ++bf;
    }
  }
}

// ------------------------------------------------------------ triangles
// (Intrepid's HGRAD_TRI_Cn / HDIV_TRI_In are not vendored.) Test spaces: any
// basis of P_n and RT_n gives the same K, rhs, Schur and energy error, so the
// oracle uses its own: orthogonal Koornwinder polynomials (unnormalised,
// ordered i outer, j inner) and RT_n = {(p,0),(0,p)} for p in P_{n-1}
// interleaved, plus x p_top for the degree-(n-1) members, x measured from
// vertex 0. Field (trial) basis: Lagrange on warp-and-blend nodes of degree p
// (Warburton 2006), rows bottom to top, x fastest.

// Jacobi P_0..P_n^(a,b)(x), three-term recurrence
void jacobi(int n, double a, double b, double x, double* P) {
  P[0] = 1.0;
  if (n == 0) return;
  P[1] = 0.5 * (a - b + (a + b + 2.0) * x);
  for (int k = 1; k < n; ++k) {
    double c = 2.0 * k + a + b;
    double a1 = 2.0 * (k + 1) * (k + a + b + 1.0) * c;
    double a2 = (c + 1.0) * (a * a - b * b);
    double a3 = c * (c + 1.0) * (c + 2.0);
    double a4 = 2.0 * (k + a) * (k + b) * (c + 2.0);
    P[k + 1] = ((a2 + a3 * x) * P[k] - a4 * P[k - 1]) / a1;
  }
}

// d/dx P_m^(a,b) = (m+a+b+1)/2 P_{m-1}^(a+1,b+1)
double jacobi_der(int m, double a, double b, double x) {
  if (m == 0) return 0.0;
  double P[24];
  jacobi(m - 1, a + 1.0, b + 1.0, x, P);
  return 0.5 * (m + a + b + 1.0) * P[m - 1];
}

int tri_card(int n) { return (n + 1) * (n + 2) / 2; }

// Koornwinder basis of P_n at (x,y): values and (d/dx, d/dy), order i outer, j inner
void koornwinder(int n, double x, double y, double* val, double* gx, double* gy) {
  double r = 2.0 * x - 1.0, s = 2.0 * y - 1.0;
  double h = 0.5 * (1.0 - s);
  double a = (h > 0.0) ? (1.0 + r) / h - 1.0 : -1.0;
  double La[24];
  jacobi(n, 0.0, 0.0, a, La);
  int m = 0;
  for (int i = 0; i <= n; ++i) {
    double Jb[24];
    jacobi(n - i, 2.0 * i + 1.0, 0.0, s, Jb);
    double hi = std::pow(h, i), him1 = i > 0 ? std::pow(h, i - 1) : 0.0;
    double dLa = jacobi_der(i, 0.0, 0.0, a);
    for (int j = 0; j + i <= n; ++j, ++m) {
      double g = Jb[j], dg = jacobi_der(j, 2.0 * i + 1.0, 0.0, s);
      val[m] = La[i] * hi * g;
      if (gx) {
        // d/dr [La(a) h^i] = La'(a) h^(i-1);  d/ds = La'(a) (1+a)/2 h^(i-1) - La(a) i/2 h^(i-1)
        double dr = dLa * him1 * g;
        double ds = (dLa * 0.5 * (1.0 + a) - La[i] * 0.5 * i) * him1 * g + La[i] * hi * dg;
        gx[m] = 2.0 * dr;
        gy[m] = 2.0 * ds;
      }
    }
  }
}

// warp-and-blend nodes of degree n (rows bottom to top, x fastest), on the unit triangle
void warp_blend_nodes(int n, vector<double>& out) {
  out.clear();
  if (n == 0) { out = {1.0 / 3.0, 1.0 / 3.0}; return; }
  static const double alpopt[15] = {0.0, 0.0, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                    1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
  double alpha = n < 16 ? alpopt[n - 1] : 5.0 / 3.0;
  vector<double> gl;
  gll_nodes(n, gl);
  vector<double> eqn(n + 1);
  for (int i = 0; i <= n; ++i) eqn[i] = -1.0 + 2.0 * i / n;
  Lagrange1D eq;
  eq.nodes = eqn;
  auto warp = [&](double t) {  // (interpolated GLL shift) / (1 - t^2), zero at the ends
    if (std::fabs(t) >= 1.0 - 1e-10) return 0.0;
    double l[32];
    eq.eval(t, l, nullptr);
    double w = 0.0;
    for (int i = 0; i <= n; ++i) w += l[i] * (gl[i] - eqn[i]);
    return w / (1.0 - t * t);
  };
  // equilateral reference: vertices v0 (-1,-1/sqrt3), v1 (1,-1/sqrt3), v2 (0, 2/sqrt3);
  // each edge's warp acts along that edge's direction
  const double r3 = std::sqrt(3.0);
  const double V[3][2] = {{-1.0, -1.0 / r3}, {1.0, -1.0 / r3}, {0.0, 2.0 / r3}};
  for (int row = 0; row <= n; ++row)
    for (int col = 0; col + row <= n; ++col) {
      double l2 = (double)row / n, l1 = (double)col / n, l0 = 1.0 - l1 - l2;  // barycentric of v0, v1, v2
      double bc[3] = {l0, l1, l2};
      double X = l0 * V[0][0] + l1 * V[1][0] + l2 * V[2][0];
      double Y = l0 * V[0][1] + l1 * V[1][1] + l2 * V[2][1];
      // edge (p -> q), opposite vertex o: blend 4 l_p l_q, warp of (l_q - l_p), direction (v_q - v_p)/2
      const int E[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
      for (int k = 0; k < 3; ++k) {
        int p = E[k][0], q = E[k][1], o = E[k][2];
        double w = 4.0 * bc[p] * bc[q] * warp(bc[q] - bc[p]) * (1.0 + (alpha * bc[o]) * (alpha * bc[o]));
        X += w * 0.5 * (V[q][0] - V[p][0]);
        Y += w * 0.5 * (V[q][1] - V[p][1]);
      }
      // back to the unit triangle through the barycentric coordinates of (X, Y)
      double b2 = (Y - V[0][1]) / (V[2][1] - V[0][1]);
      double b1 = 0.5 * (X - V[0][0] - b2 * (V[2][0] - V[0][0]));
      out.push_back(b1);
      out.push_back(b2);
    }
}

// Lagrange basis on the warp-and-blend nodes of degree n: phi = psi^T C,
// C = V^-1 with V[i][m] = psi_m(node i), solved in long double
struct LagrangeTri {
  int n = 0;
  vector<long double> C;  // (card, card), phi_k = sum_m psi_m C[m][k]
  void build(int deg) {
    n = deg;
    int m = tri_card(n);
    vector<double> nodes;
    warp_blend_nodes(n, nodes);
    vector<long double> A((size_t)m * m), X((size_t)m * m, 0.0L);
    vector<double> psi(m);
    for (int i = 0; i < m; ++i) {
      koornwinder(n, nodes[2 * i], nodes[2 * i + 1], psi.data(), nullptr, nullptr);
      for (int k = 0; k < m; ++k) A[(size_t)i * m + k] = psi[k];
      X[(size_t)i * m + i] = 1.0L;
    }
    for (int c = 0; c < m; ++c) {
      int p = c;
      for (int r = c + 1; r < m; ++r)
        if (std::fabs((double)A[(size_t)r * m + c]) > std::fabs((double)A[(size_t)p * m + c])) p = r;
      for (int j = 0; j < m; ++j) {
        std::swap(A[(size_t)c * m + j], A[(size_t)p * m + j]);
        std::swap(X[(size_t)c * m + j], X[(size_t)p * m + j]);
      }
      long double d = A[(size_t)c * m + c];
      for (int j = 0; j < m; ++j) { A[(size_t)c * m + j] /= d; X[(size_t)c * m + j] /= d; }
      for (int r = 0; r < m; ++r) {
        if (r == c) continue;
        long double f = A[(size_t)r * m + c];
        for (int j = 0; j < m; ++j) { A[(size_t)r * m + j] -= f * A[(size_t)c * m + j]; X[(size_t)r * m + j] -= f * X[(size_t)c * m + j]; }
      }
    }
    // A V^-1 = I: X holds V^-1 (rows m, cols i); phi_k = sum_m psi_m Vinv[m][k]
    C.assign(X.begin(), X.end());
  }
  void eval(double x, double y, double* val) const {
    int m = tri_card(n);
    vector<double> psi(m);
    koornwinder(n, x, y, psi.data(), nullptr, nullptr);
    for (int k = 0; k < m; ++k) {
      long double v = 0.0L;  // 80-bit sum: the expansion cancels
      for (int a = 0; a < m; ++a) v += (long double)psi[a] * C[(size_t)a * m + k];
      val[k] = (double)v;
    }
  }
};

// collapsed Gauss rule on the unit triangle: n points in a (x-like) and in b (y), weight (1-b)/8
void tri_rule(int n, vector<double>& pts, vector<double>& wts) {
  vector<double> x, w;
  gauss_rule(n, x, w);
  pts.assign((size_t)n * n * 2, 0.0);
  wts.assign((size_t)n * n, 0.0);
  for (int jb = 0; jb < n; ++jb)
    for (int ia = 0; ia < n; ++ia) {
      int q = jb * n + ia;
      double a = x[ia], b = x[jb];
      pts[2 * q] = 0.25 * (1.0 + a) * (1.0 - b);
      pts[2 * q + 1] = 0.5 * (1.0 + b);
      wts[q] = w[ia] * w[jb] * 0.125 * (1.0 - b);
    }
}

// Triangle test bases. Variant 0 (default): the Koornwinder order and scaling
// above, RT interleaved, top-degree complement x p measured from vertex 0.
// Variant 1: the same polynomials orthonormalised (norm^2 = 1/(2(2i+1)(i+j+1))),
// ordered by total degree then j, RT as all (p,0), all (p,0)... then (x - c) p_top
// with c the centroid. Both span P_n and RT_n, so K does not depend on the variant;
// the two differ only in how G and B round.
static int g_tri_basis = 0;

// Koornwinder values/gradients in the order and scaling of the active variant
void tri_poly(int n, double x, double y, double* val, double* gx, double* gy) {
  int m = tri_card(n);
  if (g_tri_basis == 0) { koornwinder(n, x, y, val, gx, gy); return; }
  vector<double> v(m), a(m), b(m);
  koornwinder(n, x, y, v.data(), a.data(), b.data());
  // i-outer index of (i, j): sum_{i' < i} (n - i' + 1) + j
  auto io = [&](int i, int j) { return i * (n + 1) - i * (i - 1) / 2 + j; };
  int k = 0;
  for (int d = 0; d <= n; ++d)
    for (int j = 0; j <= d; ++j, ++k) {
      int i = d - j, src = io(i, j);
      double sc = std::sqrt(2.0 * (2.0 * i + 1.0) * (i + j + 1.0));
      val[k] = sc * v[src];
      if (gx) { gx[k] = sc * a[src]; gy[k] = sc * b[src]; }
    }
}

// triangle test values in the reference frame: HGRAD (card) + grads (card, 2); HDIV (card, 2) + div (card)
void tri_hgrad(int n, const double* xi, double* val, double* grad) {
  int m = tri_card(n);
  vector<double> gx(m), gy(m);
  tri_poly(n, xi[0], xi[1], val, gx.data(), gy.data());
  for (int i = 0; i < m; ++i) { grad[2 * i] = gx[i]; grad[2 * i + 1] = gy[i]; }
}
void tri_hdiv(int n, const double* xi, double* val, double* div) {
  int m = tri_card(n - 1);
  vector<double> p(m), gx(m), gy(m);
  tri_poly(n - 1, xi[0], xi[1], p.data(), gx.data(), gy.data());
  int bf = 0;
  if (g_tri_basis == 0) {
    for (int k = 0; k < m; ++k) {
      val[2 * bf] = p[k]; val[2 * bf + 1] = 0.0; div[bf] = gx[k]; ++bf;
      val[2 * bf] = 0.0; val[2 * bf + 1] = p[k]; div[bf] = gy[k]; ++bf;
    }
  } else {
    for (int k = 0; k < m; ++k) { val[2 * bf] = p[k]; val[2 * bf + 1] = 0.0; div[bf] = gx[k]; ++bf; }
    for (int k = 0; k < m; ++k) { val[2 * bf] = 0.0; val[2 * bf + 1] = p[k]; div[bf] = gy[k]; ++bf; }
  }
  // members of exact degree n-1, times the position vector
  double cx = g_tri_basis == 0 ? 0.0 : 1.0 / 3.0;
  double x = xi[0] - cx, y = xi[1] - cx;
  int k = 0;
  if (g_tri_basis == 0) {
    for (int i = 0; i <= n - 1; ++i)
      for (int j = 0; j + i <= n - 1; ++j, ++k) {
        if (i + j != n - 1) continue;
        val[2 * bf] = x * p[k]; val[2 * bf + 1] = y * p[k];
        div[bf] = 2.0 * p[k] + x * gx[k] + y * gy[k];
        ++bf;
      }
  } else {
    for (k = m - n; k < m; ++k) {  // the last n in degree order have degree n-1
      val[2 * bf] = x * p[k]; val[2 * bf + 1] = y * p[k];
      div[bf] = 2.0 * p[k] + x * gx[k] + y * gy[k];
      ++bf;
    }
  }
}

// ------------------------------------------------------------ point geometry
struct Geo {
  double x[3];
  double J[3][3], Jinv[3][3], detJ;
};

void geometry(const RefCell& rc, const double* nodes, const double* xi, Geo& g) {
  double N[8], dN[24];
  vertex_basis(rc, xi, N, dN);
  int dim = rc.dim;
  for (int d = 0; d < dim; ++d) {
    g.x[d] = 0.0;
    for (int e = 0; e < dim; ++e) g.J[d][e] = 0.0;
  }
  for (int v = 0; v < rc.nverts; ++v)
    for (int d = 0; d < dim; ++d) {
      g.x[d] += nodes[v * dim + d] * N[v];
      for (int e = 0; e < dim; ++e) g.J[d][e] += nodes[v * dim + d] * dN[v * dim + e];
    }
  if (dim == 2) {
    g.detJ = g.J[0][0] * g.J[1][1] - g.J[0][1] * g.J[1][0];
    g.Jinv[0][0] = g.J[1][1] / g.detJ;
    g.Jinv[0][1] = -g.J[0][1] / g.detJ;
    g.Jinv[1][0] = -g.J[1][0] / g.detJ;
    g.Jinv[1][1] = g.J[0][0] / g.detJ;
  } else {
    const double(*A)[3] = g.J;
    double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    g.detJ = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    double id = 1.0 / g.detJ;
    g.Jinv[0][0] = c00 * id;
    g.Jinv[1][0] = c01 * id;
    g.Jinv[2][0] = c02 * id;
    g.Jinv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
    g.Jinv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
    g.Jinv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
    g.Jinv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
    g.Jinv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
    g.Jinv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  }
}

// ------------------------------------------------------------------ values
// Physical values of every test operator at one point: tv[op][dof] (N dofs)
struct TestVals {
  int N;
  vector<double> v;  // (9, N)
  double* op(int o) { return &v[(size_t)o * N]; }
};

struct Tabulator {
  const dpg_form_desc* f;
  const Element* e;
  Lagrange1D testC, testO, fieldL, traceL, fluxL;
  LagrangeTri fieldT;
  Tabulator(const dpg_form_desc* f_, const Element* e_) : f(f_), e(e_) {
    testC = closed_basis(e->kt);
    testO = open_basis(e->kt);
    fieldL = closed_basis(e->field_deg);
    if (e->topo == DPG_TRI) fieldT.build(e->field_deg);
    traceL = closed_basis(e->trace_deg);
    if (e->flux_deg >= 1) fluxL = closed_basis(e->flux_deg);
  }
  // field basis values at reference point xi
  void field_values(const double* xi, double* out) const {
    if (e->topo == DPG_TRI) fieldT.eval(xi[0], xi[1], out);
    else scalar_tensor(e->dim, fieldL, xi, out, nullptr);
  }
  // volume cubature with n points per direction
  void volume_rule(int n, vector<double>& pts, vector<double>& wts) const {
    if (e->topo == DPG_TRI) tri_rule(n, pts, wts);
    else tensor_rule(e->dim, n, pts, wts);
  }
  // test values at reference point xi with geometry g; normal may be null (volume)
  void test_values(const double* xi, const Geo& g, const double* normal, TestVals& tv) const {
    int dim = e->dim;
    tv.N = e->N;
    tv.v.assign((size_t)9 * e->N, 0.0);
    vector<double> val((size_t)3 * e->N), grad((size_t)3 * e->N), div(e->N);
    for (int t = 0; t < f->n_test; ++t) {
      int off = e->test_offset[t], card = e->test_card[t];
      if (f->test_space[t] == DPG_HGRAD) {
        if (e->topo == DPG_TRI) tri_hgrad(e->kt, xi, val.data(), grad.data());
        else scalar_tensor(dim, testC, xi, val.data(), grad.data());
        for (int i = 0; i < card; ++i) {
          tv.op(DPG_OP_VALUE)[off + i] = val[i];
          for (int d = 0; d < dim; ++d) {
            double s = 0.0;  // (J^-T grad_ref)_d = sum_e Jinv[e][d] * dref_e
            for (int ee = 0; ee < dim; ++ee) s += g.Jinv[ee][d] * grad[i * dim + ee];
            tv.op(DPG_OP_DX + d)[off + i] = s;
          }
        }
      } else {
        if (e->topo == DPG_TRI) tri_hdiv(e->kt, xi, val.data(), div.data());
        else hdiv_tensor(dim, testC, testO, xi, val.data(), div.data());
        for (int i = 0; i < card; ++i) {
          double phys[3];
          for (int d = 0; d < dim; ++d) {
            double s = 0.0;
            for (int ee = 0; ee < dim; ++ee) s += g.J[d][ee] * val[i * dim + ee];
            phys[d] = s / g.detJ;
            tv.op(DPG_OP_X + d)[off + i] = phys[d];
          }
          tv.op(DPG_OP_DIV)[off + i] = div[i] / g.detJ;
          if (normal) {
            double s = 0.0;
            for (int d = 0; d < dim; ++d) s += phys[d] * normal[d];
            tv.op(DPG_OP_DOT_NORMAL)[off + i] = s;
          }
        }
      }
    }
  }
};

double coef_value(const dpg_coef& c, const double* cellCoefs, const double* normal) {
  double v = c.scale;
  if (c.coef_slot >= 0) {
    double a = cellCoefs[c.coef_slot];
    v *= (c.coef_fn == DPG_FN_SQRT) ? std::sqrt(a) : (c.coef_fn == DPG_FN_INV ? 1.0 / a : a);
  }
  if (c.normal_comp >= 0) v *= normal ? normal[c.normal_comp] : 0.0;
  return v;
}

bool bterm_on_side(const dpg_form_desc* f, const dpg_bterm& t) {
  int kind = f->trial_kind[t.trial_var];
  return kind == DPG_TRACE || kind == DPG_FLUX || t.test_op == DPG_OP_DOT_NORMAL || t.coef.normal_comp >= 0;
}

double load_value(const dpg_load_desc* L, int comp, const double* x, int dim) {
  if (!L || comp >= L->ncomp) return 0.0;
  double f = L->constant[comp];
  for (int m = 0; m < L->nmodes[comp]; ++m) {
    double p = L->amp[comp][m];
    for (int d = 0; d < dim; ++d) p *= std::sin(L->k[comp][m][d] * M_PI * x[d]);
    f += p;
  }
  return f;
}

// ----------------------------------------------------- per-cell B, G, l
// B is (M, N) row-major = stiffnessEnriched(c, trial, test); G (N, N); l (N)
template <typename Acc>
void cell_BGl_t(const dpg_form_desc* f, const Element& e, const Tabulator& tab, const double* nodes,
                const int8_t* par, const double* coefs, const dpg_load_desc* load, Acc* B, Acc* G, Acc* l) {
  int dim = e.dim, N = e.N, M = e.M;
  const RefCell& rc = e.rc;
  TestVals tv;
  Geo g;
  vector<double> fv(128), tr(64);
  // ---- volume terms on the stiffness cache
  if (B || l) {
    vector<double> pts, wts;
    tab.volume_rule(e.Q1, pts, wts);
    if (B) std::fill(B, B + (size_t)M * N, Acc(0));
    if (l) std::fill(l, l + N, Acc(0));
    for (size_t q = 0; q < wts.size(); ++q) {
      const double* xi = &pts[q * dim];
      geometry(rc, nodes, xi, g);
      double wm = wts[q] * g.detJ;  // computeCellMeasure
      tab.test_values(xi, g, nullptr, tv);
      if (B) {
        tab.field_values(xi, fv.data());
        for (int ti = 0; ti < f->n_bterms; ++ti) {
          const dpg_bterm& t = f->bterms[ti];
          if (bterm_on_side(f, t)) continue;
          double cf = coef_value(t.coef, coefs, nullptr) * wm;
          int kind = f->trial_kind[t.trial_var];
          int toff = e.trial_offset[t.trial_var] + (kind == DPG_FIELD_VECTOR ? t.trial_comp * e.trial_card[t.trial_var] : 0);
          int card = e.trial_card[t.trial_var];
          int voff = e.test_offset[t.test_var], vcard = e.test_card[t.test_var];
          const double* tvals = tv.op(t.test_op);
          for (int j = 0; j < card; ++j) {
            Acc uj = (Acc)cf * (Acc)fv[j];
            Acc* Brow = B + (size_t)(toff + j) * N;
            for (int i = voff; i < voff + vcard; ++i) Brow[i] += uj * (Acc)tvals[i];
          }
        }
      }
      if (l) {
        for (int ri = 0; ri < f->n_rhs; ++ri) {
          const dpg_rhsterm& r = f->rhs[ri];
          double fx = load_value(load, r.load_comp, g.x, dim) * r.scale * wm;
          int voff = e.test_offset[r.test_var], vcard = e.test_card[r.test_var];
          const double* tvals = tv.op(r.test_op);
          for (int i = voff; i < voff + vcard; ++i) l[i] += (Acc)fx * (Acc)tvals[i];
        }
      }
    }
    // ---- side terms, side by side
    if (B) {
      for (int s = 0; s < rc.nsides; ++s) {
        double c[3][3];
        side_param(rc, s, c);
        vector<double> spts, swts;
        tensor_rule(dim - 1, e.Qs1, spts, swts);
        for (size_t q = 0; q < swts.size(); ++q) {
          double xi[3];
          const double* sp = &spts[q * (dim - 1)];
          for (int d = 0; d < dim; ++d) {
            xi[d] = c[0][d];
            for (int a = 0; a < dim - 1; ++a) xi[d] += c[a + 1][d] * sp[a];
          }
          geometry(rc, nodes, xi, g);
          // physical tangents J c_{a+1}; measure and outward unit normal
          double T[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
          for (int a = 0; a < dim - 1; ++a)
            for (int d = 0; d < dim; ++d) {
              double sacc = 0.0;
              for (int ee = 0; ee < dim; ++ee) sacc += g.J[d][ee] * c[a + 1][ee];
              T[a][d] = sacc;
            }
          double nrm[3], len;
          if (dim == 2) {
            nrm[0] = T[0][1];
            nrm[1] = -T[0][0];
          } else {
            nrm[0] = T[0][1] * T[1][2] - T[0][2] * T[1][1];
            nrm[1] = T[0][2] * T[1][0] - T[0][0] * T[1][2];
            nrm[2] = T[0][0] * T[1][1] - T[0][1] * T[1][0];
          }
          len = 0.0;
          for (int d = 0; d < dim; ++d) len += nrm[d] * nrm[d];
          len = std::sqrt(len);
          for (int d = 0; d < dim; ++d) nrm[d] /= len;
          double wm = swts[q] * len;
          tab.test_values(xi, g, nrm, tv);
          tab.field_values(xi, fv.data());
          for (int ti = 0; ti < f->n_bterms; ++ti) {
            const dpg_bterm& t = f->bterms[ti];
            if (!bterm_on_side(f, t)) continue;
            double cf = coef_value(t.coef, coefs, nrm) * wm;
            int kind = f->trial_kind[t.trial_var];
            int card = e.trial_card[t.trial_var];
            int toff;
            const double* uv;
            if (kind == DPG_TRACE || kind == DPG_FLUX) {
              const Lagrange1D& L = (kind == DPG_TRACE) ? tab.traceL : tab.fluxL;
              scalar_tensor(dim - 1, L, sp, tr.data(), nullptr);
              uv = tr.data();
              toff = e.trial_offset[t.trial_var] + s * card;
              if (kind == DPG_FLUX) cf *= (double)par[s];
            } else {
              uv = fv.data();
              toff = e.trial_offset[t.trial_var] + (kind == DPG_FIELD_VECTOR ? t.trial_comp * card : 0);
            }
            int voff = e.test_offset[t.test_var], vcard = e.test_card[t.test_var];
            const double* tvals = tv.op(t.test_op);
            for (int j = 0; j < card; ++j) {
              Acc uj = (Acc)cf * (Acc)uv[j];
              Acc* Brow = B + (size_t)(toff + j) * N;
              for (int i = voff; i < voff + vcard; ++i) Brow[i] += uj * (Acc)tvals[i];
            }
          }
        }
      }
    }
  }
  // ---- Gram matrix on the IP cache
  if (G) {
    vector<double> pts, wts;
    tab.volume_rule(e.QG1, pts, wts);
    std::fill(G, G + (size_t)N * N, Acc(0));
    vector<double> a(N);
    for (size_t q = 0; q < wts.size(); ++q) {
      const double* xi = &pts[q * dim];
      geometry(rc, nodes, xi, g);
      double wm = wts[q] * g.detJ;
      tab.test_values(xi, g, nullptr, tv);
      for (int li = 0; li < f->n_ipterms; ++li) {
        const dpg_ipterm& lt = f->ipterms[li];
        std::fill(a.begin(), a.end(), 0.0);
        int lo = N, hi = 0;
        for (int si = 0; si < lt.n; ++si) {
          const dpg_lt_summand& s = lt.s[si];
          double cf = coef_value(s.coef, coefs, nullptr);
          int voff = e.test_offset[s.test_var], vcard = e.test_card[s.test_var];
          const double* tvals = tv.op(s.test_op);
          for (int i = voff; i < voff + vcard; ++i) a[i] += cf * tvals[i];
          lo = std::min(lo, voff);
          hi = std::max(hi, voff + vcard);
        }
        for (int i = lo; i < hi; ++i) {
          Acc ai = (Acc)wm * (Acc)a[i];
          if (ai == Acc(0)) continue;
          Acc* Grow = G + (size_t)i * N;
          for (int k = lo; k < hi; ++k) Grow[k] += ai * (Acc)a[k];
        }
      }
    }
  }
}

// Accumulation of B, G, l: FP64 (the reference's arithmetic) or, for the
// extended-precision evaluation, 80-bit sums of the same FP64 point values,
// rounded once at the end.
static int g_accum_extended = 0;

void cell_BGl(const dpg_form_desc* f, const Element& e, const Tabulator& tab, const double* nodes,
              const int8_t* par, const double* coefs, const dpg_load_desc* load, double* B, double* G,
              double* l) {
  if (!g_accum_extended) {
    cell_BGl_t<double>(f, e, tab, nodes, par, coefs, load, B, G, l);
    return;
  }
  vector<long double> Bx(B ? (size_t)e.M * e.N : 0), Gx(G ? (size_t)e.N * e.N : 0), lx(l ? e.N : 0);
  cell_BGl_t<long double>(f, e, tab, nodes, par, coefs, load, B ? Bx.data() : nullptr, G ? Gx.data() : nullptr,
                          l ? lx.data() : nullptr);
  for (size_t i = 0; i < Bx.size(); ++i) B[i] = (double)Bx[i];
  for (size_t i = 0; i < Gx.size(); ++i) G[i] = (double)Gx[i];
  for (size_t i = 0; i < lx.size(); ++i) l[i] = (double)lx[i];
}

// ------------------------------------------------------------ dense kernels
// Column-major lower Cholesky in place; returns LAPACK INFO.
int potrf_lower(int n, double* A) {
  for (int j = 0; j < n; ++j) {
    double* colj = A + (size_t)j * n;
    for (int k = 0; k < j; ++k) {
      const double* colk = A + (size_t)k * n;
      double ljk = colk[j];
      for (int i = j; i < n; ++i) colj[i] -= colk[i] * ljk;
    }
    double d = colj[j];
    if (!(d > 0.0)) return j + 1;
    d = std::sqrt(d);
    colj[j] = d;
    for (int i = j + 1; i < n; ++i) colj[i] /= d;
  }
  return 0;
}

// B (n x m col-major) <- L^-1 B
void trsm_lower(int n, int m, const double* L, double* Bm) {
  for (int c = 0; c < m; ++c) {
    double* b = Bm + (size_t)c * n;
    for (int j = 0; j < n; ++j) {
      double bj = b[j] / L[(size_t)j * n + j];
      b[j] = bj;
      const double* lj = L + (size_t)j * n;
      for (int i = j + 1; i < n; ++i) b[i] -= lj[i] * bj;
    }
  }
}

// b <- L^-T b
void trsv_lower_trans(int n, const double* L, double* b) {
  for (int j = n - 1; j >= 0; --j) {
    const double* lj = L + (size_t)j * n;
    double s = b[j];
    for (int i = j + 1; i < n; ++i) s -= lj[i] * b[i];
    b[j] = s / lj[j];
  }
}

// The same factored solve in 80-bit extended precision (x87 long double), used
// by the tests to measure how far an FP64 evaluation of B^T G^-1 B is from the
// exact value for given G, B, l (a property of the input, not of the code).
template <class R>
void factored_solve_t(int N, int M, const double* ip, const double* Bm, const double* l, double* K, double* rhs) {
  vector<R> G((size_t)N * N), B((size_t)N * M), y(N), s(N);
  for (size_t i = 0; i < G.size(); ++i) G[i] = ip[i];
  for (size_t i = 0; i < B.size(); ++i) B[i] = Bm[i];
  for (int i = 0; i < N; ++i) y[i] = l[i];
  for (int i = 0; i < N; ++i) s[i] = R(1) / std::sqrt((R)G[(size_t)i * N + i]);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) G[(size_t)j * N + i] *= s[i] * s[j];
  for (int j = 0; j < N; ++j) {
    R* colj = &G[(size_t)j * N];
    for (int k = 0; k < j; ++k) {
      const R* colk = &G[(size_t)k * N];
      R ljk = colk[j];
      for (int i = j; i < N; ++i) colj[i] -= colk[i] * ljk;
    }
    R d = std::sqrt(colj[j]);
    colj[j] = d;
    for (int i = j + 1; i < N; ++i) colj[i] /= d;
  }
  for (int j = 0; j < N; ++j)
    for (int i = j; i < N; ++i) G[(size_t)j * N + i] /= s[i];
  auto solve = [&](R* b) {
    for (int j = 0; j < N; ++j) {
      R bj = b[j] / G[(size_t)j * N + j];
      b[j] = bj;
      for (int i = j + 1; i < N; ++i) b[i] -= G[(size_t)j * N + i] * bj;
    }
  };
  for (int c = 0; c < M; ++c) solve(&B[(size_t)c * N]);
  solve(y.data());
  for (int i = 0; i < M; ++i) {
    for (int j = 0; j < M; ++j) {
      R acc = 0;
      for (int k = 0; k < N; ++k) acc += B[(size_t)i * N + k] * B[(size_t)j * N + k];
      K[(size_t)i * M + j] = (double)acc;
    }
    R acc = 0;
    for (int k = 0; k < N; ++k) acc += B[(size_t)i * N + k] * y[k];
    rhs[i] = (double)acc;
  }
}

// LU-based static condensation, templated on the arithmetic
template <class R>
int schur_t(int M, int F, const double* K, const double* rhs, double* S, double* g) {
  int T = M - F;
  // LU with partial pivoting of D = K_FF (row-major copy)
  vector<R> D((size_t)F * F), X((size_t)F * (T + 1));
  for (int i = 0; i < F; ++i)
    for (int j = 0; j < F; ++j) D[(size_t)i * F + j] = K[(size_t)i * M + j];
  for (int i = 0; i < F; ++i) {
    for (int j = 0; j < T; ++j) X[(size_t)i * (T + 1) + j] = K[(size_t)i * M + F + j];
    X[(size_t)i * (T + 1) + T] = rhs[i];
  }
  int info = 0;
  vector<int> piv(F);
  for (int k = 0; k < F; ++k) {
    int p = k;
    R mx = std::fabs(D[(size_t)k * F + k]);
    for (int i = k + 1; i < F; ++i)
      if (std::fabs(D[(size_t)i * F + k]) > mx) { mx = std::fabs(D[(size_t)i * F + k]); p = i; }
    piv[k] = p;
    if (mx == 0.0) { if (!info) info = k + 1; continue; }
    if (p != k) {
      for (int j = 0; j < F; ++j) std::swap(D[(size_t)k * F + j], D[(size_t)p * F + j]);
      for (int j = 0; j <= T; ++j) std::swap(X[(size_t)k * (T + 1) + j], X[(size_t)p * (T + 1) + j]);
    }
    R dkk = D[(size_t)k * F + k];
    for (int i = k + 1; i < F; ++i) {
      R m = D[(size_t)i * F + k] / dkk;
      D[(size_t)i * F + k] = m;
      for (int j = k + 1; j < F; ++j) D[(size_t)i * F + j] -= m * D[(size_t)k * F + j];
      for (int j = 0; j <= T; ++j) X[(size_t)i * (T + 1) + j] -= m * X[(size_t)k * (T + 1) + j];
    }
  }
  // back substitution U X = Y
  for (int i = F - 1; i >= 0; --i) {
    for (int j = 0; j <= T; ++j) {
      R sacc = X[(size_t)i * (T + 1) + j];
      for (int k = i + 1; k < F; ++k) sacc -= D[(size_t)i * F + k] * X[(size_t)k * (T + 1) + j];
      X[(size_t)i * (T + 1) + j] = sacc / D[(size_t)i * F + i];
    }
  }
  for (int a = 0; a < T; ++a) {
    for (int b = 0; b < T; ++b) {
      R sacc = K[(size_t)(F + a) * M + F + b];
      for (int k = 0; k < F; ++k) sacc -= K[(size_t)k * M + F + a] * X[(size_t)k * (T + 1) + b];
      S[(size_t)a * T + b] = (double)sacc;
    }
    R gacc = rhs[F + a];
    for (int k = 0; k < F; ++k) gacc -= K[(size_t)k * M + F + a] * X[(size_t)k * (T + 1) + T];
    g[a] = (double)gacc;
  }
  return info;
}

}  // namespace

// =================================================================== C API
extern "C" {

// TBF::factoredCholeskySolve (BF.cpp:923-1018). ip: N x N column-major (lower
// part used, overwritten by L); B: M x N row-major = N x M column-major
// (overwritten by L^-1 B); l: N (overwritten by L^-1 l); K: M x M; rhs: M.
int oracle_factored_cholesky_solve(int N, int M, double* ip, double* Bm, double* l, double* K, double* rhs) {
  vector<double> s(N);
  for (int i = 0; i < N; ++i) s[i] = 1.0 / std::sqrt(ip[(size_t)i * N + i]);  // DPOEQU
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) ip[(size_t)j * N + i] *= s[i] * s[j];
  int info = potrf_lower(N, ip);
  // unequilibrate L: row i divided by s_i
  for (int j = 0; j < N; ++j)
    for (int i = j; i < N; ++i) ip[(size_t)j * N + i] /= s[i];
  trsm_lower(N, M, ip, Bm);
  for (int i = 0; i < M; ++i)
    for (int j = 0; j <= i; ++j) {
      const double* bi = Bm + (size_t)i * N;
      const double* bj = Bm + (size_t)j * N;
      double acc = 0.0;
      for (int k = 0; k < N; ++k) acc += bi[k] * bj[k];
      K[(size_t)i * M + j] = acc;
      K[(size_t)j * M + i] = acc;
    }
  trsm_lower(N, 1, ip, l);
  for (int i = 0; i < M; ++i) {
    const double* bi = Bm + (size_t)i * N;
    double acc = 0.0;
    for (int k = 0; k < N; ++k) acc += bi[k] * l[k];
    rhs[i] = acc;
  }
  return info;
}

// Extended-precision reference of factoredCholeskySolve; inputs are not modified.
int oracle_factored_solve_extended(int N, int M, const double* ip, const double* Bm, const double* l, double* K,
                                   double* rhs) {
  factored_solve_t<long double>(N, M, ip, Bm, l, K, rhs);
  return 0;
}

// Static condensation of one cell (CondensedDofInterpreter.cpp:1063-1192):
// the first F dofs are fields, the remaining T = M - F are traces/fluxes.
int oracle_schur(int M, int F, const double* K, const double* rhs, double* S, double* g) {
  return schur_t<double>(M, F, K, rhs, S, g);
}

// the same in 80-bit extended precision (test reference)
int oracle_schur_extended(int M, int F, const double* K, const double* rhs, double* S, double* g) {
  return schur_t<long double>(M, F, K, rhs, S, g);
}

int oracle_sizes(const dpg_form_desc* f, const dpg_bucket_desc* b, int32_t* out /*N,M,F,Q1,Qs1,QG1*/) {
  Element e;
  if (!build_element(f, b, e)) return -1;
  out[0] = e.N;
  out[1] = e.M;
  out[2] = e.F;
  out[3] = e.Q1;
  out[4] = e.Qs1;
  out[5] = e.QG1;
  return 0;
}

// 1D rules for tests
int oracle_gauss_rule(int n, double* x, double* w) {
  vector<double> xx, ww;
  gauss_rule(n, xx, ww);
  for (int i = 0; i < n; ++i) { x[i] = xx[i]; w[i] = ww[i]; }
  return 0;
}
// warp-and-blend nodes of degree n on the unit triangle: (n+1)(n+2)/2 (x, y) pairs
int oracle_tri_nodes(int n, double* xy) {
  vector<double> v;
  warp_blend_nodes(n, v);
  for (size_t i = 0; i < v.size(); ++i) xy[i] = v[i];
  return (int)(v.size() / 2);
}

// collapsed Gauss rule with n x n points on the unit triangle
int oracle_tri_rule(int n, double* xy, double* w) {
  vector<double> p, ww;
  tri_rule(n, p, ww);
  for (size_t i = 0; i < p.size(); ++i) xy[i] = p[i];
  for (size_t i = 0; i < ww.size(); ++i) w[i] = ww[i];
  return (int)ww.size();
}

// 1: accumulate B, G, l in 80-bit (extended-precision evaluation); 0: FP64
int oracle_set_extended_accumulation(int on) {
  g_accum_extended = on ? 1 : 0;
  return 0;
}

// selects the triangle test-basis variant (0 default, 1 orthonormal, centroid complement)
int oracle_set_tri_test_basis(int variant) {
  if (variant != 0 && variant != 1) return -1;
  g_tri_basis = variant;
  return 0;
}

int oracle_gll_nodes(int n, double* x) {
  vector<double> xx;
  gll_nodes(n, xx);
  for (int i = 0; i <= n; ++i) x[i] = xx[i];
  return 0;
}

// Full local assembly of `ncells` cells of one element type. Any output may
// be NULL. B is (C, M, N) (stiffnessEnriched), G (C, N, N), l (C, N),
// K (C, M, M), rhs (C, M), S (C, T, T), g (C, T), info (C).
int oracle_assemble(const dpg_form_desc* f, const dpg_bucket_desc* b, int ncells, const double* nodes,
                    const int8_t* parities, const double* coefs, const dpg_load_desc* load, double* Bout,
                    double* Gout, double* lout, double* Kout, double* rhsout, double* Sout, double* gout,
                    int32_t* info, int nthreads) {
  Element e;
  if (!build_element(f, b, e)) return -1;
  int N = e.N, M = e.M, F = e.F, T = M - F;
  int nv = e.rc.nverts, dim = e.dim, ns = e.rc.nsides;
  int nbad = 0;
  Tabulator tab(f, &e);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nbad)
#endif
  for (int c = 0; c < ncells; ++c) {
    vector<double> B((size_t)M * N), G((size_t)N * N), l(N), K((size_t)M * M), rhs(M);
    const double* cn = nodes + (size_t)c * nv * dim;
    const int8_t* cp = parities + (size_t)c * ns;
    const double* cc = coefs ? coefs + (size_t)c * f->n_coefs : nullptr;
    cell_BGl(f, e, tab, cn, cp, cc, load, B.data(), G.data(), l.data());
    if (Bout) std::memcpy(Bout + (size_t)c * M * N, B.data(), sizeof(double) * M * N);
    if (Gout) std::memcpy(Gout + (size_t)c * N * N, G.data(), sizeof(double) * N * N);
    if (lout) std::memcpy(lout + (size_t)c * N, l.data(), sizeof(double) * N);
    if (Kout || rhsout || Sout || gout || info) {
      int inf = oracle_factored_cholesky_solve(N, M, G.data(), B.data(), l.data(), K.data(), rhs.data());
      if (info) info[c] = inf;
      if (inf) nbad += 1;
      if (Kout) std::memcpy(Kout + (size_t)c * M * M, K.data(), sizeof(double) * M * M);
      if (rhsout) std::memcpy(rhsout + (size_t)c * M, rhs.data(), sizeof(double) * M);
      if (Sout || gout) {
        vector<double> S((size_t)T * T), g(T);
        oracle_schur(M, F, K.data(), rhs.data(), S.data(), g.data());
        if (Sout) std::memcpy(Sout + (size_t)c * T * T, S.data(), sizeof(double) * T * T);
        if (gout) std::memcpy(gout + (size_t)c * T, g.data(), sizeof(double) * T);
      }
    }
  }
  return nbad;
}

// Energy error (Solution.cpp:2653-2798): eta2[c] = r^T G^-1 r, r = l - B u.
int oracle_energy_error(const dpg_form_desc* f, const dpg_bucket_desc* b, int ncells, const double* nodes,
                        const int8_t* parities, const double* coefs, const dpg_load_desc* load, const double* u,
                        double* eta2, int nthreads) {
  Element e;
  if (!build_element(f, b, e)) return -1;
  int N = e.N, M = e.M;
  int nv = e.rc.nverts, dim = e.dim, ns = e.rc.nsides;
  int nbad = 0;
  Tabulator tab(f, &e);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nbad)
#endif
  for (int c = 0; c < ncells; ++c) {
    vector<double> B((size_t)M * N), G((size_t)N * N), l(N);
    cell_BGl(f, e, tab, nodes + (size_t)c * nv * dim, parities + (size_t)c * ns,
             coefs ? coefs + (size_t)c * f->n_coefs : nullptr, load, B.data(), G.data(), l.data());
    const double* uc = u + (size_t)c * M;
    vector<double> r(l), ev;
    for (int j = 0; j < M; ++j) {
      const double* Bj = &B[(size_t)j * N];
      for (int i = 0; i < N; ++i) r[i] -= Bj[i] * uc[j];
    }
    ev = r;
    int inf = potrf_lower(N, G.data());  // solveSPDSystemLAPACKCholesky: POTRF + POTRS
    if (inf) nbad += 1;
    trsm_lower(N, 1, G.data(), ev.data());
    trsv_lower_trans(N, G.data(), ev.data());
    double acc = 0.0;
    for (int i = 0; i < N; ++i) acc += r[i] * ev[i];
    eta2[c] = acc;
  }
  return nbad;
}

// TBF::graphNorm (BF.cpp:1039-1141) with unit trial weights and unit L2 test weights.
int oracle_graph_norm(dpg_form_desc* f) {
  int n = 0;
  int dim = f->spatial_dim;
  for (int v = 0; v < f->n_trial; ++v) {
    int kind = f->trial_kind[v];
    if (kind != DPG_FIELD && kind != DPG_FIELD_VECTOR) continue;
    int ncomp = kind == DPG_FIELD_VECTOR ? dim : 1;
    for (int comp = 0; comp < ncomp; ++comp) {
      dpg_ipterm lt;
      std::memset(&lt, 0, sizeof(lt));
      for (int t = 0; t < f->n_bterms; ++t) {
        const dpg_bterm& b = f->bterms[t];
        if (b.trial_var != v || b.trial_comp != comp) continue;
        if (b.test_op == DPG_OP_DOT_NORMAL) continue;  // boundary-only test summand
        if (b.coef.normal_comp >= 0) continue;         // boundary-only weight
        if (lt.n >= DPG_MAX_LT_SUMMANDS) return -2;
        lt.s[lt.n].test_var = b.test_var;
        lt.s[lt.n].test_op = b.test_op;
        lt.s[lt.n].coef = b.coef;
        lt.n++;
      }
      if (lt.n == 0) continue;
      if (n >= DPG_MAX_IPTERMS) return -2;
      f->ipterms[n++] = lt;
    }
  }
  for (int t = 0; t < f->n_test; ++t) {
    int nc = f->test_space[t] == DPG_HGRAD ? 1 : dim;
    for (int comp = 0; comp < nc; ++comp) {
      dpg_ipterm lt;
      std::memset(&lt, 0, sizeof(lt));
      lt.n = 1;
      lt.s[0].test_var = t;
      lt.s[0].test_op = f->test_space[t] == DPG_HGRAD ? DPG_OP_VALUE : DPG_OP_X + comp;
      lt.s[0].coef.scale = 1.0;
      lt.s[0].coef.coef_slot = -1;
      lt.s[0].coef.normal_comp = -1;
      if (n >= DPG_MAX_IPTERMS) return -2;
      f->ipterms[n++] = lt;
    }
  }
  f->n_ipterms = n;
  return 0;
}

// ------------------------------------------------- formulation term tables
static dpg_coef C(double s, int slot = -1, int fn = DPG_FN_VALUE, int nc = -1) {
  dpg_coef c;
  c.scale = s;
  c.coef_slot = slot;
  c.coef_fn = fn;
  c.normal_comp = nc;
  c.pad = 0;
  return c;
}
static void addB(dpg_form_desc* f, int tv, int tc, int sv, int op, dpg_coef c) {
  dpg_bterm& b = f->bterms[f->n_bterms++];
  b.trial_var = tv;
  b.trial_comp = tc;
  b.test_var = sv;
  b.test_op = op;
  b.coef = c;
}

// PoissonFormulation ULTRAWEAK (PoissonFormulation.cpp:30-88)
// trial: u, sigma (vector), u_hat (trace), sigma_n_hat (flux); test: v, tau
int oracle_form_poisson_uw(int dim, dpg_form_desc* f) {
  std::memset(f, 0, sizeof(*f));
  f->spatial_dim = dim;
  f->n_trial = 4;
  f->trial_kind[0] = DPG_FIELD;
  f->trial_kind[1] = DPG_FIELD_VECTOR;
  f->trial_kind[2] = DPG_TRACE;
  f->trial_kind[3] = DPG_FLUX;
  f->n_test = 2;
  f->test_space[0] = DPG_HGRAD;
  f->test_space[1] = DPG_HDIV;
  addB(f, 0, 0, 1, DPG_OP_DIV, C(1.0));                                        // (u, div tau)
  for (int d = 0; d < dim; ++d) addB(f, 1, d, 1, DPG_OP_X + d, C(1.0));        // (sigma, tau)
  addB(f, 2, 0, 1, DPG_OP_DOT_NORMAL, C(-1.0));                                // -(u_hat, tau.n)
  for (int d = 0; d < dim; ++d) addB(f, 1, d, 0, DPG_OP_DX + d, C(-1.0));      // -(sigma, grad v)
  addB(f, 3, 0, 0, DPG_OP_VALUE, C(1.0));                                      // (sigma_n_hat, v)
  f->n_rhs = 1;
  f->rhs[0].test_var = 0;
  f->rhs[0].test_op = DPG_OP_VALUE;
  f->rhs[0].load_comp = 0;
  f->rhs[0].scale = 1.0;
  return oracle_graph_norm(f);
}

// ConvectionDiffusionReactionFormulation ULTRAWEAK (ConvectionDiffusionReactionFormulation.cpp:71-111)
// per-cell coefficient slots: 0 = eps, 1..dim = beta, dim+1 = alpha
int oracle_form_cdr_uw(int dim, dpg_form_desc* f) {
  std::memset(f, 0, sizeof(*f));
  f->spatial_dim = dim;
  f->n_coefs = dim + 2;
  f->n_trial = 4;
  f->trial_kind[0] = DPG_FIELD;
  f->trial_kind[1] = DPG_FIELD_VECTOR;
  f->trial_kind[2] = DPG_TRACE;
  f->trial_kind[3] = DPG_FLUX;
  f->n_test = 2;
  f->test_space[0] = DPG_HGRAD;
  f->test_space[1] = DPG_HDIV;
  for (int d = 0; d < dim; ++d) addB(f, 1, d, 1, DPG_OP_X + d, C(1.0));                     // (sigma, tau)
  addB(f, 0, 0, 1, DPG_OP_DIV, C(1.0, 0, DPG_FN_SQRT));                                   // (sqrt(eps) u, div tau)
  addB(f, 2, 0, 1, DPG_OP_DOT_NORMAL, C(-1.0, 0, DPG_FN_SQRT));                           // -(sqrt(eps) uhat, tau.n)
  for (int d = 0; d < dim; ++d) addB(f, 1, d, 0, DPG_OP_DX + d, C(1.0, 0, DPG_FN_SQRT));  // (sqrt(eps) sigma, grad v)
  for (int d = 0; d < dim; ++d) addB(f, 0, 0, 0, DPG_OP_DX + d, C(-1.0, 1 + d));          // -(beta u, grad v)
  addB(f, 3, 0, 0, DPG_OP_VALUE, C(-1.0));                                                // -(sigma_n, v)
  addB(f, 0, 0, 0, DPG_OP_VALUE, C(1.0, dim + 1));                                        // (alpha u, v)
  f->n_rhs = 1;
  f->rhs[0].test_var = 0;
  f->rhs[0].test_op = DPG_OP_VALUE;
  f->rhs[0].load_comp = 0;
  f->rhs[0].scale = 1.0;
  return oracle_graph_norm(f);
}

// ConvectionDiffusionFormulation (ConvectionDiffusionFormulation.cpp:23-96), per-cell [eps, beta]
int oracle_form_convection_diffusion_uw(int dim, dpg_form_desc* f) {
  std::memset(f, 0, sizeof(*f));
  f->spatial_dim = dim;
  f->n_coefs = dim + 1;
  f->n_trial = 4;
  f->trial_kind[0] = DPG_FIELD;
  f->trial_kind[1] = DPG_FIELD_VECTOR;
  f->trial_kind[2] = DPG_TRACE;
  f->trial_kind[3] = DPG_FLUX;
  f->n_test = 2;
  f->test_space[0] = DPG_HGRAD;
  f->test_space[1] = DPG_HDIV;
  for (int d = 0; d < dim; ++d) addB(f, 1, d, 1, DPG_OP_X + d, C(1.0, 0, DPG_FN_INV));  // (sigma/eps, tau)
  addB(f, 0, 0, 1, DPG_OP_DIV, C(1.0));                                                // (u, div tau)
  addB(f, 2, 0, 1, DPG_OP_DOT_NORMAL, C(-1.0));                                        // -(uhat, tau.n)
  for (int d = 0; d < dim; ++d) addB(f, 0, 0, 0, DPG_OP_DX + d, C(-1.0, 1 + d));        // (-beta u, grad v)
  for (int d = 0; d < dim; ++d) addB(f, 1, d, 0, DPG_OP_DX + d, C(1.0));                // (sigma, grad v)
  addB(f, 3, 0, 0, DPG_OP_VALUE, C(1.0));                                              // (tc, v)
  f->n_rhs = 1;
  f->rhs[0].test_var = 0;
  f->rhs[0].test_op = DPG_OP_VALUE;
  f->rhs[0].load_comp = 0;
  f->rhs[0].scale = 1.0;
  return oracle_graph_norm(f);
}

// StokesVGPFormulation, steady, 2D (StokesVGPFormulation.cpp:372-408, rhs :1015-1028)
// trial: u1 u2 p s11 s12 s21 s22 u1hat u2hat t1n t2n ; test: v1 v2 tau1 tau2 q
int oracle_form_stokes_vgp_uw(int dim, double mu, dpg_form_desc* f) {
  if (dim != 2) return -1;
  std::memset(f, 0, sizeof(*f));
  f->spatial_dim = 2;
  f->n_trial = 11;
  for (int i = 0; i < 7; ++i) f->trial_kind[i] = DPG_FIELD;
  f->trial_kind[7] = f->trial_kind[8] = DPG_TRACE;
  f->trial_kind[9] = f->trial_kind[10] = DPG_FLUX;
  f->n_test = 5;
  f->test_space[0] = f->test_space[1] = DPG_HGRAD;
  f->test_space[2] = f->test_space[3] = DPG_HDIV;
  f->test_space[4] = DPG_HGRAD;
  const int U[2] = {0, 1}, P = 2, SIG[2][2] = {{3, 4}, {5, 6}}, UH[2] = {7, 8}, TN[2] = {9, 10};
  const int V[2] = {0, 1}, TAU[2] = {2, 3}, Q = 4;
  for (int d1 = 0; d1 < 2; ++d1) {
    addB(f, U[d1], 0, TAU[d1], DPG_OP_DIV, C(1.0));
    for (int d2 = 0; d2 < 2; ++d2) addB(f, SIG[d1][d2], 0, TAU[d1], DPG_OP_X + d2, C(1.0 / mu));
    addB(f, UH[d1], 0, TAU[d1], DPG_OP_DOT_NORMAL, C(-1.0));
  }
  for (int d1 = 0; d1 < 2; ++d1) {
    for (int d2 = 0; d2 < 2; ++d2) addB(f, SIG[d1][d2], 0, V[d1], DPG_OP_DX + d2, C(1.0));
    addB(f, P, 0, V[d1], DPG_OP_DX + d1, C(-1.0));
    addB(f, TN[d1], 0, V[d1], DPG_OP_VALUE, C(1.0));
  }
  for (int d1 = 0; d1 < 2; ++d1) {
    addB(f, U[d1], 0, Q, DPG_OP_DX + d1, C(-1.0));
    addB(f, UH[d1], 0, Q, DPG_OP_VALUE, C(1.0, -1, DPG_FN_VALUE, d1));
  }
  f->n_rhs = 2;
  for (int d = 0; d < 2; ++d) {
    f->rhs[d].test_var = V[d];
    f->rhs[d].test_op = DPG_OP_VALUE;
    f->rhs[d].load_comp = d;
    f->rhs[d].scale = 1.0;
  }
  return oracle_graph_norm(f);
}

}  // extern "C"
