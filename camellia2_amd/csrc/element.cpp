// element.cpp -- element layout, cubature and reference tables (host setup).
#include "element.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>

namespace dpg {
namespace {

// Symmetric eigen-decomposition by cyclic Jacobi rotations (n <= ~24 here).
// A is (n,n) row-major and is destroyed; V receives the eigenvectors as columns.
void jacobi_eigen(int n, std::vector<double>& A, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
    if (off < 1e-34) break;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double apq = A[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- J^T A J
          double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
    }
  }
}

// Legendre P_0..P_{n-1} and derivatives at t
void legendre_all(int n, double t, double* P, double* dP) {
  if (n <= 0) return;
  P[0] = 1.0;
  if (dP) dP[0] = 0.0;
  if (n == 1) return;
  P[1] = t;
  if (dP) dP[1] = 1.0;
  for (int j = 1; j + 1 < n; ++j) {
    P[j + 1] = ((2.0 * j + 1.0) * t * P[j] - j * P[j - 1]) / (j + 1.0);
    if (dP) dP[j + 1] = dP[j - 1] + (2.0 * j + 1.0) * P[j];
  }
}

int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

}  // namespace

Rule1D gauss_legendre(int n) {
  // Golub-Welsch: eigenvalues of the Jacobi matrix of the Legendre recurrence
  std::vector<double> A((size_t)n * n, 0.0), V;
  for (int k = 1; k < n; ++k) {
    double b = k / std::sqrt(4.0 * k * k - 1.0);
    A[(size_t)(k - 1) * n + k] = b;
    A[(size_t)k * n + (k - 1)] = b;
  }
  jacobi_eigen(n, A, V);
  std::vector<std::pair<double, double>> xw(n);
  for (int i = 0; i < n; ++i) xw[i] = {A[(size_t)i * n + i], 2.0 * V[i] * V[i]};
  std::sort(xw.begin(), xw.end());
  Rule1D r;
  for (auto& p : xw) {
    r.x.push_back(p.first);
    r.w.push_back(p.second);
  }
  // symmetric rule: symmetrise to remove the last-bit asymmetry
  for (int i = 0; i < n / 2; ++i) {
    double x = 0.5 * (r.x[n - 1 - i] - r.x[i]);
    double w = 0.5 * (r.w[n - 1 - i] + r.w[i]);
    r.x[i] = -x;
    r.x[n - 1 - i] = x;
    r.w[i] = r.w[n - 1 - i] = w;
  }
  if (n % 2) r.x[n / 2] = 0.0;
  return r;
}

std::vector<double> gll_points(int n) {
  // interior points: zeros of P_n' = Gauss-Jacobi(1,1) nodes (Gegenbauer 3/2 recurrence)
  std::vector<double> x;
  x.push_back(-1.0);
  int m = n - 1;
  if (m > 0) {
    std::vector<double> A((size_t)m * m, 0.0), V;
    for (int k = 1; k < m; ++k) {
      double b = std::sqrt(k * (k + 2.0) / ((2.0 * k + 1.0) * (2.0 * k + 3.0)));
      A[(size_t)(k - 1) * m + k] = b;
      A[(size_t)k * m + (k - 1)] = b;
    }
    jacobi_eigen(m, A, V);
    std::vector<double> in(m);
    for (int i = 0; i < m; ++i) in[i] = A[(size_t)i * m + i];
    std::sort(in.begin(), in.end());
    for (int i = 0; i < m / 2; ++i) {
      double s = 0.5 * (in[m - 1 - i] - in[i]);
      in[i] = -s;
      in[m - 1 - i] = s;
    }
    if (m % 2) in[m / 2] = 0.0;
    x.insert(x.end(), in.begin(), in.end());
  }
  x.push_back(1.0);
  return x;
}

LagrangeLine::LagrangeLine(const std::vector<double>& nodes) : n_((int)nodes.size()) {
  int n = n_;
  // Vandermonde V[i][j] = P_j(x_i); coef = V^{-1} (Gauss-Jordan, partial pivoting)
  std::vector<double> V((size_t)n * n), I((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    legendre_all(n, nodes[i], &V[(size_t)i * n], nullptr);
    I[(size_t)i * n + i] = 1.0;
  }
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(V[(size_t)r * n + c]) > std::fabs(V[(size_t)p * n + c])) p = r;
    for (int j = 0; j < n; ++j) {
      std::swap(V[(size_t)c * n + j], V[(size_t)p * n + j]);
      std::swap(I[(size_t)c * n + j], I[(size_t)p * n + j]);
    }
    double d = V[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) {
      V[(size_t)c * n + j] /= d;
      I[(size_t)c * n + j] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = V[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        V[(size_t)r * n + j] -= f * V[(size_t)c * n + j];
        I[(size_t)r * n + j] -= f * I[(size_t)c * n + j];
      }
    }
  }
  // phi_i(x) = sum_j Vinv[j][i] P_j(x)
  coef_.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) coef_[(size_t)i * n + j] = I[(size_t)j * n + i];
}

void LagrangeLine::eval(double t, double* val, double* der) const {
  double P[32], dP[32];
  legendre_all(n_, t, P, dP);
  for (int i = 0; i < n_; ++i) {
    double v = 0.0, d = 0.0;
    for (int j = 0; j < n_; ++j) {
      v += coef_[(size_t)i * n_ + j] * P[j];
      d += coef_[(size_t)i * n_ + j] * dP[j];
    }
    val[i] = v;
    if (der) der[i] = d;
  }
}

namespace {
// Jacobi P_0..P_n^(alpha,0) and derivatives at x (three-term recurrence and its derivative)
void jacobi_a0(int n, double alpha, double x, double* P, double* dP) {
  P[0] = 1.0;
  dP[0] = 0.0;
  if (n == 0) return;
  P[1] = 0.5 * ((alpha + 2.0) * x + alpha);
  dP[1] = 0.5 * (alpha + 2.0);
  for (int k = 2; k <= n; ++k) {
    double a = 2.0 * k + alpha;
    double c0 = 2.0 * k * (k + alpha) * (a - 2.0);
    double c1 = (a - 1.0) * a * (a - 2.0), c2 = (a - 1.0) * alpha * alpha;
    double c3 = 2.0 * (k + alpha - 1.0) * (k - 1.0) * a;
    P[k] = ((c1 * x + c2) * P[k - 1] - c3 * P[k - 2]) / c0;
    dP[k] = (c1 * P[k - 1] + (c1 * x + c2) * dP[k - 1] - c3 * dP[k - 2]) / c0;
  }
}
}  // namespace

DubinerTri::DubinerTri(int n) : n_(n) {
  for (int d = 0; d <= n; ++d)
    for (int j = 0; j <= d; ++j) {
      int i = d - j;
      i_.push_back(i);
      j_.push_back(j);
      // ||psi_ij||^2 = 1 / (2 (2i+1) (i+j+1)) on the unit triangle
      scale_.push_back(std::sqrt(2.0 * (2.0 * i + 1.0) * (i + j + 1.0)));
    }
}

void DubinerTri::eval(double x, double y, double* val, double* dx, double* dy) const {
  double r = 2.0 * x - 1.0, s = 2.0 * y - 1.0;
  double q = 0.5 * (1.0 - s);  // (1-b)/2, b = s
  double a = (s < 1.0) ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0;
  double Pa[16], dPa[16], Pb[16], dPb[16], qp[17];
  jacobi_a0(n_, 0.0, a, Pa, dPa);
  qp[0] = 1.0;
  for (int k = 1; k <= n_; ++k) qp[k] = qp[k - 1] * q;
  int m = 0;
  for (int d = 0; d <= n_; ++d)
    for (int j = 0; j <= d; ++j, ++m) {
      int i = d - j;
      jacobi_a0(j, 2.0 * i + 1.0, s, Pb, dPb);
      double fa = Pa[i], dfa = dPa[i], g = Pb[j], dg = dPb[j];
      val[m] = scale_[m] * fa * qp[i] * g;
      if (dx || dy) {
        double dr = 0.0, ds = fa * dg * qp[i];
        if (i > 0) {
          dr = dfa * g * qp[i - 1];
          ds += dfa * 0.5 * (1.0 + a) * g * qp[i - 1] - fa * 0.5 * i * g * qp[i - 1];
        }
        if (dx) dx[m] = 2.0 * scale_[m] * dr;
        if (dy) dy[m] = 2.0 * scale_[m] * ds;
      }
    }
}

namespace {
// warp of the 1D equispaced nodes onto the GLL nodes, divided by the edge bubble
double warp_factor(int n, const LagrangeLine& eq, const std::vector<double>& gll, const std::vector<double>& req,
                   double r) {
  if (std::fabs(r) >= 1.0 - 1e-10) return 0.0;
  double l[32];
  eq.eval(r, l, nullptr);
  double w = 0.0;
  for (int i = 0; i <= n; ++i) w += l[i] * (gll[i] - req[i]);
  return w / (1.0 - r * r);
}
}  // namespace

std::vector<double> warp_blend_tri(int n) {
  std::vector<double> out;
  if (n == 0) return {1.0 / 3.0, 1.0 / 3.0};
  static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                    1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
  double alpha = n < 16 ? alpopt[n - 1] : 5.0 / 3.0;
  std::vector<double> gll = gll_points(n), req(n + 1);
  for (int i = 0; i <= n; ++i) req[i] = -1.0 + 2.0 * i / n;
  LagrangeLine eq(req);
  const double s3 = std::sqrt(3.0), c2 = std::cos(2.0 * M_PI / 3.0), s2 = std::sin(2.0 * M_PI / 3.0);
  const double c4 = std::cos(4.0 * M_PI / 3.0), s4 = std::sin(4.0 * M_PI / 3.0);
  for (int row = 0; row <= n; ++row)
    for (int col = 0; col + row <= n; ++col) {
      double L1 = (double)row / n, L3 = (double)col / n, L2 = 1.0 - L1 - L3;
      // equilateral triangle (-1,-1/sqrt3), (1,-1/sqrt3), (0,2/sqrt3)
      double x = -L2 + L3, y = (-L2 - L3 + 2.0 * L1) / s3;
      double w1 = 4.0 * L2 * L3 * warp_factor(n, eq, gll, req, L3 - L2) * (1.0 + (alpha * L1) * (alpha * L1));
      double w2 = 4.0 * L1 * L3 * warp_factor(n, eq, gll, req, L1 - L3) * (1.0 + (alpha * L2) * (alpha * L2));
      double w3 = 4.0 * L1 * L2 * warp_factor(n, eq, gll, req, L2 - L1) * (1.0 + (alpha * L3) * (alpha * L3));
      x += w1 + c2 * w2 + c4 * w3;
      y += s2 * w2 + s4 * w3;
      // back to barycentric coordinates, then to the unit triangle
      double b1 = (s3 * y + 1.0) / 3.0, b2 = (-3.0 * x - s3 * y + 2.0) / 6.0, b3 = (3.0 * x - s3 * y + 2.0) / 6.0;
      double rr = -b2 + b3 - b1, ss = -b2 - b3 + b1;
      out.push_back(0.5 * (rr + 1.0));
      out.push_back(0.5 * (ss + 1.0));
    }
  return out;
}

LagrangeTri::LagrangeTri(int n, const std::vector<double>& nodes) : P_(n) {
  int m = P_.size();
  if ((int)nodes.size() != 2 * m) throw std::invalid_argument("triangle node count");
  // V[i][k] = psi_k(node i); coef = V^-1, by Gauss-Jordan with partial pivoting in
  // 80-bit arithmetic: the nodal (trial) values must be accurate to the last bit
  // because they enter B directly
  std::vector<long double> V((size_t)m * m), I((size_t)m * m, 0.0L);
  std::vector<double> row(m);
  for (int i = 0; i < m; ++i) {
    P_.eval(nodes[2 * i], nodes[2 * i + 1], row.data(), nullptr, nullptr);
    for (int k = 0; k < m; ++k) V[(size_t)i * m + k] = row[k];
    I[(size_t)i * m + i] = 1.0L;
  }
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int r = c + 1; r < m; ++r)
      if (fabsl(V[(size_t)r * m + c]) > fabsl(V[(size_t)p * m + c])) p = r;
    for (int j = 0; j < m; ++j) {
      std::swap(V[(size_t)c * m + j], V[(size_t)p * m + j]);
      std::swap(I[(size_t)c * m + j], I[(size_t)p * m + j]);
    }
    long double d = V[(size_t)c * m + c];
    for (int j = 0; j < m; ++j) {
      V[(size_t)c * m + j] /= d;
      I[(size_t)c * m + j] /= d;
    }
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      long double f = V[(size_t)r * m + c];
      if (f == 0.0L) continue;
      for (int j = 0; j < m; ++j) {
        V[(size_t)r * m + j] -= f * V[(size_t)c * m + j];
        I[(size_t)r * m + j] -= f * I[(size_t)c * m + j];
      }
    }
  }
  // phi_k(x) = sum_j psi_j(x) Vinv[j][k]
  coef_.assign((size_t)m * m, 0.0L);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j) coef_[(size_t)k * m + j] = I[(size_t)j * m + k];
}

void LagrangeTri::eval(double x, double y, double* val) const {
  int m = P_.size();
  std::vector<double> psi(m);
  P_.eval(x, y, psi.data(), nullptr, nullptr);
  for (int k = 0; k < m; ++k) {
    long double v = 0.0L;  // the expansion cancels: sum in 80-bit, round once
    for (int j = 0; j < m; ++j) v += coef_[(size_t)k * m + j] * (long double)psi[j];
    val[k] = (double)v;
  }
}

RefCell make_ref_cell(int topo) {
  RefCell c;
  c.topo = topo;
  if (topo == DPG_QUAD) {
    c.dim = 2;
    c.nverts = 4;
    c.nsides = 4;
    c.verts = {-1, -1, 1, -1, 1, 1, -1, 1};
    c.nodes_per_side = 2;
    c.side_nodes = {0, 1, 1, 2, 2, 3, 3, 0};
  } else if (topo == DPG_HEX) {
    c.dim = 3;
    c.nverts = 8;
    c.nsides = 6;
    c.verts = {-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1};
    c.nodes_per_side = 4;  // shards Hexahedron<8> face node map
    c.side_nodes = {0, 1, 5, 4, 1, 2, 6, 5, 2, 3, 7, 6, 0, 4, 7, 3, 0, 3, 2, 1, 4, 5, 6, 7};
  } else if (topo == DPG_TRI) {
    c.dim = 2;
    c.nverts = 3;
    c.nsides = 3;
    c.verts = {0, 0, 1, 0, 0, 1};  // shards Triangle<3>
    c.nodes_per_side = 2;
    c.side_nodes = {0, 1, 1, 2, 2, 0};
  } else {
    throw std::invalid_argument("unsupported cell topology");
  }
  // affine side maps (CamelliaCellTools::getSubcellParametrization): exact fits
  // of the side's reference nodes (line: -1,1; quad: CCW corners)
  c.side_param.assign((size_t)c.nsides * 3 * c.dim, 0.0);
  for (int s = 0; s < c.nsides; ++s) {
    const int* sn = &c.side_nodes[(size_t)s * c.nodes_per_side];
    for (int d = 0; d < c.dim; ++d) {
      double* p = &c.side_param[(size_t)s * 3 * c.dim];
      if (c.dim == 2) {
        double a = c.verts[sn[0] * 2 + d], b = c.verts[sn[1] * 2 + d];
        p[0 * c.dim + d] = 0.5 * (a + b);
        p[1 * c.dim + d] = 0.5 * (b - a);
      } else {
        double v[4];
        for (int k = 0; k < 4; ++k) v[k] = c.verts[sn[k] * 3 + d];
        p[0 * c.dim + d] = 0.25 * (v[0] + v[1] + v[2] + v[3]);
        p[1 * c.dim + d] = 0.25 * (-v[0] + v[1] + v[2] - v[3]);
        p[2 * c.dim + d] = 0.25 * (-v[0] - v[1] + v[2] + v[3]);
      }
    }
  }
  return c;
}

ElementLayout make_layout(const dpg_form_desc& f, const dpg_bucket_desc& b, bool skeleton_only) {
  ElementLayout L;
  L.topo = b.topology;
  L.rc = make_ref_cell(b.topology);
  L.dim = L.rc.dim;
  if (f.spatial_dim != L.dim) throw std::invalid_argument("form spatial_dim does not match the cell topology");
  if (f.n_trial <= 0 || f.n_trial > DPG_MAX_VARS || f.n_test <= 0 || f.n_test > DPG_MAX_VARS)
    throw std::invalid_argument("bad variable counts");
  L.k = b.h1_order;
  L.kt = b.test_degree;
  if (L.k < 1 || L.kt < 1 || L.kt > 10 || L.k > 10) throw std::invalid_argument("unsupported polynomial order");
  L.field_deg = L.k - 1;
  L.trace_deg = L.k;
  L.flux_deg = L.k - 1;
  int sd = L.dim - 1;
  const bool tri = L.topo == DPG_TRI;
  L.nf = tri ? (L.field_deg + 1) * (L.field_deg + 2) / 2 : ipow(L.field_deg + 1, L.dim);
  L.ntr = ipow(L.trace_deg + 1, sd);
  L.nfl = ipow(L.flux_deg + 1, sd);
  bool hasField = false, hasTrace = false, hasFlux = false, sawSide = false;
  int off = 0;
  for (int v = 0; v < f.n_trial; ++v) {
    int kind = f.trial_kind[v];
    if (kind == DPG_FIELD || kind == DPG_FIELD_VECTOR) {
      if (sawSide) throw std::invalid_argument("field variables must precede traces and fluxes");
      if (!skeleton_only && L.field_deg < 1 && L.dim > 0 && L.k < 2)
        throw std::invalid_argument("field order must be >= 1");
      int nc = kind == DPG_FIELD_VECTOR ? L.dim : 1;
      for (int c = 0; c < nc; ++c) {
        L.blocks.push_back({v, c, -1, kind, off, L.nf});
        off += L.nf;
      }
      hasField = true;
    } else if (kind == DPG_TRACE || kind == DPG_FLUX) {
      sawSide = true;
      int card = kind == DPG_TRACE ? L.ntr : L.nfl;
      if (!skeleton_only && kind == DPG_FLUX && L.flux_deg < 1) throw std::invalid_argument("flux order must be >= 1");
      for (int s = 0; s < L.rc.nsides; ++s) {
        L.blocks.push_back({v, 0, s, kind, off, card});
        off += card;
      }
      (kind == DPG_TRACE ? hasTrace : hasFlux) = true;
    } else {
      throw std::invalid_argument("unknown trial kind");
    }
  }
  L.M = off;
  L.F = 0;
  for (auto& blk : L.blocks)
    if (blk.side < 0) L.F += blk.card;
  L.T = L.M - L.F;
  off = 0;
  for (int v = 0; v < f.n_test; ++v) {
    int card;
    if (f.test_space[v] == DPG_HGRAD)
      card = tri ? (L.kt + 1) * (L.kt + 2) / 2 : ipow(L.kt + 1, L.dim);
    else if (f.test_space[v] == DPG_HDIV)
      card = tri ? L.kt * (L.kt + 2) : (L.dim == 2 ? 2 * L.kt * (L.kt + 1) : 3 * L.kt * L.kt * (L.kt + 1));
    else
      throw std::invalid_argument("unknown test space");
    L.test_off.push_back(off);
    L.test_card.push_back(card);
    for (int i = 0; i < card; ++i) L.var_of_test_dof.push_back(v);
    off += card;
  }
  L.N = off;
  // cubature degrees (BasisCache.cpp:68-84, 236-267, 438-470, 696-731)
  int maxTrial = 0;
  if (hasField) maxTrial = std::max(maxTrial, L.field_deg);
  if (hasFlux) maxTrial = std::max(maxTrial, L.flux_deg);
  if (hasTrace) maxTrial = std::max(maxTrial, L.trace_deg);
  int maxTest = L.kt;
  int sideTrial = hasTrace ? L.trace_deg : (hasFlux ? L.flux_deg : maxTrial);
  L.Q1 = (maxTrial + maxTest) / 2 + 1;
  L.QG1 = (2 * maxTest) / 2 + 1;
  L.Qs1 = (sideTrial + maxTest) / 2 + 1;
  if (tri) {
    // collapsed (Duffy) Gauss rule: the extra (1-b)/2 Jacobian raises the degree
    // in b by one, so floor(d/2)+2 points per direction integrate degree d exactly
    L.Q1 = (maxTrial + maxTest) / 2 + 2;
    L.QG1 = (2 * maxTest) / 2 + 2;
  }
  L.Q = ipow(L.Q1, L.dim);
  L.QG = ipow(L.QG1, L.dim);
  L.Qs = ipow(L.Qs1, sd);
  return L;
}

double flops_per_cell(const dpg_form_desc& f, const ElementLayout& L) {
  double FG = 0.0;
  for (int i = 0; i < f.n_ipterms; ++i) {
    bool used[DPG_MAX_VARS] = {false};
    for (int s = 0; s < f.ipterms[i].n; ++s) used[f.ipterms[i].s[s].test_var] = true;
    double n = 0;
    for (int v = 0; v < f.n_test; ++v)
      if (used[v]) n += L.test_card[v];
    FG += 2.0 * L.QG * n * n;
  }
  double FB = 0.0;
  for (int i = 0; i < f.n_bterms; ++i) {
    const dpg_bterm& t = f.bterms[i];
    int kind = f.trial_kind[t.trial_var];
    bool side = kind == DPG_TRACE || kind == DPG_FLUX || t.test_op == DPG_OP_DOT_NORMAL || t.coef.normal_comp >= 0;
    double card = kind == DPG_TRACE ? L.ntr : (kind == DPG_FLUX ? L.nfl : L.nf);
    double nv = L.test_card[t.test_var];
    FB += side ? L.rc.nsides * 2.0 * L.Qs * card * nv : 2.0 * L.Q * card * nv;
  }
  double N = L.N, M = L.M, F = L.F, T = L.T;
  double Fl = 2.0 * L.Q * N;
  double Fch = N * N * N / 3.0 + N * N * M + N * M * M + N * N + 2.0 * N * M;
  double FS = 2.0 * F * F * F / 3.0 + 2.0 * F * F * T + 2.0 * T * T * F;
  return FG + FB + Fl + Fch + FS;
}

namespace {

// test prims of every test dof at reference point xi: out (dim+1, N)
void test_prims(const dpg_form_desc& f, const ElementLayout& L, const LagrangeLine& Cl, const LagrangeLine& Op,
                const double* xi, double* out) {
  int dim = L.dim, N = L.N;
  double cv[3][16], cd[3][16], ov[3][16];
  for (int a = 0; a < dim; ++a) {
    Cl.eval(xi[a], cv[a], cd[a]);
    Op.eval(xi[a], ov[a], nullptr);
  }
  int nc = Cl.size(), no = Op.size();
  for (int v = 0; v < f.n_test; ++v) {
    int off = L.test_off[v];
    if (f.test_space[v] == DPG_HGRAD) {
      int card = L.test_card[v];
      for (int idx = 0; idx < card; ++idx) {
        int ii[3] = {idx % nc, (idx / nc) % nc, idx / (nc * nc)};
        double val = 1.0;
        for (int a = 0; a < dim; ++a) val *= cv[a][ii[a]];
        out[0 * N + off + idx] = val;
        for (int e = 0; e < dim; ++e) {
          double g = 1.0;
          for (int a = 0; a < dim; ++a) g *= (a == e) ? cd[a][ii[a]] : cv[a][ii[a]];
          out[(1 + e) * N + off + idx] = g;
        }
      }
    } else {
      // HDIV In: component c is closed (GLL) in direction c, open (Gauss) elsewhere,
      // components in order, x fastest within a component
      int bf = 0;
      for (int comp = 0; comp < dim; ++comp) {
        int n[3] = {1, 1, 1};
        for (int a = 0; a < dim; ++a) n[a] = (a == comp) ? nc : no;
        int cnt = n[0] * n[1] * n[2];
        for (int idx = 0; idx < cnt; ++idx) {
          int ii[3] = {idx % n[0], (idx / n[0]) % n[1], idx / (n[0] * n[1])};
          double val = 1.0, dv = 1.0;
          for (int a = 0; a < dim; ++a) {
            if (a == comp) {
              val *= cv[a][ii[a]];
              dv *= cd[a][ii[a]];
            } else {
              val *= ov[a][ii[a]];
              dv *= ov[a][ii[a]];
            }
          }
          for (int a = 0; a < dim; ++a) out[a * N + off + bf] = (a == comp) ? val : 0.0;
          out[dim * N + off + bf] = dv;
          ++bf;
        }
      }
    }
  }
}

// triangle test prims: HGRAD = Dubiner P_kt (v, dv/dx, dv/dy); HDIV = RT_kt as
// (psi, 0), (0, psi) for psi in P_{kt-1}, then (x - xc) psi for the kt Dubiner
// functions of degree exactly kt-1 (their top-degree parts span the homogeneous
// polynomials of degree kt-1, so the set is a basis of RT_kt = P_{kt-1}^2 + x P~_{kt-1}).
// K, rhs, Schur and the energy error do not depend on the choice of test basis.
void test_prims_tri(const dpg_form_desc& f, const ElementLayout& L, const DubinerTri& Pt, const DubinerTri& Pr,
                    const double* xi, double* out) {
  const int N = L.N;
  std::vector<double> v(Pt.size()), dx(Pt.size()), dy(Pt.size());
  std::vector<double> rv(Pr.size()), rdx(Pr.size()), rdy(Pr.size());
  Pt.eval(xi[0], xi[1], v.data(), dx.data(), dy.data());
  Pr.eval(xi[0], xi[1], rv.data(), rdx.data(), rdy.data());
  const double xc = xi[0] - 1.0 / 3.0, yc = xi[1] - 1.0 / 3.0;
  const int nlow = Pr.size(), ntop = Pr.degree() + 1;
  for (int tv = 0; tv < f.n_test; ++tv) {
    int off = L.test_off[tv];
    if (f.test_space[tv] == DPG_HGRAD) {
      for (int m = 0; m < Pt.size(); ++m) {
        out[0 * N + off + m] = v[m];
        out[1 * N + off + m] = dx[m];
        out[2 * N + off + m] = dy[m];
      }
    } else {
      for (int m = 0; m < nlow; ++m) {
        out[0 * N + off + m] = rv[m];
        out[1 * N + off + m] = 0.0;
        out[2 * N + off + m] = rdx[m];
        out[0 * N + off + nlow + m] = 0.0;
        out[1 * N + off + nlow + m] = rv[m];
        out[2 * N + off + nlow + m] = rdy[m];
      }
      for (int t = 0; t < ntop; ++t) {
        int m = nlow - ntop + t, b = off + 2 * nlow + t;
        out[0 * N + b] = xc * rv[m];
        out[1 * N + b] = yc * rv[m];
        out[2 * N + b] = 2.0 * rv[m] + xc * rdx[m] + yc * rdy[m];
      }
    }
  }
}

void scalar_values(int dim, const LagrangeLine& B, const double* xi, double* out) {
  double v[3][16];
  for (int a = 0; a < dim; ++a) B.eval(xi[a], v[a], nullptr);
  int n = B.size();
  int card = ipow(n, dim);
  for (int idx = 0; idx < card; ++idx) {
    int ii[3] = {idx % n, (idx / n) % n, idx / (n * n)};
    double val = 1.0;
    for (int a = 0; a < dim; ++a) val *= v[a][ii[a]];
    out[idx] = val;
  }
}

void vertex_basis(const RefCell& rc, const double* xi, double* Nv, double* dN) {
  if (rc.topo == DPG_TRI) {  // barycentric: 1-x-y, x, y
    Nv[0] = 1.0 - xi[0] - xi[1];
    Nv[1] = xi[0];
    Nv[2] = xi[1];
    const double g[6] = {-1.0, -1.0, 1.0, 0.0, 0.0, 1.0};
    for (int k = 0; k < 6; ++k) dN[k] = g[k];
    return;
  }
  for (int v = 0; v < rc.nverts; ++v) {
    const double* p = &rc.verts[(size_t)v * rc.dim];
    double fac[3], g[3];
    for (int d = 0; d < rc.dim; ++d) {
      fac[d] = 0.5 * (1.0 + p[d] * xi[d]);
      g[d] = 0.5 * p[d];
    }
    double val = 1.0;
    for (int d = 0; d < rc.dim; ++d) val *= fac[d];
    Nv[v] = val;
    for (int e = 0; e < rc.dim; ++e) {
      double gr = g[e];
      for (int d = 0; d < rc.dim; ++d)
        if (d != e) gr *= fac[d];
      dN[(size_t)v * rc.dim + e] = gr;
    }
  }
}

struct TriBases {
  DubinerTri Pt, Pr;  // test degree kt and kt-1
  LagrangeTri Fld;    // field: nodal on warp-and-blend nodes of degree k-1
};

void fill_pointset(const dpg_form_desc& f, const ElementLayout& L, const LagrangeLine& Cl, const LagrangeLine& Op,
                   const LagrangeLine& Fld, const TriBases* tb, PointSet& ps) {
  int dim = L.dim, V = L.rc.nverts;
  ps.R.assign((size_t)ps.npts * (dim + 1) * L.N, 0.0);
  ps.dN.assign((size_t)ps.npts * V * dim, 0.0);
  ps.Nv.assign((size_t)ps.npts * V, 0.0);
  ps.field.assign((size_t)ps.npts * L.nf, 0.0);
  for (int q = 0; q < ps.npts; ++q) {
    const double* xi = &ps.xi[(size_t)q * dim];
    if (tb) {
      test_prims_tri(f, L, tb->Pt, tb->Pr, xi, &ps.R[(size_t)q * (dim + 1) * L.N]);
      tb->Fld.eval(xi[0], xi[1], &ps.field[(size_t)q * L.nf]);
    } else {
      test_prims(f, L, Cl, Op, xi, &ps.R[(size_t)q * (dim + 1) * L.N]);
      if (L.field_deg >= 1 || L.nf == 1) scalar_values(dim, Fld, xi, &ps.field[(size_t)q * L.nf]);
    }
    vertex_basis(L.rc, xi, &ps.Nv[(size_t)q * V], &ps.dN[(size_t)q * V * dim]);
  }
}

// collapsed Gauss rule on the unit triangle, n x n points, b (y) outer
void tri_points(int n, std::vector<double>& xi, std::vector<double>& w) {
  Rule1D r = gauss_legendre(n);
  xi.assign((size_t)n * n * 2, 0.0);
  w.assign((size_t)n * n, 0.0);
  for (int jb = 0; jb < n; ++jb)
    for (int ia = 0; ia < n; ++ia) {
      double a = r.x[ia], b = r.x[jb];
      int q = jb * n + ia;
      xi[2 * q + 0] = 0.25 * (1.0 + a) * (1.0 - b);
      xi[2 * q + 1] = 0.5 * (1.0 + b);
      w[q] = r.w[ia] * r.w[jb] * (1.0 - b) * 0.125;  // (1-b)/2 Jacobian, 1/4 to the unit triangle
    }
}

void tensor_points(int dim, const Rule1D& r, std::vector<double>& xi, std::vector<double>& w) {
  int n = (int)r.x.size();
  int np = ipow(n, dim);
  xi.assign((size_t)np * dim, 0.0);
  w.assign(np, 0.0);
  for (int q = 0; q < np; ++q) {
    int rem = q;
    double wt = 1.0;
    for (int d = 0; d < dim; ++d) {
      int id = rem % n;
      rem /= n;
      xi[(size_t)q * dim + d] = r.x[id];
      wt *= r.w[id];
    }
    w[q] = wt;
  }
}

}  // namespace

RefTables make_tables(const dpg_form_desc& f, const ElementLayout& L) {
  RefTables T;
  int dim = L.dim;
  LagrangeLine Cl(gll_points(L.kt));
  LagrangeLine Op(gauss_legendre(L.kt).x);
  LagrangeLine Fld(gll_points(std::max(L.field_deg, 1)));
  std::unique_ptr<TriBases> tb;
  if (L.topo == DPG_TRI) {
    tb.reset(new TriBases());
    tb->Pt = DubinerTri(L.kt);
    tb->Pr = DubinerTri(L.kt - 1);
    tb->Fld = LagrangeTri(L.field_deg, warp_blend_tri(L.field_deg));
  }
  // volume stiffness cache
  if (tb) tri_points(L.Q1, T.vol.xi, T.vol.w);
  else tensor_points(dim, gauss_legendre(L.Q1), T.vol.xi, T.vol.w);
  T.vol.npts = L.Q;
  fill_pointset(f, L, Cl, Op, Fld, tb.get(), T.vol);
  // Gram cache
  if (tb) tri_points(L.QG1, T.ip.xi, T.ip.w);
  else tensor_points(dim, gauss_legendre(L.QG1), T.ip.xi, T.ip.w);
  T.ip.npts = L.QG;
  fill_pointset(f, L, Cl, Op, Fld, tb.get(), T.ip);
  // sides
  std::vector<double> sxi, sw;
  tensor_points(dim - 1, gauss_legendre(L.Qs1), sxi, sw);
  int S = L.rc.nsides;
  T.side.npts = S * L.Qs;
  T.side.xi.assign((size_t)T.side.npts * dim, 0.0);
  T.side.w.assign(T.side.npts, 0.0);
  for (int s = 0; s < S; ++s) {
    const double* p = &L.rc.side_param[(size_t)s * 3 * dim];
    for (int q = 0; q < L.Qs; ++q) {
      double* xi = &T.side.xi[((size_t)s * L.Qs + q) * dim];
      for (int d = 0; d < dim; ++d) {
        xi[d] = p[0 * dim + d];
        for (int a = 0; a < dim - 1; ++a) xi[d] += p[(1 + a) * dim + d] * sxi[(size_t)q * (dim - 1) + a];
      }
      T.side.w[(size_t)s * L.Qs + q] = sw[q];
    }
  }
  fill_pointset(f, L, Cl, Op, Fld, tb.get(), T.side);
  // side trial bases on the side reference points
  LagrangeLine Tr(gll_points(L.trace_deg));
  T.trace.assign((size_t)L.Qs * L.ntr, 0.0);
  for (int q = 0; q < L.Qs; ++q) scalar_values(dim - 1, Tr, &sxi[(size_t)q * (dim - 1)], &T.trace[(size_t)q * L.ntr]);
  T.flux.assign((size_t)L.Qs * L.nfl, 0.0);
  if (L.flux_deg >= 1) {
    LagrangeLine Fl(gll_points(L.flux_deg));
    for (int q = 0; q < L.Qs; ++q) scalar_values(dim - 1, Fl, &sxi[(size_t)q * (dim - 1)], &T.flux[(size_t)q * L.nfl]);
  }
  return T;
}

}  // namespace dpg
