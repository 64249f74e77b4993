// element.hpp -- element types (buckets), dof layout and reference tables.
//
// The host side of the BasisCache role (BasisCache.cpp:68-267): for one
// element type it fixes the dof layout (DofOrderingFactory.cpp:53-211), the
// cubature (CubatureFactory.cpp:25-53, floor(d/2)+1 Gauss points per
// direction) and tabulates every reference basis at every cubature point, so
// that the device only has to apply per-cell geometry.
#pragma once

#include <string>
#include <vector>

#include "../../include/dpg_capi.h"

namespace dpg {

// 1D rules and Lagrange bases ------------------------------------------------
struct Rule1D {
  std::vector<double> x, w;
};
Rule1D gauss_legendre(int n);         // n points, increasing
std::vector<double> gll_points(int n);  // n+1 Gauss-Lobatto-Legendre points, increasing

// Lagrange basis on arbitrary nodes, represented in the Legendre basis
// (phi_i = sum_j coef[i][j] P_j, from the inverse Vandermonde matrix)
class LagrangeLine {
 public:
  LagrangeLine() = default;
  explicit LagrangeLine(const std::vector<double>& nodes);
  int size() const { return n_; }
  void eval(double t, double* val, double* der) const;

 private:
  int n_ = 0;
  std::vector<double> coef_;  // (n, n)
};

// Triangle bases (reference triangle (0,0),(1,0),(0,1)) -------------------------
// Orthonormal Dubiner basis of P_n: psi_ij(x,y) = P_i(a) ((1-b)/2)^i P_j^(2i+1,0)(b)
// in collapsed coordinates a = 2(1+r)/(1-s)-1, b = s, r = 2x-1, s = 2y-1
// (Koornwinder; Sherwin-Karniadakis, Hesthaven-Warburton 2008, sec. 6.1),
// ordered by total degree then j; normalised to unit L2 norm on the triangle.
class DubinerTri {
 public:
  DubinerTri() = default;
  explicit DubinerTri(int n);
  int degree() const { return n_; }
  int size() const { return (n_ + 1) * (n_ + 2) / 2; }
  // values, d/dx, d/dy (each size()) at (x, y); dx/dy may be NULL
  void eval(double x, double y, double* val, double* dx, double* dy) const;

 private:
  int n_ = 0;
  std::vector<int> i_, j_;
  std::vector<double> scale_;
};

// Warp-and-blend interpolation nodes of degree n on the reference triangle
// (Warburton 2006; Hesthaven-Warburton 2008, Nodes2D), rows bottom to top,
// x fastest within a row: (n+1)(n+2)/2 points (x, y).
std::vector<double> warp_blend_tri(int n);

// Nodal Lagrange basis on given triangle nodes, through the Dubiner Vandermonde.
class LagrangeTri {
 public:
  LagrangeTri() = default;
  LagrangeTri(int n, const std::vector<double>& nodes);
  int size() const { return P_.size(); }
  void eval(double x, double y, double* val) const;

 private:
  DubinerTri P_;
  std::vector<long double> coef_;  // (size, size): phi_k = sum_m coef[k][m] psi_m
};

// Reference cell data ---------------------------------------------------------
struct RefCell {
  int topo = 0, dim = 0, nverts = 0, nsides = 0;
  std::vector<double> verts;  // (V, dim)
  std::vector<int> side_nodes;  // (S, nodes per side)
  int nodes_per_side = 0;
  // affine side parametrisation xi = c0 + c1 s0 + c2 s1: (S, 3, dim)
  std::vector<double> side_param;
};
RefCell make_ref_cell(int topo);

// Element layout (one bucket) -------------------------------------------------
struct TrialBlock {  // a contiguous run of trial dofs sharing one scalar basis
  int var, comp, side;  // side = -1 for field blocks
  int kind;             // DPG_FIELD / DPG_FIELD_VECTOR / DPG_TRACE / DPG_FLUX
  int offset, card;
};

struct ElementLayout {
  int topo = 0, dim = 0, k = 0, kt = 0;
  RefCell rc;
  int N = 0, M = 0, F = 0, T = 0;
  int field_deg = 0, trace_deg = 0, flux_deg = 0;
  int nf = 0, ntr = 0, nfl = 0;  // scalar cardinalities (field, trace, flux)
  int Q1 = 0, Qs1 = 0, QG1 = 0;  // points per direction
  int Q = 0, Qs = 0, QG = 0;
  std::vector<int> test_off, test_card;
  std::vector<int> var_of_test_dof;  // (N)
  std::vector<TrialBlock> blocks;    // trial blocks in dof order
};

// Throws std::invalid_argument on an unsupported form/bucket.
// skeleton_only: the layout's block structure for numbering the trace and flux
// dofs only, which also admits H1 order 1 (constant fields and fluxes; the
// assembly itself needs order >= 2)
ElementLayout make_layout(const dpg_form_desc& form, const dpg_bucket_desc& bucket, bool skeleton_only = false);

// Algorithmic flop count per cell (the roofline numerator):
//   F_G  = sum_lt 2 QG (sum of the cardinalities of the test vars of lt)^2
//   F_B  = sum_terms 2 Q card_trial N_testvar  (+ side terms with Qs, per side)
//   F_l  = 2 Q N
//   F_ch = N^3/3 + N^2 M + N M^2 + N^2 + 2 N M
//   F_S  = 2F^3/3 + 2F^2 T + 2 T^2 F
double flops_per_cell(const dpg_form_desc& form, const ElementLayout& L);

// Reference tables ------------------------------------------------------------
// "prims" of a test dof: HGRAD -> (v, dv/dxi_1..dim); HDIV -> (tau_1..dim, div tau)
struct PointSet {
  int npts = 0;
  std::vector<double> xi;     // (npts, dim) volume reference coordinates
  std::vector<double> w;      // (npts) cubature weights
  std::vector<double> R;      // (npts, dim+1, N) test prims
  std::vector<double> dN;     // (npts, V, dim) vertex-basis gradients
  std::vector<double> Nv;     // (npts, V) vertex-basis values
  std::vector<double> field;  // (npts, nf) field scalar basis values
};

struct RefTables {
  PointSet vol;   // stiffness/rhs cache, Q points
  PointSet ip;    // Gram cache, QG points
  PointSet side;  // all sides, S*Qs points (side-major)
  std::vector<double> trace;  // (Qs, ntr) side basis of traces
  std::vector<double> flux;   // (Qs, nfl) side basis of fluxes
};
RefTables make_tables(const dpg_form_desc& form, const ElementLayout& L);

}  // namespace dpg
