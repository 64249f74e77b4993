// hpsystem.hip -- the condensed (trace and flux) system of a 2D hp mesh:
// quads and triangles, hanging sides of a 2:1 refinement, a polynomial order
// per cell. Replaces, for such meshes, the minimum-rule interpretation of the
// cells' local trace and flux dofs (GDAMinimumRule::interpretLocalData,
// GDAMinimumRule.cpp:714-780, with the side constraints of
// BasisReconciliation, BasisReconciliation.cpp:259) and the global insertion of
// TSolution::populateStiffnessAndLoad (Solution.cpp:1263-1293).
//
// The rule, as built here:
//   - a macro edge is a side of the mesh that no other side is part of: a
//     side shared by two cells, a boundary side, or the side of a coarse cell
//     whose neighbours are two fine cells (each fine side is one half of it);
//   - the polynomial order on a macro edge is the minimum of the orders of the
//     cells touching it (the trace order k, the flux order k - 1);
//   - H1 traces have one global dof per regular vertex and k_e - 1 per macro
//     edge (GLL nodes of order k_e, oriented from the smaller vertex id); a
//     hanging vertex (the midpoint of a coarse side) has no dof of its own:
//     its value is the coarse side's trace at its midpoint;
//   - L2 fluxes have k_e dofs per macro edge (GLL nodes of order k_e - 1; the
//     cells' side parities are inside their local matrices).
// Every local trace and flux dof of a cell is the global function of its side
// evaluated at the dof's node: local = M_c global, with M_c from Lagrange
// interpolation (exact, since the global space on a side is a subspace of the
// cell's). The constrained cell matrices M_c^T S_c M_c are assembled through a
// ragged skeleton (dpg_skeleton_create_ragged) into one CSR matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dpg_capi.h"
#include "element.hpp"
#include "errors.hpp"

#define HP_CHECK(expr)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace dpg {
namespace {

template <class T>
struct HpBuf {
  T* p = nullptr;
  size_t n = 0;
  HpBuf() = default;
  HpBuf(const HpBuf&) = delete;
  HpBuf& operator=(const HpBuf&) = delete;
  ~HpBuf() {
    if (p) (void)hipFree(p);
  }
  void upload(const std::vector<T>& h, hipStream_t st) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = h.size();
    HP_CHECK(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)));
    if (n) HP_CHECK(hipMemcpyAsync(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
  }
};

// Lagrange basis function m of the nodes x at t (product formula)
double lagrange(const std::vector<double>& x, int m, double t) {
  double v = 1.0;
  for (size_t j = 0; j < x.size(); ++j)
    if ((int)j != m) v *= (t - x[j]) / (x[m] - x[j]);
  return v;
}

// nodes of a nodal basis of degree n on [0, 1] (GLL; the midpoint for n = 0)
std::vector<double> unit_nodes(int n) {
  if (n == 0) return {0.5};
  std::vector<double> g = gll_points(n);
  for (double& v : g) v = 0.5 * (v + 1.0);
  return g;
}

struct Edge {
  int P = 0, Q = 0;  // vertex ids, P < Q (canonical orientation)
  int order = 1 << 30;
  bool boundary = false;
  int hanging = -1;  // midpoint vertex when the edge is the side of a coarse cell with two fine neighbours
};

struct Entry {
  int32_t g;
  double w;
};

// the constrained numbering of an hp mesh (host)
struct HpNumbering {
  int64_t nglobal = 0;
  int nentities = 0;
  std::vector<int> T, M, F, R, G, bucket_cells;  // per bucket
  std::vector<int32_t> off;                      // (C + 1): the cells' global dofs in dofs
  std::vector<int32_t> dofs;
  std::vector<std::vector<std::vector<Entry>>> rows;  // [cell][local trace/flux dof] -> (global dof, weight)
  std::vector<int32_t> fixed, entity, rep_cell, rep_local;
};

}  // namespace
}  // namespace dpg

struct dpg_hp_system {
  int ncells = 0, nbuckets = 0;
  int64_t nglobal = 0;
  std::vector<int> bucket_cells;              // cells per bucket
  std::vector<int> T, M, F;                   // per bucket: trailing dofs, trial dofs, field dofs
  std::vector<int> R;                         // per bucket: entries per local row (padded)
  std::vector<int> G;                         // per bucket: max global dofs of a cell
  std::vector<int32_t> off_h;                 // (C + 1) offsets into dofs (global cell order)
  std::vector<int32_t> fixed_h, rep_cell_h, rep_local_h, entity_h;
  int nentities = 0;
  // device
  std::vector<std::unique_ptr<dpg::HpBuf<int32_t>>> cell_of;  // per bucket: global cell index of each of its cells
  std::vector<std::unique_ptr<dpg::HpBuf<int32_t>>> mpos;     // per bucket (C_b, T_b, R_b): position in the cell's dofs
  std::vector<std::unique_ptr<dpg::HpBuf<double>>> mw;        // per bucket (C_b, T_b, R_b): weight
  dpg::HpBuf<int32_t> off, dofs;
  dpg::HpBuf<int64_t> off2;
  dpg::HpBuf<double> Shat, ghat;  // constrained blocks and loads, ragged
  dpg_skeleton* sk = nullptr;
  ~dpg_hp_system() {
    if (sk) dpg_skeleton_destroy(sk);
  }
};

namespace dpg {
namespace {

// M_c^T S_c M_c and M_c^T g_c of every cell of one bucket, written at the
// cell's ragged offsets. One workgroup per cell: W = S M (T x G) in LDS with
// lane i owning row i, then Shat = M^T W accumulated in place with lane b
// owning column b; every sum runs in a fixed order.
__global__ __launch_bounds__(256) void constrain_blocks(const double* __restrict__ S, const double* __restrict__ g,
                                                        int T, int R, int Gmax, const int32_t* __restrict__ cell_of,
                                                        const int32_t* __restrict__ mpos,
                                                        const double* __restrict__ mw, const int32_t* __restrict__ off,
                                                        const int64_t* __restrict__ off2, double* __restrict__ Shat,
                                                        double* __restrict__ ghat) {
  extern __shared__ double W[];  // T x Gmax
  const int cb = blockIdx.x, t = threadIdx.x;
  const int c = cell_of[cb];
  const int G = off[c + 1] - off[c];
  const double* Sc = S + (size_t)cb * T * T;
  const int32_t* P = mpos + (size_t)cb * T * R;
  const double* w = mw + (size_t)cb * T * R;
  double* out = Shat + off2[c];
  for (int e = t; e < T * Gmax; e += 256) W[e] = 0.0;
  for (int e = t; e < G * G; e += 256) out[e] = 0.0;
  __syncthreads();
  for (int i = t; i < T; i += 256)
    for (int j = 0; j < T; ++j) {
      const double sij = Sc[(size_t)i * T + j];
      for (int r = 0; r < R; ++r) {
        const int pj = P[j * R + r];
        if (pj >= 0) W[(size_t)i * Gmax + pj] = fma(sij, w[j * R + r], W[(size_t)i * Gmax + pj]);
      }
    }
  __syncthreads();
  for (int b = t; b < G; b += 256)
    for (int i = 0; i < T; ++i) {
      const double wib = W[(size_t)i * Gmax + b];
      for (int r = 0; r < R; ++r) {
        const int pi = P[i * R + r];
        if (pi >= 0) out[(size_t)pi * G + b] = fma(w[i * R + r], wib, out[(size_t)pi * G + b]);
      }
    }
  if (g && t == 0) {
    const double* gc = g + (size_t)cb * T;
    double* go = ghat + off[c];
    for (int a = 0; a < G; ++a) go[a] = 0.0;
    for (int i = 0; i < T; ++i)
      for (int r = 0; r < R; ++r) {
        const int pi = P[i * R + r];
        if (pi >= 0) go[pi] = fma(w[i * R + r], gc[i], go[pi]);
      }
  }
}

// u[cb][F + i] = sum_r w * x[dofs[off[c] + pos]]: the cells' local trace and
// flux coefficients from a global vector
__global__ void local_from_global(const double* __restrict__ x, int ncb, int T, int R, int M, int F,
                                  const int32_t* __restrict__ cell_of, const int32_t* __restrict__ mpos,
                                  const double* __restrict__ mw, const int32_t* __restrict__ off,
                                  const int32_t* __restrict__ dofs, double* __restrict__ u) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)ncb * T) return;
  const int cb = (int)(q / T), i = (int)(q % T);
  const int c = cell_of[cb];
  double v = 0.0;
  for (int r = 0; r < R; ++r) {
    const int p = mpos[q * R + r];
    if (p >= 0) v = fma(mw[q * R + r], x[dofs[off[c] + p]], v);
  }
  u[(size_t)cb * M + F + i] = v;
}

uint64_t pair_key(int a, int b) {
  return ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
}

HpNumbering number_hp(const dpg_form_desc& f, const dpg_bucket_desc* buckets, int nb, int C,
                      const int32_t* cell_bucket, const int32_t* cv, int nvertices, const double* X) {
  HpNumbering hs;
  if (f.spatial_dim != 2) throw std::invalid_argument("hp systems are built for 2D meshes");
  std::vector<ElementLayout> L;
  for (int b = 0; b < nb; ++b) {
    if (buckets[b].topology != DPG_QUAD && buckets[b].topology != DPG_TRI)
      throw std::invalid_argument("hp systems take quads and triangles");
    L.push_back(make_layout(f, buckets[b]));
    if (L.back().T <= 0) throw std::invalid_argument("the form has no trace or flux variables");
  }
  for (int c = 0; c < C; ++c)
    if (cell_bucket[c] < 0 || cell_bucket[c] >= nb) throw std::invalid_argument("cell bucket out of range");
  auto nverts = [&](int c) { return L[cell_bucket[c]].rc.nverts; };
  auto vert = [&](int c, int k) { return cv[(size_t)c * 4 + k]; };
  const int nvtx = nvertices;
  for (int c = 0; c < C; ++c)
    for (int k = 0; k < nverts(c); ++k)
      if (vert(c, k) < 0 || vert(c, k) >= nvtx) throw std::invalid_argument("vertex id out of range");
  // H is the midpoint of the side (A, B): the hanging configuration
  auto midpoint = [&](int a, int b, int h) {
    const double dx = X[2 * b] - X[2 * a], dy = X[2 * b + 1] - X[2 * a + 1];
    const double mx = 0.5 * (X[2 * a] + X[2 * b]) - X[2 * h], my = 0.5 * (X[2 * a + 1] + X[2 * b + 1]) - X[2 * h + 1];
    return std::sqrt(mx * mx + my * my) <= 1e-9 * std::sqrt(dx * dx + dy * dy);
  };
  // sides of the cells by their vertex pair
  std::unordered_map<uint64_t, std::vector<std::pair<int, int>>> sides;  // key -> (cell, side)
  for (int c = 0; c < C; ++c) {
    const RefCell& rc = L[cell_bucket[c]].rc;
    for (int s = 0; s < rc.nsides; ++s)
      sides[pair_key(vert(c, rc.side_nodes[s * 2]), vert(c, rc.side_nodes[s * 2 + 1]))].push_back({c, s});
  }
  // sides with one cell: boundary sides, sides of coarse cells next to two fine
  // cells, and the fine halves; a coarse side (A, B) is split at H when (A, H)
  // and (H, B) are one-cell sides too
  std::unordered_map<int, std::vector<int>> single_adj;  // vertex -> other ends of its one-cell sides
  for (auto& kv : sides) {
    if (kv.second.size() > 2) throw std::invalid_argument("a side belongs to more than two cells");
    if (kv.second.size() == 1) {
      const int a = (int)(kv.first >> 32), b = (int)(kv.first & 0xffffffffu);
      single_adj[a].push_back(b);
      single_adj[b].push_back(a);
    }
  }
  std::unordered_map<uint64_t, int> edge_of;          // side key -> macro edge
  std::unordered_map<uint64_t, int> parent_of;        // half side key -> macro edge
  std::vector<Edge> edges;
  std::vector<int> hanging_edge(nvtx, -1);             // hanging vertex -> its coarse edge
  for (auto& kv : sides) {
    if (kv.second.size() != 1) continue;
    const int a = (int)(kv.first >> 32), b = (int)(kv.first & 0xffffffffu);
    for (int h : single_adj[a]) {
      if (h == b) continue;
      if (!sides.count(pair_key(h, b)) || sides.at(pair_key(h, b)).size() != 1) continue;
      const int cc = kv.second[0].first;
      if (sides.at(pair_key(a, h))[0].first == cc || sides.at(pair_key(h, b))[0].first == cc) continue;
      if (!midpoint(a, b, h)) continue;
      Edge e;
      e.P = a;
      e.Q = b;
      e.hanging = h;
      const int id = (int)edges.size();
      edges.push_back(e);
      edge_of[kv.first] = id;
      parent_of[pair_key(a, h)] = id;
      parent_of[pair_key(h, b)] = id;
      if (hanging_edge[h] >= 0) throw std::invalid_argument("a vertex hangs on two sides");
      hanging_edge[h] = id;
      break;
    }
  }
  for (auto& kv : sides) {
    if (edge_of.count(kv.first) || parent_of.count(kv.first)) continue;
    Edge e;
    e.P = (int)(kv.first >> 32);
    e.Q = (int)(kv.first & 0xffffffffu);
    e.boundary = kv.second.size() == 1;
    edge_of[kv.first] = (int)edges.size();
    edges.push_back(e);
  }
  // a deterministic edge order (sorted by vertex pair) for the numbering
  {
    std::vector<int> order(edges.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
      return pair_key(edges[x].P, edges[x].Q) < pair_key(edges[y].P, edges[y].Q);
    });
    std::vector<int> rank(edges.size());
    for (size_t i = 0; i < order.size(); ++i) rank[order[i]] = (int)i;
    std::vector<Edge> sorted(edges.size());
    for (size_t i = 0; i < edges.size(); ++i) sorted[rank[i]] = edges[i];
    edges.swap(sorted);
    for (auto& kv : edge_of) kv.second = rank[kv.second];
    for (auto& kv : parent_of) kv.second = rank[kv.second];
    for (int& v : hanging_edge)
      if (v >= 0) v = rank[v];
  }
  // minimum rule: the order of a macro edge is the smallest of its cells' orders
  for (int c = 0; c < C; ++c) {
    const ElementLayout& Lc = L[cell_bucket[c]];
    for (int s = 0; s < Lc.rc.nsides; ++s) {
      const uint64_t key = pair_key(vert(c, Lc.rc.side_nodes[s * 2]), vert(c, Lc.rc.side_nodes[s * 2 + 1]));
      const int e = edge_of.count(key) ? edge_of.at(key) : parent_of.at(key);
      edges[e].order = std::min(edges[e].order, Lc.k);
    }
  }
  for (const Edge& e : edges)
    if (e.hanging >= 0 && (hanging_edge[e.P] >= 0 || hanging_edge[e.Q] >= 0))
      throw std::invalid_argument("the end of a coarse side hangs (the mesh is not 2:1 balanced)");
  // global numbering: per trace variable regular vertices then edge interiors;
  // per flux variable the edges' nodes
  std::vector<int> trace_vars, flux_vars;
  for (int v = 0; v < f.n_trial; ++v) {
    if (f.trial_kind[v] == DPG_TRACE) trace_vars.push_back(v);
    if (f.trial_kind[v] == DPG_FLUX) flux_vars.push_back(v);
  }
  const int ne = (int)edges.size();
  std::vector<int> vdof(nvtx, -1), ebase(ne), fbase(ne);
  std::map<int, int> var_base;
  int next = 0, nreg = 0;
  for (int v = 0; v < nvtx; ++v)
    if (hanging_edge[v] < 0) vdof[v] = nreg++;
  int eint = 0, eflux = 0;
  for (int e = 0; e < ne; ++e) {
    ebase[e] = eint;
    eint += edges[e].order - 1;
    fbase[e] = eflux;
    eflux += edges[e].order;
  }
  for (int v : trace_vars) {
    var_base[v] = next;
    next += nreg + eint;
  }
  for (int v : flux_vars) {
    var_base[v] = next;
    next += eflux;
  }
  hs.nglobal = next;
  if (next <= 0) throw std::invalid_argument("no global dofs");
  // the global trace function of edge e as (dof, weight) at parameter t
  // (canonical orientation), hanging end vertices expanded through their edge
  auto trace_at = [&](int var, int e, double t, std::vector<Entry>& out) {
    const Edge& E = edges[e];
    const std::vector<double> x = unit_nodes(E.order);
    for (int m = 0; m <= E.order; ++m) {
      double w = lagrange(x, m, t);
      if (std::fabs(w) < 1e-14) continue;
      if (std::fabs(w - 1.0) < 1e-14) w = 1.0;
      if (m == 0 || m == E.order) {
        const int vx = m == 0 ? E.P : E.Q;
        if (vdof[vx] >= 0) {
          out.push_back({var_base[var] + vdof[vx], w});
        } else {  // a hanging end: the coarse side's trace at its midpoint
          const Edge& H = edges[hanging_edge[vx]];
          const std::vector<double> xh = unit_nodes(H.order);
          for (int mh = 0; mh <= H.order; ++mh) {
            double wh = lagrange(xh, mh, 0.5);
            if (std::fabs(wh) < 1e-14) continue;
            int32_t g;
            if (mh == 0) g = var_base[var] + vdof[H.P];
            else if (mh == H.order) g = var_base[var] + vdof[H.Q];
            else g = var_base[var] + nreg + ebase[hanging_edge[vx]] + (mh - 1);
            out.push_back({g, w * wh});
          }
        }
      } else {
        out.push_back({var_base[var] + nreg + ebase[e] + (m - 1), w});
      }
    }
  };
  auto flux_at = [&](int var, int e, double t, std::vector<Entry>& out) {
    const Edge& E = edges[e];
    const std::vector<double> x = unit_nodes(E.order - 1);
    for (int m = 0; m < E.order; ++m) {
      double w = E.order == 1 ? 1.0 : lagrange(x, m, t);
      if (std::fabs(w) < 1e-14) continue;
      if (std::fabs(w - 1.0) < 1e-14) w = 1.0;
      out.push_back({var_base[var] + fbase[e] + m, w});
    }
  };
  // per cell: rows of M_c, the cell's global dofs
  hs.T.resize(nb);
  hs.M.resize(nb);
  hs.F.resize(nb);
  hs.R.assign(nb, 1);
  hs.G.assign(nb, 1);
  hs.bucket_cells.assign(nb, 0);
  for (int b = 0; b < nb; ++b) {
    hs.T[b] = L[b].T;
    hs.M[b] = L[b].M;
    hs.F[b] = L[b].F;
  }
  std::vector<std::vector<std::vector<Entry>>>& rows = hs.rows;
  rows.resize(C);
  std::vector<std::vector<int32_t>> cdofs(C);
  hs.rep_cell.assign((size_t)next, -1);
  hs.rep_local.assign((size_t)next, -1);
  hs.fixed.assign((size_t)next, 0);
  for (int c = 0; c < C; ++c) {
    const int b = cell_bucket[c];
    const ElementLayout& Lc = L[b];
    hs.bucket_cells[b]++;
    rows[c].assign(Lc.T, {});
    for (const TrialBlock& blk : Lc.blocks) {
      if (blk.side < 0) continue;
      const int a0 = vert(c, Lc.rc.side_nodes[blk.side * 2]), a1 = vert(c, Lc.rc.side_nodes[blk.side * 2 + 1]);
      const uint64_t key = pair_key(a0, a1);
      int e;
      double t0, t1;
      auto tpar = [&](int e_, int vx) {  // parameter of a vertex on edge e_
        const Edge& E = edges[e_];
        return vx == E.P ? 0.0 : (vx == E.Q ? 1.0 : 0.5);
      };
      if (edge_of.count(key)) {
        e = edge_of.at(key);
      } else {
        e = parent_of.at(key);
      }
      t0 = tpar(e, a0);
      t1 = tpar(e, a1);
      const int deg = blk.kind == DPG_TRACE ? Lc.trace_deg : Lc.flux_deg;
      const std::vector<double> xl = unit_nodes(deg);
      for (int j = 0; j < blk.card; ++j) {
        const double t = t0 + (t1 - t0) * xl[j];
        std::vector<Entry> ent;
        if (blk.kind == DPG_TRACE) trace_at(blk.var, e, t, ent);
        else flux_at(blk.var, e, t, ent);
        // merge repeated dofs
        std::vector<Entry> merged;
        for (const Entry& x : ent) {
          bool found = false;
          for (Entry& y : merged)
            if (y.g == x.g) {
              y.w += x.w;
              found = true;
            }
          if (!found) merged.push_back(x);
        }
        const int i = blk.offset - Lc.F + j;
        rows[c][i] = merged;
        if (merged.size() == 1 && merged[0].w == 1.0 && hs.rep_cell[merged[0].g] < 0) {
          hs.rep_cell[merged[0].g] = c;
          hs.rep_local[merged[0].g] = i;
        }
        if (blk.kind == DPG_TRACE && edges[e].boundary)
          for (const Entry& x : merged) hs.fixed[x.g] = 1;
        hs.R[b] = std::max<int>(hs.R[b], (int)merged.size());
      }
    }
    for (const auto& row : rows[c])
      for (const Entry& x : row)
        if (std::find(cdofs[c].begin(), cdofs[c].end(), x.g) == cdofs[c].end()) cdofs[c].push_back(x.g);
    hs.G[b] = std::max<int>(hs.G[b], (int)cdofs[c].size());
  }
  // entities for block Jacobi: regular vertices, then edges (interior traces and fluxes)
  hs.entity.assign((size_t)next, -1);
  for (int v : trace_vars) {
    for (int x = 0; x < nvtx; ++x)
      if (vdof[x] >= 0) hs.entity[var_base[v] + vdof[x]] = vdof[x];
    for (int e = 0; e < ne; ++e)
      for (int m = 0; m < edges[e].order - 1; ++m) hs.entity[var_base[v] + nreg + ebase[e] + m] = nreg + e;
  }
  for (int v : flux_vars)
    for (int e = 0; e < ne; ++e)
      for (int m = 0; m < edges[e].order; ++m) hs.entity[var_base[v] + fbase[e] + m] = nreg + e;
  hs.nentities = nreg + ne;
  for (int32_t x : hs.entity)
    if (x < 0) throw std::logic_error("hp numbering left a dof without entity");
  hs.off.assign((size_t)C + 1, 0);
  for (int c = 0; c < C; ++c) {
    hs.off[(size_t)c + 1] = hs.off[c] + (int32_t)cdofs[c].size();
    hs.dofs.insert(hs.dofs.end(), cdofs[c].begin(), cdofs[c].end());
  }
  // positions of the row entries within the cell's dof list (replace g by it)
  for (int c = 0; c < C; ++c)
    for (auto& row : rows[c])
      for (Entry& x : row) x.g = (int32_t)(std::find(cdofs[c].begin(), cdofs[c].end(), x.g) - cdofs[c].begin());
  return hs;
}

// the device side of an hp system from its numbering
void build_hp(dpg_hp_system& hs, const dpg_form_desc& f, const dpg_bucket_desc* buckets, int nb, int C,
              const int32_t* cell_bucket, const int32_t* cv, int nvertices, const double* X, hipStream_t st) {
  const HpNumbering h = number_hp(f, buckets, nb, C, cell_bucket, cv, nvertices, X);
  hs.ncells = C;
  hs.nbuckets = nb;
  hs.nglobal = h.nglobal;
  hs.nentities = h.nentities;
  hs.T = h.T;
  hs.M = h.M;
  hs.F = h.F;
  hs.R = h.R;
  hs.G = h.G;
  hs.bucket_cells = h.bucket_cells;
  hs.off_h = h.off;
  hs.fixed_h = h.fixed;
  hs.entity_h = h.entity;
  hs.rep_cell_h = h.rep_cell;
  hs.rep_local_h = h.rep_local;
  const std::vector<std::vector<std::vector<Entry>>>& rows = h.rows;
  // ragged offsets, dofs and per-bucket padded M
  const std::vector<int32_t>& dofs = h.dofs;
  std::vector<int64_t> off2((size_t)C + 1, 0);
  for (int c = 0; c < C; ++c) {
    const int64_t g = hs.off_h[(size_t)c + 1] - hs.off_h[c];
    off2[(size_t)c + 1] = off2[c] + g * g;
  }
  std::vector<std::vector<int32_t>> cell_of(nb), mpos(nb);
  std::vector<std::vector<double>> mw(nb);
  for (int c = 0; c < C; ++c) {
    const int b = cell_bucket[c], T = hs.T[b], R = hs.R[b];
    cell_of[b].push_back(c);
    for (int i = 0; i < T; ++i)
      for (int r = 0; r < R; ++r) {
        if (r < (int)rows[c][i].size()) {
          const Entry& x = rows[c][i][r];
          mpos[b].push_back(x.g);
          mw[b].push_back(x.w);
        } else {
          mpos[b].push_back(-1);
          mw[b].push_back(0.0);
        }
      }
  }
  for (int b = 0; b < nb; ++b) {
    hs.cell_of.emplace_back(new HpBuf<int32_t>());
    hs.mpos.emplace_back(new HpBuf<int32_t>());
    hs.mw.emplace_back(new HpBuf<double>());
    hs.cell_of[b]->upload(cell_of[b], st);
    hs.mpos[b]->upload(mpos[b], st);
    hs.mw[b]->upload(mw[b], st);
    const size_t lds = sizeof(double) * (size_t)hs.T[b] * hs.G[b];
    if (lds > 64 * 1024) throw std::invalid_argument("an element type of the hp mesh is too large for the constraint kernel");
  }
  hs.off.upload(hs.off_h, st);
  hs.dofs.upload(dofs, st);
  hs.off2.upload(off2, st);
  hs.Shat.upload(std::vector<double>((size_t)off2[C], 0.0), st);
  hs.ghat.upload(std::vector<double>((size_t)hs.off_h[C], 0.0), st);
  if (dpg_skeleton_create_ragged(C, hs.off_h.data(), dofs.data(), (int)h.nglobal, st, &hs.sk) != 0)
    throw std::runtime_error(std::string("hp skeleton: ") + dpg_last_error());
  HP_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
}

}  // namespace
}  // namespace dpg

extern "C" {

int dpg_hp_system_create(const dpg_form_desc* form, const dpg_bucket_desc* buckets, int nbuckets, int ncells,
                         const int32_t* cell_bucket, const int32_t* cell_vertices, int nvertices,
                         const double* vertex_coords, void* stream, dpg_hp_system** out) {
  return dpg::guarded([&] {
    if (!form || !buckets || nbuckets <= 0 || ncells <= 0 || !cell_bucket || !cell_vertices || !out ||
        nvertices <= 0 || !vertex_coords)
      throw std::invalid_argument("bad arguments");
    *out = nullptr;
    std::unique_ptr<dpg_hp_system> hs(new dpg_hp_system());
    dpg::build_hp(*hs, *form, buckets, nbuckets, ncells, cell_bucket, cell_vertices, nvertices, vertex_coords,
                  (hipStream_t)stream);
    *out = hs.release();
  });
}

int dpg_hp_numbering(const dpg_form_desc* form, const dpg_bucket_desc* buckets, int nbuckets, int ncells,
                     const int32_t* cell_bucket, const int32_t* cell_vertices, int nvertices,
                     const double* vertex_coords, int64_t* sizes, int32_t* offsets, int32_t* dofs, int32_t* row_ptr,
                     int32_t* entry_dof, double* entry_weight) {
  return dpg::guarded([&] {
    if (!form || !buckets || nbuckets <= 0 || ncells <= 0 || !cell_bucket || !cell_vertices || !sizes ||
        nvertices <= 0 || !vertex_coords)
      throw std::invalid_argument("bad arguments");
    const dpg::HpNumbering h =
        dpg::number_hp(*form, buckets, nbuckets, ncells, cell_bucket, cell_vertices, nvertices, vertex_coords);
    int64_t nrows = 0, nent = 0;
    for (const auto& cell : h.rows) {
      nrows += (int64_t)cell.size();
      for (const auto& row : cell) nent += (int64_t)row.size();
    }
    sizes[0] = h.nglobal;
    sizes[1] = (int64_t)h.dofs.size();
    sizes[2] = nrows;
    sizes[3] = nent;
    if (offsets) std::memcpy(offsets, h.off.data(), h.off.size() * sizeof(int32_t));
    if (dofs) std::memcpy(dofs, h.dofs.data(), h.dofs.size() * sizeof(int32_t));
    if (row_ptr || entry_dof || entry_weight) {
      int64_t r = 0, q = 0;
      if (row_ptr) row_ptr[0] = 0;
      for (const auto& cell : h.rows)
        for (const auto& row : cell) {
          for (const dpg::Entry& x : row) {
            if (entry_dof) entry_dof[q] = h.dofs[(size_t)h.off[&cell - &h.rows[0]] + x.g];
            if (entry_weight) entry_weight[q] = x.w;
            ++q;
          }
          ++r;
          if (row_ptr) row_ptr[r] = (int32_t)q;
        }
    }
  });
}

int dpg_hp_system_info(const dpg_hp_system* hs, int64_t* nglobal, int64_t* nentities, dpg_skeleton** sk) {
  return dpg::guarded([&] {
    if (!hs) throw std::invalid_argument("NULL hp system");
    if (nglobal) *nglobal = hs->nglobal;
    if (nentities) *nentities = hs->nentities;
    if (sk) *sk = hs->sk;
  });
}

int dpg_hp_system_dofs(const dpg_hp_system* hs, int32_t* fixed, int32_t* entity, int32_t* rep_cell,
                       int32_t* rep_local) {
  return dpg::guarded([&] {
    if (!hs) throw std::invalid_argument("NULL hp system");
    const size_t n = (size_t)hs->nglobal;
    if (fixed) std::memcpy(fixed, hs->fixed_h.data(), n * sizeof(int32_t));
    if (entity) std::memcpy(entity, hs->entity_h.data(), n * sizeof(int32_t));
    if (rep_cell) std::memcpy(rep_cell, hs->rep_cell_h.data(), n * sizeof(int32_t));
    if (rep_local) std::memcpy(rep_local, hs->rep_local_h.data(), n * sizeof(int32_t));
  });
}

int dpg_hp_system_assemble(dpg_hp_system* hs, const double* const* schur, const double* const* schur_rhs,
                           double* values, double* rhs, void* stream) {
  return dpg::guarded([&] {
    if (!hs || !schur) throw std::invalid_argument("NULL argument");
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < hs->nbuckets; ++b) {
      if (hs->bucket_cells[b] == 0) continue;
      if (!schur[b]) throw std::invalid_argument("a bucket's Schur blocks are missing");
      const double* g = schur_rhs ? schur_rhs[b] : nullptr;
      if (rhs && !g) throw std::invalid_argument("rhs needs schur_rhs of every bucket");
      const size_t lds = sizeof(double) * (size_t)hs->T[b] * hs->G[b];
      dpg::constrain_blocks<<<hs->bucket_cells[b], 256, lds, st>>>(
          schur[b], g, hs->T[b], hs->R[b], hs->G[b], hs->cell_of[b]->p, hs->mpos[b]->p, hs->mw[b]->p, hs->off.p,
          hs->off2.p, hs->Shat.p, hs->ghat.p);
      HP_CHECK(hipGetLastError());
    }
    if (dpg_skeleton_assemble(hs->sk, hs->Shat.p, rhs ? hs->ghat.p : nullptr, values, rhs, stream) != 0)
      throw std::runtime_error(dpg_last_error());
  });
}

int dpg_hp_system_local(const dpg_hp_system* hs, const double* x, double* const* u, void* stream) {
  return dpg::guarded([&] {
    if (!hs || !x || !u) throw std::invalid_argument("NULL argument");
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < hs->nbuckets; ++b) {
      const int n = hs->bucket_cells[b];
      if (n == 0) continue;
      if (!u[b]) throw std::invalid_argument("a bucket's output is missing");
      const int64_t q = (int64_t)n * hs->T[b];
      dpg::local_from_global<<<(unsigned)((q + 255) / 256), 256, 0, st>>>(
          x, n, hs->T[b], hs->R[b], hs->M[b], hs->F[b], hs->cell_of[b]->p, hs->mpos[b]->p, hs->mw[b]->p, hs->off.p,
          hs->dofs.p, u[b]);
      HP_CHECK(hipGetLastError());
    }
  });
}

int dpg_hp_system_destroy(dpg_hp_system* hs) {
  return dpg::guarded([&] { delete hs; });
}

}  // extern "C"
